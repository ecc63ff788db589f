"""Continuous-batching DEM server over the segmented batch engine.

Requests (a scene and a number of steps) arrive over time.  At every scheduling epoch the server
admits the requests that have arrived -- first come, first served, as long as the estimated HBM
footprint of all live scenes stays within the budget -- and re-packs the live set into one batch
(dem_batch_repack: the scenes already running move over device to device, only the arrivals are
uploaded; dem_initialize for the newly admitted bonded ones).  It then advances every live scene
together, one segmented launch per stage and one graph replay per two steps, by the smaller of the
epoch length and the fewest steps any live scene still needs, so no scene runs past its request.
With `poll_steps` below the epoch length the epoch is stepped in slices of that many steps, and a
request that arrives during an epoch ends it after the current slice once the epoch has run
`min_epoch_s` (adaptive epochs): an arrival then waits at most that long plus one slice to be
admitted, and a burst of arrivals shares one re-pack.  On the benchmark stream, whose latencies
are dominated by the scenes' own step counts, slices of 100 steps admitted arrivals sooner but
cost more in host round trips and re-packs than they saved (makespan 10.9 s against 10.6 s, p99
10.2 s against 9.8 s), so by default an epoch runs whole.
Finished scenes are read back and retired at the next epoch.  State stays on the device between
epochs.

``serve_one_at_a_time`` is the plain path beside it: each request alone, in arrival order, with
its own FP64 scene.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import scenes

# HBM bytes per grain, per contact slot, of one scene (device arrays of abi.hip allocate_scene):
# double-buffered pos/vel/omg/quat, a/dw/F/T outputs, r0, wall flags; sort arrays (bucket, rank,
# FP32 position, two orders, sorted bucket, FP64 and FP32 sorted positions, about two bucket
# counters); file index; touch count, row selector, row counts; scratch; scene id
BYTES_PER_GRAIN = (2 * 4 * 32 + 4 * 32 + 8 + 4 + (4 + 4 + 16 + 2 * 4 + 4 + 32 + 16 + 8) + 4
                   + (4 + 1 + 2 * 4) + 2 * 4 + 2)
BYTES_PER_SLOT = 2 * (4 + 15 * 8) + 4


def footprint(n: int, slots: int, nwall: int = 5) -> int:
    """Estimated device bytes of one scene with n grains and `slots` contact slots per grain."""
    return n * (BYTES_PER_GRAIN + slots * BYTES_PER_SLOT + 32 * max(nwall, 1))


@dataclass
class Request:
    rid: int
    arrival: float
    kind: str          # "cube" (bonded, initialized) or "pour" (no bonds)
    n: int
    steps: int
    seed: int
    spec: tuple = None  # (cfg, grains, walls), built on demand

    @property
    def bonded(self) -> bool:
        return self.kind == "cube"

    def build(self):
        if self.spec is None:
            if self.kind == "cube":
                nx = max(2, int(round(self.n ** (1 / 3))))
                nz = max(2, int(round(self.n / (nx * nx))))
                rng = np.random.default_rng(self.seed)
                sb = math.exp(rng.uniform(math.log(1e5), math.log(1e7)))
                self.spec = scenes.bonded_cube(nx, nx, nz, speed=rng.uniform(2, 30), dt=5e-7,
                                               seed=self.seed, bond_tensile_strength=sb,
                                               bond_shear_strength=sb)
            else:
                # the pour density: box sized for the grain count at solid fraction ~0.23
                side = (self.n * 0.694e-9 / 0.23 / 3.0) ** (1 / 3)
                box = (side, side, 3 * side)
                self.spec = scenes.pour(self.n, seed=self.seed, box=box)
        return self.spec


def make_stream(nreq: int = 40, rate: float = 20.0, seed: int = 5, nmin: int = 1000,
                nmax: int = 200_000, smin: int = 2000, smax: int = 50_000) -> List[Request]:
    """Poisson arrivals at `rate` per second; grains log-uniform in [nmin, nmax]; steps uniform
    in [smin, smax]; alternately bonded cubes and loose pours."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(nreq):
        t += rng.exponential(1.0 / rate)
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        steps = int(rng.integers(smin, smax + 1))
        out.append(Request(k, t, "cube" if k % 2 == 0 else "pour", n, steps, 10_000 + k))
    return out


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    finished: float
    n: int
    steps: int
    grains: Optional[np.ndarray] = None
    contacts: Optional[tuple] = None
    bonds_broken: int = 0

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Live:
    req: Request
    remaining: int
    admitted: float
    slot: int = -1


def _percentile(x, q):
    return float(np.percentile(np.asarray(x), q)) if len(x) else float("nan")


def summarize(results: List[Result], makespan: float) -> dict:
    lat = [r.latency for r in results]
    work = sum(r.n * r.steps for r in results)
    return {"requests": len(results), "makespan_s": round(makespan, 3),
            "scenes_per_s": round(len(results) / makespan, 3),
            "particle_steps_per_s": round(work / makespan, 1),
            "latency_p50_s": round(_percentile(lat, 50), 3),
            "latency_p99_s": round(_percentile(lat, 99), 3),
            "latency_max_s": round(max(lat), 3) if lat else float("nan")}


class Server:
    def __init__(self, precision: str = "mixed", hbm_budget_gb: float = 200.0,
                 epoch_steps: int = 1000, device: int = 0, keep_results: bool = False,
                 memory_cache_gb: float = 64.0, poll_steps: Optional[int] = None,
                 min_epoch_s: float = 0.1, near_lists: bool = True):
        self.precision = precision
        self.near_lists = near_lists
        self.poll = max(1, int(poll_steps)) if poll_steps else max(1, int(epoch_steps))
        self.min_epoch = float(min_epoch_s)
        self.budget = hbm_budget_gb * 1e9
        # released batches' device memory kept for the next re-pack while the server runs
        self.cache = int(memory_cache_gb * 1e9)
        self.epoch = epoch_steps
        self.device = device
        self.keep = keep_results
        self.repacks = 0
        self.repack_s = 0.0  # wall time spent re-packing (batch creation, state transfer, bond pass)
        self.step_s = 0.0    # wall time spent stepping
        # one record per epoch: start time, live scenes and grains, steps, and the seconds spent
        # packing the batch, initialising bonds, releasing the previous batch and stepping
        self.epochs: List[dict] = []

    def _slots(self, live: List[_Live]) -> int:
        return max(min(128, max(32, 2 * l.req.build()[0].max_coordinate_number)) for l in live)

    def run(self, requests: List[Request]) -> tuple:
        """Serve the requests in real time (arrival times are seconds after the call)."""
        from .executor import set_memory_cache
        set_memory_cache(self.cache)
        try:
            return self._run(requests)
        finally:
            set_memory_cache(0)

    def _run(self, requests: List[Request]) -> tuple:
        from .executor import Batch
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            r.build()
        live: List[_Live] = []
        done: List[Result] = []
        batch = None
        dirty = False  # the live set changed since the batch was packed
        t0 = time.perf_counter()

        def clock():
            return time.perf_counter() - t0

        while pending or live:
            now = clock()
            # admission: arrived requests, FIFO, within the HBM budget
            new = []
            used = sum(footprint(l.req.n, 64, len(l.req.build()[2])) for l in live)
            while pending and pending[0].arrival <= now:
                r = pending[0]
                need = footprint(r.n, 64, len(r.build()[2]))
                if (live or new) and used + need > self.budget:
                    break
                pending.pop(0)
                used += need
                new.append(_Live(r, r.steps, now))
            if not live and not new:
                if batch is not None:
                    batch.close()
                    batch = None
                time.sleep(max(pending[0].arrival - now, 0.0))
                continue
            rec = {"t": round(now, 4), "new": len(new)}
            if new or dirty or batch is None:
                tr = time.perf_counter()
                members = live + new
                fresh = [m.req.build() for m in new]
                if batch is None or not live:
                    nb = Batch(fresh, precision=self.precision, device=self.device,
                               near_lists=self.near_lists)
                else:
                    # the live scenes move over on the device; only arrivals are uploaded
                    nb = Batch.repack(batch, [m.slot for m in live], fresh,
                                      precision=self.precision, device=self.device,
                                      near_lists=self.near_lists)
                nb.set_bond_init([0] * len(live) + [1 if m.req.bonded else 0 for m in new])
                for k, m in enumerate(members):
                    m.slot = k
                ti = time.perf_counter()
                if any(m.req.bonded for m in new):
                    nb.initialize()
                te = time.perf_counter()
                if batch is not None:
                    batch.close()
                batch = nb
                live = members
                dirty = False
                self.repacks += 1
                self.repack_s += time.perf_counter() - tr
                rec.update(pack_s=round(ti - tr, 4), init_s=round(te - ti, 4),
                           close_s=round(time.perf_counter() - te, 4),
                           slots=int(batch.derived()["contact_slots"]))
            k = min(self.epoch, min(m.remaining for m in live))
            ts = time.perf_counter()
            # the epoch in slices; an arrival ends it early (admitted at the next epoch)
            done_k = 0
            while done_k < k:
                sl = min(self.poll, k - done_k)
                batch.step(sl)
                done_k += sl
                if pending and pending[0].arrival <= clock() and \
                        time.perf_counter() - ts >= self.min_epoch:
                    break
            k = done_k
            self.step_s += time.perf_counter() - ts
            rec.update(scenes=len(live), grains=int(batch.n), steps=int(k),
                       step_s=round(time.perf_counter() - ts, 4))
            self.epochs.append(rec)
            for m in live:
                m.remaining -= k
            finished = [m for m in live if m.remaining == 0]
            if finished:
                now = clock()
                grains = batch.read_grains() if self.keep else None
                contacts = batch.split_contacts(*batch.read_contacts()) if self.keep else None
                broken = batch.bonds_broken()
                for m in finished:
                    res = Result(m.req.rid, m.req.arrival, m.admitted, now, m.req.n, m.req.steps,
                                 bonds_broken=int(broken[m.slot]))
                    if self.keep:
                        res.grains = grains[batch.offsets[m.slot]:batch.offsets[m.slot + 1]].copy()
                        res.contacts = contacts[m.slot]
                    done.append(res)
                live = [m for m in live if m.remaining > 0]
                dirty = True  # survivors are re-packed with the next arrivals
        makespan = clock()
        if batch is not None:
            batch.close()
        return sorted(done, key=lambda r: r.rid), makespan


def serve_one_at_a_time(requests: List[Request], precision: str = "fp64", realtime: bool = True,
                        keep_results: bool = False, device: int = 0) -> tuple:
    """Each request alone, in arrival order, with its own scene (the plain drop-in path)."""
    from .executor import Scene
    order = sorted(requests, key=lambda r: r.arrival)
    for r in order:
        r.build()
    t0 = time.perf_counter()
    done = []
    for r in order:
        now = time.perf_counter() - t0
        if realtime and now < r.arrival:
            time.sleep(r.arrival - now)
        start = time.perf_counter() - t0
        cfg, g, w = r.build()
        with Scene(cfg, g, w, precision=precision, device=device) as s:
            if r.bonded:
                s.initialize()
            s.step(r.steps)
            res = Result(r.rid, r.arrival, start, time.perf_counter() - t0, r.n, r.steps,
                         bonds_broken=s.counters()["bonds_broken_total"])
            if keep_results:
                res.grains = s.read_grains()
                res.contacts = s.read_contacts()
        done.append(res)
    return sorted(done, key=lambda r: r.rid), time.perf_counter() - t0
