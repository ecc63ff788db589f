"""Benchmark: particle-steps/s of the DEM step on one MI355X.

Workload (default): the 1M-grain polydisperse pour -- log-normal radii (median 0.5 mm, sigma 0.25,
clipped 0.25-1.0 mm), density 2500 kg/m^3, random non-overlapping seeding in a 0.1 x 0.1 x 0.3 m
box with a floor and four side walls, Hertz-Mindlin only (E = 1e7 Pa, mu = 0.4), dt = 2e-7 s,
synthetic input generated in-process (seed 3).  A step is one simulate! over all grains.

Prints one JSON line.  ``value`` is the whole-job throughput (particle-steps/s) of the timed K
steps (graph replays), state resident in HBM before the timed region starts.  ``roofline``
describes the dominant kernel (largest share of the step time), its average duration measured
with HIP events on the scene's stream over K identical eagerly-launched steps just before.
``cpu_baseline`` times the FP64 CPU restatement (the oracle) of the same step on a bounded sample
on this host, serial and as its OpenMP build on every thread OpenMP is given; the faster is the
baseline and both are reported.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def _algorithmic_bytes(kernel: str, c: dict, nwall: int, precision: str, free: bool = False,
                       near: bool = True) -> float:
    """Compulsory HBM bytes of one launch of ``kernel`` in this engine's design (each array the
    kernel must touch, read or written once):
      hash            previous order 4 B + position 32 B read; bucket, rank 8 B + FP32 position
                      16 B written per grain
      scan            cell counts read + cleared 8 B, starts written 4 B per cell
      scatter         bucket, rank, previous order, FP32 position 28 B + bucket start 4 B read;
                      order, sorted bucket, sorted FP32 position 24 B written per grain
      broad_phase     sorted FP32 position, bucket and order 24 B per grain, bucket starts 4 B per
                      cell; per touching pair both FP64 positions 64 B, the partner's index 4 B and
                      the two row appends (4 B counter + 4 B entry each)
      merge           touch count 4 B + row count and selector 5 B per grain
      force_integrate grain state read 128 B + written 128 B, row counts 5 B, wall flags 4 B, with
                      near lists the FP32 rebuild position 16 B (read for the displacement check,
                      written on a rebuilding step);
                      per Hertz contact 2 rows x (partner 4 + spring 24 read + 24 written),
                      per bond 2 rows x (partner 4 + F,ti,tj 72 read + 72 written), per wall
                      contact 32 B (contact points and local forces are written only on a call's
                      last step, outside the timed replays); in the contact-free mode of a single
                      scene with no row entry (free) the row counts are not read"""
    n = c["n"]
    cells = c.get("grid_cells") or c["hash_table_size"]
    Ch, Cb, W = c["hertz"], c["bonded"], c["wall_contacts"]
    touching = Ch + Cb
    if kernel == "broad_phase":
        return n * 24.0 + 4.0 * cells + touching * (64 + 4 + 16)
    if kernel == "force_integrate":
        return (n * (128 + 128 + (0 if free else 5) + 4 + (16 if near else 0)) + 2 * Ch * (4 + 48)
                + 2 * Cb * (4 + 144) + 32.0 * W)
    if kernel == "hash":
        return n * 60.0
    if kernel == "scatter":
        return n * 56.0
    if kernel == "scan":
        return 12.0 * cells
    if kernel == "merge":
        return n * 9.0
    return 0.0


def _reference_equivalent_bytes(c: dict) -> float:
    """Bytes the reference design moves for its broad phase: 20 B per cp_list entry (write 8 +
    read 8 + one 4-B BitTable probe), the per-candidate-pair figure of the measurement plan."""
    return 20.0 * c["candidate_pairs"]


PROFILE_TAG = {"pour": "pour1m", "pour_packed": "pour1m_packed", "batch256": "batch256",
               "cube18k": "cube18k", "cube900": "cube900"}


def measured_traffic(workload: str, precision: str, n: int):
    """HBM bytes per launch of each kernel from the committed rocprofv3 summary of this workload
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the same counters' reading for a copy kernel
    of known size): the line ``traffic_bytes_per_launch = {...}`` of
    profiles/<workload>_<precision>_hbm_traffic.md.  None when no summary covers the run."""
    if workload not in PROFILE_TAG or (workload.startswith("pour") and n != 1_000_000):
        return None, None
    path = os.path.join(ROOT, "profiles", f"{PROFILE_TAG[workload]}_{precision}_hbm_traffic.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch = "):
            return json.loads(line.split("=", 1)[1]), os.path.relpath(path, ROOT)
    return None, None


def build_workload(name: str, n: int):
    from taichi_dem_julia_port_amd import scenes
    if name == "pour":
        cfg, g, w = scenes.pour(n)
        desc = (f"{n}-grain polydisperse pour (log-normal r, median 0.5 mm, sigma 0.25, clipped "
                "0.25-1.0 mm; 0.1x0.1x0.3 m box, floor + 4 side walls; E=1e7 Pa, mu=0.4, "
                "dt=2e-7 s; no bonds)")
        return cfg, g, w, False, desc
    if name == "pour_packed":
        cfg, g, w = scenes.pour_packed(n)
        desc = (f"{n}-grain pour, contact-heavy state: the pour's seeded grains with radii "
                "inflated 1.25x after placement and velocities uniform in +-0.05 m/s (Hertz-Mindlin "
                "only, E=1e7 Pa, mu=0.4, dt=2e-7 s)")
        return cfg, g, w, False, desc
    if name == "cube18k":
        cfg, g, w = scenes.cube_18k()
        return cfg, g, w, True, "18252-grain bonded cube impact (26x26x27, 20 m/s, dt=5e-7 s)"
    if name == "cube900":
        cfg, g, w = scenes.cube_small()
        return cfg, g, w, True, "900-grain bonded cube impact (10x10x9, 10 m/s, dt=1e-6 s)"
    if name == "batch256":
        members = scenes.batched_cubes(256, seed=4)
        return members, None, None, True, (
            "256 bonded-cube scenes advanced together (grains log-uniform 500-20000, E 1e7-1e9 Pa, "
            "mu 0.2-0.8, bond strength 1e5-1e7 Pa, impact 2-30 m/s, dt=5e-7 s), one segmented "
            "launch per stage")
    raise SystemExit(f"unknown workload {name}")


def progress(msg):
    """A progress line on stderr (stdout carries the one JSON result line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_engine(cfg, g, w, bonded, precision, steps, warmup, graph=True, near_lists=True):
    from taichi_dem_julia_port_amd.executor import Batch, Scene
    progress(f"engine run: {precision}, {steps} steps after {warmup} warm-up"
             f"{'' if near_lists else ', sort and broad phase on every step'}")
    s = (Batch(cfg, precision=precision, graph=graph, near_lists=near_lists) if g is None
         else Scene(cfg, g, w, precision=precision, graph=graph, near_lists=near_lists))
    if bonded:
        s.initialize()
    s.step(warmup)
    # capture the step graphs of both buffer parities (a call of 19 steps replays one
    # sixteen-step and one two-step graph and flips the parity; one of 3 steps replays a
    # two-step graph), so that no capture falls into a timed call
    for n in (19, 19, 3, 3):
        s.step(n)
    # per-kernel durations: K steps with a HIP event pair around every launch on the scene's
    # stream (events keep these steps out of the captured graph)
    s.set_timing(True)
    t0 = time.perf_counter()
    s.step(steps)
    t1 = time.perf_counter()
    times = s.kernel_times()
    s.set_timing(False)
    rebuilds0 = s.counters()["rebuilds"]
    # timed region: the next K steps on the production path (graph replays), state in HBM;
    # dem_step returns after a stream synchronisation
    t2 = time.perf_counter()
    s.step(steps)
    t3 = time.perf_counter()
    rebuilds = s.counters()["rebuilds"] - rebuilds0
    # one more step, untimed, with the diagnostic passes: the reference's cp_list length and the
    # candidates the broad phase examined
    s.set_diagnostics(True)
    s.step(1)
    c = s.counters()
    d = s.derived()
    s.close()
    return {"elapsed": t3 - t2, "elapsed_events": t1 - t0, "kernels": times, "counters": c,
            "derived": d, "rebuilds": rebuilds}


def precision_error(cfg, g, w, bonded, steps=1000):
    """Mixed precision against FP64 after `steps` steps from the same initial state (untimed):
    the largest position difference relative to the largest |k| and to the largest radius.
    Beside it, the same figures for FP64 against FP64 started from x coordinates moved by one ulp,
    by 1e-9 and by 1e-7 relative (about one FP32 rounding): the divergence the dynamics alone
    produce from a small difference (a contact that forms or breaks one step apart changes a
    grain's path by far more than the difference that decided it)."""
    import numpy as np
    from taichi_dem_julia_port_amd.executor import Batch, Scene

    def bump(gr, rel):
        gr = gr.copy()
        x = gr["k"][:, 0]
        gr["k"][:, 0] = np.nextafter(x, np.inf) if rel == 0 else x * (1.0 + rel)
        return gr

    def run(prec, perturb=None):
        if g is None:
            members = [(c, gg if perturb is None else bump(gg, perturb), ww) for (c, gg, ww) in cfg]
            s = Batch(members, precision=prec)
        else:
            s = Scene(cfg, g if perturb is None else bump(g, perturb), w, precision=prec)
        if bonded:
            s.initialize()
        s.step(steps)
        out = s.read_grains()
        s.close()
        return out

    progress(f"position error after {steps} steps: mixed and perturbed FP64 runs")
    a, b = run("fp64"), run("mixed")
    c, e, f = run("fp64", 0.0), run("fp64", 1e-9), run("fp64", 1e-7)
    scale_k = max(float(np.abs(a["k"]).max()), 1e-300)
    scale_r = max(float(a["r"].max()), 1e-300)

    def err(x, y):
        per = np.abs(x["k"] - y["k"]).max(axis=1) if len(x) else np.zeros(1)
        d = float(per.max())
        return {"max_abs_m": d, "relative_to_max_position": d / scale_k,
                "relative_to_max_radius": d / scale_r,
                "p99_abs_m": float(np.percentile(per, 99)),
                "grains_above_1e-6_of_max_position": int((per > 1e-6 * scale_k).sum())}
    return {"steps": steps, "mixed_vs_fp64": err(a, b), "fp64_one_ulp_perturbation": err(a, c),
            "fp64_1e-9_relative_perturbation": err(a, e),
            "fp64_1e-7_relative_perturbation": err(a, f)}


def one_at_a_time(members, precision, steps, sample=16, seed=0):
    """Plain path beside a batch: the same scenes stepped one scene at a time (a seeded sample),
    scaled to the whole set by particle count."""
    import numpy as np
    from taichi_dem_julia_port_amd.executor import Scene
    rng = np.random.default_rng(seed)
    idx = rng.choice(len(members), size=min(sample, len(members)), replace=False)
    tot_n, tot_t = 0, 0.0
    for k in idx:
        cfg, g, w = members[k]
        with Scene(cfg, g, w, precision=precision) as s:
            s.initialize()
            s.step(5)
            t0 = time.perf_counter()
            s.step(steps)
            tot_t += time.perf_counter() - t0
        tot_n += len(g)
    return tot_n * steps / tot_t, [int(k) for k in idx]


def cpu_baseline(cfg, g, w, bonded, steps, parallel=False):
    """The oracle (FP64 C restatement of the reference step) on a bounded sample: the serial
    build, or with ``parallel`` its OpenMP build on all threads OpenMP is given."""
    from oracle.oracle import OracleScene
    if g is None:  # a batch: time the oracle on its scenes one after another (first 8)
        tot_n, tot_t = 0, 0.0
        for (c, gg, ww) in cfg[:8]:
            o = OracleScene(c, gg, ww, parallel=parallel)
            o.initialize()
            o.step(1)
            t0 = time.perf_counter()
            o.step(steps)
            tot_t += time.perf_counter() - t0
            tot_n += len(gg)
            o.close()
        return tot_n * steps / tot_t, tot_t
    o = OracleScene(cfg, g, w, parallel=parallel)
    if bonded:
        o.initialize()
    o.step(1)  # warm-up: allocations and first contact pass
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    return len(g) * steps / dt, dt


def run_serving(args):
    """Serving stream: Poisson arrivals, continuous batching vs one scene at a time."""
    from taichi_dem_julia_port_amd import serving
    reqs = serving.make_stream(args.requests, rate=args.rate, seed=5)
    work = sum(r.n * r.steps for r in reqs)
    srv = serving.Server(precision=args.precision, epoch_steps=args.epoch,
                         poll_steps=args.poll or None)
    done, makespan = srv.run(reqs)
    out_b = serving.summarize(done, makespan)
    base, makespan0 = serving.serve_one_at_a_time(serving.make_stream(args.requests, rate=args.rate, seed=5),
                                                   precision="fp64")
    out_0 = serving.summarize(base, makespan0)
    out = {
        "metric": "served scenes/s (continuous batching), p50/p99 completion latency",
        "value": out_b["scenes_per_s"], "unit": "scenes/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": (f"{args.requests} requests, Poisson arrivals at {args.rate}/s, grains "
                                f"log-uniform 1k-200k, half bonded cubes / half loose pours, steps "
                                f"uniform 2k-50k, epochs of up to {args.epoch} steps"
                                + (f", ended early by an arrival after a slice of {args.poll} "
                                   f"steps" if args.poll else "")),
                   "precision": args.precision, "particle_steps_requested": work},
        "batched": out_b, "repacks": srv.repacks,
        "repack_s": round(srv.repack_s, 3), "stepping_s": round(srv.step_s, 3),
        "one_at_a_time_fp64": out_0,
        "speedup_scenes_per_s": round(out_b["scenes_per_s"] / out_0["scenes_per_s"], 2),
    }
    print(json.dumps(out), flush=True)


def summarize(r, workload, precision, nwall, n, steps, near=True):
    """Throughput, per-kernel durations and the dominant kernel's roofline of one engine run."""
    kt = r["kernels"]
    step_ms = {k: (ms / max(l, 1)) for k, (ms, l) in kt.items() if l}
    total_kernel_ms = sum(ms for ms, _ in kt.values())
    dom = max(((k, ms) for k, (ms, l) in kt.items() if l), key=lambda kv: kv[1])[0]
    c = r["counters"]
    # a single scene whose state has no contact runs the contact-free force kernel
    free = workload != "batch256" and c["contacts"] == 0
    alg = _algorithmic_bytes(dom, c, nwall, precision, free, near)
    achieved = alg / (step_ms[dom] * 1e-3) / 1e9 if step_ms[dom] > 0 else 0.0
    traffic, tsrc = measured_traffic(workload, precision, n)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic.get(dom) if traffic else None, "algorithmic_bytes": alg,
                "avg_launch_ms": round(step_ms[dom], 5)}
    if traffic and traffic.get(dom):
        roofline["traffic_source"] = tsrc
        roofline["traffic_GB_s"] = round(traffic[dom] / (step_ms[dom] * 1e-3) / 1e9, 1)
    if dom == "broad_phase":
        # candidates examined are re-read from L2/MALL, not HBM: the kernel's practical limiter
        roofline["candidates_examined"] = c["examined"]
        roofline["cache_GB_s"] = round(c["examined"] * 16 / (step_ms[dom] * 1e-3) / 1e9, 1)
        roofline["reference_equivalent_GB_s"] = round(
            _reference_equivalent_bytes(c) / (step_ms[dom] * 1e-3) / 1e9, 1)
    per_kernel = {}
    for k, ms in step_ms.items():
        b = _algorithmic_bytes(k, c, nwall, precision, free, near)
        if b > 0 and ms > 0:
            per_kernel[k] = round(b / (ms * 1e-3) / 1e9, 1)
    return {"value": n * steps / r["elapsed"],
            "ms_per_step": round(1e3 * r["elapsed"] / steps, 5),
            "kernel_ms_per_launch": {k: round(v, 5) for k, v in step_ms.items()},
            "kernel_share": {k: round(ms / total_kernel_ms, 4) for k, (ms, l) in kt.items() if l},
            "ms_per_step_with_kernel_events": round(1e3 * r["elapsed_events"] / steps, 5),
            "roofline": roofline, "kernel_GB_s": per_kernel,
            "rebuilds_in_timed_steps": r["rebuilds"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="pour",
                    choices=["pour", "pour_packed", "cube18k", "cube900", "batch256", "serve"])
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--epoch", type=int, default=1000)
    ap.add_argument("--poll", type=int, default=0,
                    help="serving: steps per slice of an epoch, after which an arrival ends the "
                         "epoch (default: the whole epoch)")
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--precision", default=None, choices=["fp64", "mixed"],
                    help="default: mixed for the serving stream, fp64 otherwise")
    ap.add_argument("--no-compare-precision", action="store_true",
                    help="skip timing the other precision and the 1000-step position error")
    ap.add_argument("--no-contact-heavy", action="store_true",
                    help="skip the pour's contact-heavy state (reported beside the pour)")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="steps of the serial oracle's CPU baseline (default: 2 for the pours, 20 "
                         "for the 18k cube, 2000 for the 900-grain cube, 5 per batch scene); the "
                         "OpenMP build runs 10, 6, 1000, all 10000 and 100")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step eagerly (no hipGraph replay)")
    ap.add_argument("--no-near-lists", action="store_true",
                    help="sort and run the broad phase on every step (no near lists)")
    ap.add_argument("--no-every-step", action="store_true",
                    help="skip timing the every-step sort and broad phase path beside the "
                         "near-list one")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="first run a plain copy kernel of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; scenes are not split across GPUs "
                                   "and multi-GPU runs are not supported (use --gpus 1)"}))
        sys.exit(2)

    if args.precision is None:
        args.precision = "mixed" if args.workload == "serve" else "fp64"
    if args.workload == "serve":
        run_serving(args)
        return
    copy_cal = None
    if args.copy_calibration_mb > 0:
        import ctypes
        from taichi_dem_julia_port_amd import _native as N
        ms = ctypes.c_double(0.0)
        nbytes = args.copy_calibration_mb << 20
        rc = N.load_library().dem_debug_copy(nbytes, 3, ctypes.byref(ms))
        if rc == 0:
            copy_cal = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms.value,
                        "GB_s": round(2 * nbytes / (ms.value * 1e-3) / 1e9, 1)}
    t_build = time.perf_counter()
    cfg, g, w, bonded, desc = build_workload(args.workload, args.n)
    t_build = time.perf_counter() - t_build
    n = len(g) if g is not None else sum(len(m[1]) for m in cfg)
    if g is None:
        w = cfg[0][2]
    near = not args.no_near_lists
    r = run_engine(cfg, g, w, bonded, args.precision, args.steps, args.warmup,
                   graph=not args.no_graph, near_lists=near)
    m = summarize(r, args.workload, args.precision, len(w), n, args.steps, near)
    c = r["counters"]
    out = {
        "metric": "particle-steps/s (1M grains, FP32 vs FP64)",
        "value": round(m["value"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": desc, "n": n, "precision": args.precision,
                   "hash_table_size": r["derived"]["hash_table_size"],
                   "contact_slots": r["derived"]["contact_slots"]},
        "counters": {k: c[k] for k in ("candidate_pairs", "examined", "grid_cells", "contacts",
                                       "bonded", "hertz", "wall_contacts", "bonds_broken_total",
                                       "overflow")},
        "kernel_ms_per_launch": m["kernel_ms_per_launch"],
        "kernel_share": m["kernel_share"],
        "ms_per_step_with_kernel_events": m["ms_per_step_with_kernel_events"],
        "graph": not args.no_graph,
        "near_lists": near,
        "rebuilds_in_timed_steps": m["rebuilds_in_timed_steps"],
        "roofline": m["roofline"],
        "kernel_GB_s": m["kernel_GB_s"],
    }
    if near and not args.no_every_step:
        # the same run with the sort and the broad phase on every step: a step that rebuilds the
        # near lists costs about this much
        r0 = run_engine(cfg, g, w, bonded, args.precision, args.steps, args.warmup,
                        graph=not args.no_graph, near_lists=False)
        m0 = summarize(r0, args.workload, args.precision, len(w), n, args.steps, near=False)
        out["every_step_rebuild"] = {"value": round(m0["value"], 1), "ms_per_step": m0["ms_per_step"],
                                     "kernel_ms_per_launch": m0["kernel_ms_per_launch"],
                                     "near_list_speedup": round(m["value"] / m0["value"], 3)}
    if g is None:
        out["scenes"] = len(cfg)
        out["scene_steps_per_s"] = round(len(cfg) * args.steps / r["elapsed"], 1)
        v1, idx = one_at_a_time(cfg, args.precision, args.steps)
        out["one_at_a_time"] = {"value": round(v1, 1), "unit": "particle-steps/s",
                                "sample_scenes": idx, "speedup_batched": round(m["value"] / v1, 2)}
    if not args.no_compare_precision:
        other = "mixed" if args.precision == "fp64" else "fp64"
        r2 = run_engine(cfg, g, w, bonded, other, args.steps, args.warmup,
                        graph=not args.no_graph, near_lists=near)
        m2 = summarize(r2, args.workload, other, len(w), n, args.steps, near)
        out[f"value_{other}"] = round(m2["value"], 1)
        out[f"ms_per_step_{other}"] = m2["ms_per_step"]
        out[f"{other}_speedup_over_{args.precision}"] = round(m2["value"] / m["value"], 3)
        out[f"kernel_ms_per_launch_{other}"] = m2["kernel_ms_per_launch"]
        out[f"rebuilds_in_timed_steps_{other}"] = m2["rebuilds_in_timed_steps"]
        out["position_error_after_1000_steps"] = precision_error(cfg, g, w, bonded)
    if args.workload == "pour" and not args.no_contact_heavy:
        # the same grains in their contact-heavy state (the pour has no contacts for its first
        # thousands of steps): where the contact kernels are measured at this size
        progress("contact-heavy state of the pour")
        cfg_p, g_p, w_p, _, desc_p = build_workload("pour_packed", args.n)
        heavy = {"workload": desc_p}
        for prec in ("fp64", "mixed"):
            rp = run_engine(cfg_p, g_p, w_p, False, prec, args.steps, args.warmup,
                            graph=not args.no_graph, near_lists=near)
            mp = summarize(rp, "pour_packed", prec, len(w_p), len(g_p), args.steps, near)
            heavy[prec] = {"value": round(mp["value"], 1), "ms_per_step": mp["ms_per_step"],
                           "kernel_ms_per_launch": mp["kernel_ms_per_launch"],
                           "roofline": mp["roofline"],
                           "rebuilds_in_timed_steps": mp["rebuilds_in_timed_steps"]}
            heavy["counters"] = {k: rp["counters"][k] for k in ("contacts", "hertz",
                                                                "wall_contacts", "examined",
                                                                "candidate_pairs")}
        heavy["mixed_speedup_over_fp64"] = round(heavy["mixed"]["value"] / heavy["fp64"]["value"], 3)
        out["contact_heavy"] = heavy
    if not args.no_cpu_baseline:
        from oracle.oracle import num_threads
        if args.cpu_steps <= 0:
            args.cpu_steps = {"pour": 2, "pour_packed": 2, "cube18k": 20, "cube900": 2000,
                              "batch256": 5}[args.workload]
        omp_steps = {"pour": 10, "pour_packed": 6, "cube18k": 1000, "cube900": 10000,
                     "batch256": 100}[args.workload]

        def sample(steps, dt, how):
            return (f"same workload ({n} grains), {steps} steps after 1 warm-up step, {how}, "
                    f"{dt:.1f} s") if g is not None else (
                    f"first 8 scenes of the batch one after another, {steps} steps each after "
                    f"initialize + 1 warm-up step, {how}, {dt:.1f} s")
        progress("CPU baseline: serial and OpenMP oracle")
        v1, dt1 = cpu_baseline(cfg, g, w, bonded, args.cpu_steps)
        threads = num_threads(True)
        vp, dtp = cpu_baseline(cfg, g, w, bonded, omp_steps, parallel=True)
        # the faster of the two builds is the baseline; the other is reported beside it
        serial = {"value": round(v1, 1), "cores": 1,
                  "sample": sample(args.cpu_steps, dt1, "serial FP64 oracle")}
        omp = {"value": round(vp, 1), "cores": threads,
               "sample": sample(omp_steps, dtp, f"OpenMP build of the FP64 oracle, {threads} threads")}
        best = omp if vp >= v1 else serial
        out["cpu_baseline"] = {"value": best["value"], "unit": "particle-steps/s",
                               "cores": best["cores"], "kind": "port", "sample": best["sample"],
                               "serial": serial, "openmp": omp}
    out["setup_s"] = round(t_build, 2)
    if copy_cal:
        out["copy_calibration"] = copy_cal
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
