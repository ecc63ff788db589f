"""DEMPort executor API on the MI355X engine.

``Scene`` owns one device-resident GlobalData (src/executor.jl:1-29) held by libdemhip;
``Scene.initialize`` is ``initialize!`` (executor.jl:155-160), ``Scene.step(n)`` is n calls of
``simulate!`` (executor.jl:162-172), and ``solve`` mirrors ``DEMPort.solve`` (executor.jl:174-394):
TOML config, p4p input relative to the config file, optional bond initialisation, the
save-interval loop and the p4p/p4c output.

Every call goes through the C ABI of include/demhip.h; there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes
import logging
import os
import time
from typing import Optional, Sequence

import numpy as np

from . import _native as N
from .config import DEMConfig, from_toml
from .p4io import read_p4p_rows, write_p4c_frame, write_p4p_frame
from .types import CONTACT_DTYPE, GRAIN_DTYPE, GRAIN_IN_DTYPE, WALL_DTYPE, make_walls

log = logging.getLogger("demhip")

PRECISIONS = {"fp64": N.DEM_PRECISION_FP64, "mixed": N.DEM_PRECISION_MIXED}


class DEMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libdemhip error {code}: {msg}")
        self.code = code


def _ptr(arr: np.ndarray, ctype):
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


class Scene:
    """One simulation scene on the GPU (GlobalData)."""

    def __init__(self, cfg: DEMConfig, grains: np.ndarray, walls: Optional[Sequence] = None,
                 precision: str = "fp64", contact_slots: int = 0, device: int = 0,
                 index_owned_broad_phase: bool = False, reference_grid: bool = False,
                 graph: bool = True, near_lists: bool = True):
        self._lib = N.load_library()
        self.cfg = cfg
        grains = np.ascontiguousarray(grains, dtype=GRAIN_IN_DTYPE)
        self.n = len(grains)
        if walls is None:  # executor.jl:313 -- the single configured wall
            walls = [(cfg.wall_normal, cfg.wall_distance, 2)]
        w = walls if isinstance(walls, np.ndarray) and walls.dtype == WALL_DTYPE else make_walls(walls)
        self.walls = np.ascontiguousarray(w)
        self.nwall = len(self.walls)
        self._c_cfg = cfg.to_c()
        flags = (N.DEM_FLAG_INDEX_OWNED_BROAD_PHASE if index_owned_broad_phase else 0) | \
                (N.DEM_FLAG_REFERENCE_GRID if reference_grid else 0) | \
                (0 if graph else N.DEM_FLAG_NO_GRAPH) | \
                (0 if near_lists else N.DEM_FLAG_NO_NEAR_LISTS)
        opt = N.DemOptions(PRECISIONS[precision], int(contact_slots), int(device), flags)
        h = ctypes.c_void_p()
        rc = self._lib.dem_create(ctypes.byref(self._c_cfg), _ptr(self.walls, N.DemWall),
                                  self.nwall, _ptr(grains, N.DemGrainIn), self.n,
                                  ctypes.byref(opt), ctypes.byref(h))
        if rc != 0:
            raise DEMError(rc, self._lib.dem_last_error(None).decode())
        self._h = h
        self.precision = precision

    # -- lifecycle ---------------------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.dem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise DEMError(rc, self._lib.dem_last_error(self._h).decode())

    # -- hot path ------------------------------------------------------------------------------
    def initialize(self) -> None:
        """initialize! (executor.jl:155-160): bonds between pairs closer than r0_i + r0_j."""
        self._check(self._lib.dem_initialize(self._h))

    def step(self, nsteps: int = 1) -> None:
        """nsteps x simulate! (executor.jl:162-172)."""
        self._check(self._lib.dem_step(self._h, int(nsteps)))

    # -- read-back -----------------------------------------------------------------------------
    def derived(self) -> dict:
        d = N.DemDerived()
        self._check(self._lib.dem_get_derived(self._h, ctypes.byref(d)))
        return {f: getattr(d, f) for f, _ in N.DemDerived._fields_}

    def set_diagnostics(self, on: bool) -> None:
        """Count the reference's cp_list length (executor.jl:62-63) and the broad phase's
        examined candidates on the last step of every call (off by default; costs extra kernels
        per call).  Without them ``counters()`` reports both as -1."""
        self._check(self._lib.dem_set_diagnostics(self._h, 1 if on else 0))

    def counters(self) -> dict:
        c = N.DemCounters()
        self._check(self._lib.dem_get_counters(self._h, ctypes.byref(c)))
        return {f: getattr(c, f) for f, _ in N.DemCounters._fields_}

    def read_grains(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=GRAIN_DTYPE)
        self._check(self._lib.dem_read_grains(self._h, _ptr(out, N.DemGrain), self.n))
        return out

    def write_grains(self, grains: np.ndarray) -> None:
        g = np.ascontiguousarray(grains, dtype=GRAIN_DTYPE)
        self._check(self._lib.dem_write_grains(self._h, _ptr(g, N.DemGrain), self.n))

    def read_contacts(self):
        """Active contacts sorted by (i, j) and their bonded flags."""
        cnt = ctypes.c_int64(0)
        self._check(self._lib.dem_read_contacts(self._h, None, None, 0, ctypes.byref(cnt)))
        m = cnt.value
        out = np.zeros(m, dtype=CONTACT_DTYPE)
        bd = np.zeros(m, dtype=np.uint8)
        if m:
            self._check(self._lib.dem_read_contacts(self._h, _ptr(out, N.DemContact),
                                                    _ptr(bd, ctypes.c_uint8), m, ctypes.byref(cnt)))
            if cnt.value != m:
                raise DEMError(-6, "contact count changed between calls")
        return out, bd.astype(bool)

    def write_contacts(self, contacts: np.ndarray, bonded: np.ndarray) -> None:
        c = np.ascontiguousarray(contacts, dtype=CONTACT_DTYPE)
        b = np.ascontiguousarray(bonded, dtype=np.uint8)
        self._check(self._lib.dem_write_contacts(self._h, _ptr(c, N.DemContact),
                                                 _ptr(b, ctypes.c_uint8), len(c)))

    def read_wall_shear(self) -> np.ndarray:
        out = np.zeros((self.n, self.nwall, 3), dtype=np.float64)
        self._check(self._lib.dem_read_wall_shear(self._h, _ptr(out, ctypes.c_double), self.n))
        return out

    def write_wall_shear(self, s: np.ndarray) -> None:
        a = np.ascontiguousarray(s, dtype=np.float64).reshape(self.n, self.nwall, 3)
        self._check(self._lib.dem_write_wall_shear(self._h, _ptr(a, ctypes.c_double), self.n))

    def read_forces(self):
        F = np.zeros((self.n, 3))
        T = np.zeros((self.n, 3))
        self._check(self._lib.dem_read_forces(self._h, _ptr(F, ctypes.c_double),
                                              _ptr(T, ctypes.c_double), self.n))
        return F, T

    # -- intermediate arrays of the contact pass -----------------------------------------------
    def debug_cells(self):
        kid = np.zeros((self.n, 3), dtype=np.int32)
        hid = np.zeros(self.n, dtype=np.uint32)
        kc = np.zeros((self.n, 3), dtype=np.float64)
        nbr = np.zeros((self.n, 8), dtype=np.uint32)
        self._check(self._lib.dem_debug_cells(self._h, _ptr(kid, ctypes.c_int32),
                                              _ptr(hid, ctypes.c_uint32), _ptr(kc, ctypes.c_double),
                                              _ptr(nbr, ctypes.c_uint32)))
        return kid, hid, kc, nbr

    def debug_hash_table(self):
        H = self.derived()["hash_table_size"]
        cnt = np.zeros(H, dtype=np.uint32)
        start = np.zeros(H, dtype=np.uint32)
        pid = np.zeros(self.n, dtype=np.uint32)
        self._check(self._lib.dem_debug_hash_table(self._h, _ptr(cnt, ctypes.c_uint32),
                                                   _ptr(start, ctypes.c_uint32),
                                                   _ptr(pid, ctypes.c_uint32)))
        return cnt, start, pid

    def debug_cp_list(self) -> np.ndarray:
        cnt = ctypes.c_int64(0)
        self._check(self._lib.dem_debug_cp_list(self._h, None, 0, ctypes.byref(cnt)))
        m = cnt.value
        out = np.zeros((m, 2), dtype=np.int32)
        if m:
            self._check(self._lib.dem_debug_cp_list(self._h, _ptr(out, ctypes.c_int32), m,
                                                    ctypes.byref(cnt)))
        return out

    # -- timing --------------------------------------------------------------------------------
    def set_timing(self, on: bool) -> None:
        self._check(self._lib.dem_set_timing(self._h, 1 if on else 0))

    def kernel_times(self) -> dict:
        t = N.DemKernelTimes()
        self._check(self._lib.dem_get_kernel_times(self._h, ctypes.byref(t)))
        return {self._lib.dem_kernel_name(k).decode(): (t.ms[k], t.launches[k])
                for k in range(N.DEM_NUM_KERNELS)}


class Batch(Scene):
    """Several independent scenes advanced together: one launch per kernel stage for all of them
    (dem_create_batch).  Particle indices are global; ``offsets[k]:offsets[k+1]`` is scene k."""

    def __init__(self, members, precision: str = "fp64", contact_slots: int = 0, device: int = 0,
                 graph: bool = True, near_lists: bool = True):
        self._lib = N.load_library()
        members = list(members)
        cfgs, nwalls, walls, counts, allw, allg = self._pack(members)
        opt = self._options(precision, contact_slots, device, graph, near_lists)
        h = ctypes.c_void_p()
        rc = self._lib.dem_create_batch(len(members), cfgs, _ptr(nwalls, ctypes.c_int32),
                                        _ptr(allw, N.DemWall), _ptr(counts, ctypes.c_int64),
                                        _ptr(allg, N.DemGrainIn), ctypes.byref(opt), ctypes.byref(h))
        if rc != 0:
            raise DEMError(rc, self._lib.dem_last_error(None).decode())
        self._attach(h, [m[0] for m in members], walls, precision)

    @classmethod
    def repack(cls, src: "Scene", keep, members=(), precision: Optional[str] = None,
               contact_slots: int = 0, device: int = 0, graph: bool = True,
               near_lists: bool = True) -> "Batch":
        """dem_batch_repack: a new batch of scenes ``keep`` of ``src`` (a Batch or a Scene), their
        state moved over on the device, followed by the new ``members`` (cfg, grains, walls).
        ``src`` is left unchanged; the caller closes it."""
        self = cls.__new__(cls)
        self._lib = N.load_library()
        keep = np.ascontiguousarray(np.asarray(keep, dtype=np.int32).reshape(-1))
        members = list(members)
        precision = precision or src.precision
        opt = self._options(precision, contact_slots, device, graph, near_lists)
        h = ctypes.c_void_p()
        if members:
            cfgs, nwalls, walls, counts, allw, allg = self._pack(members)
            args = (cfgs, _ptr(nwalls, ctypes.c_int32), _ptr(allw, N.DemWall),
                    _ptr(counts, ctypes.c_int64), _ptr(allg, N.DemGrainIn))
        else:
            walls, args = [], (None, None, None, None, None)
        rc = self._lib.dem_batch_repack(src._h, len(keep), _ptr(keep, ctypes.c_int32),
                                        len(members), *args, ctypes.byref(opt), ctypes.byref(h))
        if rc != 0:
            raise DEMError(rc, self._lib.dem_last_error(None).decode())
        src_cfgs = getattr(src, "scene_cfgs", [src.cfg])
        src_walls = getattr(src, "scene_walls", [src.walls])
        self._attach(h, [src_cfgs[k] for k in keep] + [m[0] for m in members],
                     [src_walls[k] for k in keep] + walls, precision)
        return self

    @staticmethod
    def _options(precision, contact_slots, device, graph, near_lists=True):
        flags = (0 if graph else N.DEM_FLAG_NO_GRAPH) | (0 if near_lists else N.DEM_FLAG_NO_NEAR_LISTS)
        return N.DemOptions(PRECISIONS[precision], int(contact_slots), int(device), flags)

    @staticmethod
    def _pack(members):
        """The concatenated C arrays of dem_create_batch for scenes (cfg, grains, walls)."""
        S = len(members)
        cfgs = (N.DemConfig * S)()
        nwalls = np.zeros(S, dtype=np.int32)
        counts = np.zeros(S, dtype=np.int64)
        walls, grains = [], []
        for k, (cfg, g, w) in enumerate(members):
            cfgs[k] = cfg.to_c()
            if w is None:
                w = [(cfg.wall_normal, cfg.wall_distance, 2)]
            wa = w if isinstance(w, np.ndarray) and w.dtype == WALL_DTYPE else make_walls(w)
            walls.append(wa)
            nwalls[k] = len(wa)
            g = np.ascontiguousarray(g, dtype=GRAIN_IN_DTYPE)
            grains.append(g)
            counts[k] = len(g)
        allw = np.ascontiguousarray(np.concatenate(walls)) if sum(map(len, walls)) else \
            np.zeros(1, dtype=WALL_DTYPE)
        allg = np.ascontiguousarray(np.concatenate(grains))
        return cfgs, nwalls, walls, counts, allw, allg

    def _attach(self, h, cfgs, walls, precision) -> None:
        self._h = h
        self.scene_cfgs = list(cfgs)
        self.scene_walls = list(walls)
        self.cfg = self.scene_cfgs[0]
        self.nscenes = len(self.scene_cfgs)
        self.nwall = max(len(w) for w in self.scene_walls)
        self.walls = np.ascontiguousarray(np.concatenate(self.scene_walls)) \
            if sum(map(len, self.scene_walls)) else np.zeros(0, dtype=WALL_DTYPE)
        self.precision = precision
        off = np.zeros(self.nscenes + 1, dtype=np.int64)
        self._check(self._lib.dem_batch_offsets(self._h, _ptr(off, ctypes.c_int64)))
        self.offsets = off
        self.n = int(off[-1])

    def set_bond_init(self, mask) -> None:
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8))
        self._check(self._lib.dem_batch_set_bond_init(self._h, _ptr(m, ctypes.c_uint8)))

    def bonds_broken(self) -> np.ndarray:
        out = np.zeros(self.nscenes, dtype=np.int64)
        self._check(self._lib.dem_batch_bonds_broken(self._h, _ptr(out, ctypes.c_int64)))
        return out

    def split_grains(self, grains: np.ndarray):
        return [grains[self.offsets[k]:self.offsets[k + 1]] for k in range(self.nscenes)]

    def split_contacts(self, contacts: np.ndarray, bonded: np.ndarray):
        """Per-scene contact lists with scene-local 1-based indices."""
        scene = np.searchsorted(self.offsets, contacts["i"].astype(np.int64) - 1, side="right") - 1
        out = []
        for k in range(self.nscenes):
            m = scene == k
            c = contacts[m].copy()
            c["i"] -= self.offsets[k]
            c["j"] -= self.offsets[k]
            out.append((c, bonded[m]))
        return out


def set_memory_cache(max_bytes: int) -> None:
    """dem_set_memory_cache: keep up to max_bytes of destroyed scenes' device memory for reuse
    by later scenes (0 frees it and turns the cache off)."""
    lib = N.load_library()
    rc = lib.dem_set_memory_cache(int(max_bytes))
    if rc != 0:
        raise DEMError(rc, "dem_set_memory_cache: max_bytes must be >= 0")


def device_scan(values: np.ndarray):
    """Exclusive prefix sum with the engine's decoupled-lookback scan."""
    lib = N.load_library()
    a = np.ascontiguousarray(values, dtype=np.uint32)
    out = np.zeros_like(a)
    tot = ctypes.c_uint64(0)
    rc = lib.dem_debug_scan(_ptr(a, ctypes.c_uint32), _ptr(out, ctypes.c_uint32), len(a),
                            ctypes.byref(tot))
    if rc != 0:
        raise DEMError(rc, lib.dem_last_error(None).decode())
    return out, tot.value


def solve(cfg_filename: str, no_bond: bool = False, save_snapshot: bool = False,
          save_information: bool = True, precision: str = "fp64", out_dir: str = ".",
          max_steps: Optional[int] = None) -> Scene:
    """DEMPort.solve (executor.jl:174-394) on the MI355X engine.

    ``save_snapshot`` writes ``snapshot_<step>.png`` into ``out_dir`` after every save interval
    (executor.jl:376-378; pictures drawn by ``snapshot.snapshot``).  ``max_steps`` optionally stops
    the loop early (the reference's commented debug switch, executor.jl:360-362).
    """
    t0 = time.perf_counter()
    cfg = from_toml(cfg_filename)
    rows = read_p4p_rows(os.path.join(os.path.dirname(cfg_filename), cfg.init_particles))
    scene = Scene(cfg, rows, precision=precision)
    der = scene.derived()
    log.info("Setting: hash_table_size=%d cell_size=%g", der["hash_table_size"], der["cell_size"])
    total_steps = der["total_steps"] if max_steps is None else min(der["total_steps"], max_steps)
    save_per_steps = max(der["save_per_steps"], 1)
    if not no_bond:
        scene.initialize()
    p4p = p4c = None
    if save_information:
        os.makedirs(out_dir, exist_ok=True)
        p4p = open(os.path.join(out_dir, "output.p4p"), "w")
        p4c = open(os.path.join(out_dir, "output.p4c"), "w")
        write_p4p_frame(p4p, scene.read_grains(), 0.0)
        c, b = scene.read_contacts()
        write_p4c_frame(p4c, c, b, 0.0)
    step = 0
    while step < total_steps:
        t1 = time.perf_counter()
        scene.step(save_per_steps)
        step += save_per_steps
        log.info("Solving... step=%d dt=%.3fs", step, time.perf_counter() - t1)
        if save_snapshot:
            from .snapshot import snapshot
            snapshot(scene.read_grains(), step, out_dir)
        if save_information:
            write_p4p_frame(p4p, scene.read_grains(), step * cfg.dt)
            c, b = scene.read_contacts()
            write_p4c_frame(p4c, c, b, step * cfg.dt)
    if save_information:
        p4p.close()
        p4c.close()
    log.info("Total time: %.3f s", time.perf_counter() - t0)
    return scene
