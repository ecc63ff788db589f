"""ctypes wrapper of the CPU oracle (oracle/dem_oracle.c) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the checker.
``OracleScene`` exposes the same method names as the engine's ``Scene`` so parity tests can drive
both with one script.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from taichi_dem_julia_port_amd import _native as N
from taichi_dem_julia_port_amd.types import (CONTACT_DTYPE, GRAIN_DTYPE, GRAIN_IN_DTYPE,
                                             WALL_DTYPE, make_walls)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
# the same source built with OpenMP: the multi-core CPU baseline (force sums in atomic order)
LIB_OMP_PATH = os.path.join(HERE, "_build", "liboracle_omp.so")


class OrcDerived(ctypes.Structure):
    _fields_ = [("r_max", ctypes.c_double), ("cell_size", ctypes.c_double),
                ("material_density", ctypes.c_double), ("hash_table_size", ctypes.c_int64),
                ("grid_product_i32", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("total_steps", ctypes.c_int64), ("save_per_steps", ctypes.c_int64),
                ("cp_list_capacity", ctypes.c_int64), ("contact_capacity", ctypes.c_int64)]


class OrcCounters(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in (
        "step", "total_cps", "cp_list_capacity", "contact_ptr0", "contact_ptr1",
        "active_contacts", "bonded_contacts", "wall_contacts", "bonds_broken_total",
        "deactivated_total", "fatal_gap", "overflow")]


_LIBS = {}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load(parallel: bool = False) -> ctypes.CDLL:
    """The serial oracle, or with ``parallel`` its OpenMP build."""
    if parallel in _LIBS:
        return _LIBS[parallel]
    path = LIB_OMP_PATH if parallel else LIB_PATH
    if not os.path.exists(path):
        build()
    _LIBS[parallel] = load_path(path)
    return _LIBS[parallel]


def build_mutant(k: int, out_dir: str) -> str:
    """The serial oracle built with planted mistake k (-DORC_MUTATION=k; see dem_oracle.c) into
    out_dir; returns the library path."""
    path = os.path.join(out_dir, f"liboracle_mut{k}.so")
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
                    "-Wno-unknown-pragmas", f"-DORC_MUTATION={int(k)}", "-shared", "-o", path,
                    os.path.join(HERE, "dem_oracle.c"), "-lm"], check=True)
    return path


def load_path(path: str) -> ctypes.CDLL:
    """An oracle build at path, with every function's signature declared."""
    lib = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    sig = {
        "orc_morton3": (ctypes.c_uint64, [ctypes.c_int64] * 3),
        "orc_cartesian3morton": (ctypes.c_int64, [P(ctypes.c_int32)]),
        "orc_cell": (None, [P(ctypes.c_double), P(ctypes.c_double), ctypes.c_double, P(ctypes.c_int32)]),
        "orc_hashcode": (ctypes.c_uint32, [P(ctypes.c_int32), ctypes.c_int64]),
        "orc_center": (None, [P(ctypes.c_int32), P(ctypes.c_double), ctypes.c_double, P(ctypes.c_double)]),
        "orc_neighbors": (None, [P(ctypes.c_double), P(ctypes.c_int32), P(ctypes.c_double),
                                 ctypes.c_int64, P(ctypes.c_uint32)]),
        "orc_derive": (ctypes.c_int32, [P(N.DemConfig), P(N.DemGrainIn), ctypes.c_int64, P(OrcDerived)]),
        "orc_create": (vp, [P(N.DemConfig), P(N.DemWall), ctypes.c_int32, P(N.DemGrainIn), ctypes.c_int64]),
        "orc_destroy": (None, [vp]),
        "orc_get_derived": (None, [vp, P(OrcDerived)]),
        "orc_initialize": (ctypes.c_int32, [vp]),
        "orc_step": (ctypes.c_int32, [vp, ctypes.c_int64]),
        "orc_num_threads": (ctypes.c_int32, []),
        "orc_contact_pass": (None, [vp]),
        "orc_get_counters": (None, [vp, P(OrcCounters)]),
        "orc_get_grains": (None, [vp, P(N.DemGrain)]),
        "orc_set_grains": (None, [vp, P(N.DemGrain)]),
        "orc_num_active_contacts": (ctypes.c_int64, [vp]),
        "orc_get_contacts": (ctypes.c_int64, [vp, P(N.DemContact), P(ctypes.c_uint8), ctypes.c_int64]),
        "orc_set_contacts": (ctypes.c_int32, [vp, P(N.DemContact), P(ctypes.c_uint8), ctypes.c_int64]),
        "orc_get_wall_shear": (None, [vp, P(ctypes.c_double)]),
        "orc_set_wall_shear": (None, [vp, P(ctypes.c_double)]),
        "orc_get_forces": (None, [vp, P(ctypes.c_double), P(ctypes.c_double)]),
        "orc_set_track_scale": (None, [vp, ctypes.c_int32]),
        "orc_get_force_scale": (None, [vp, P(ctypes.c_double), P(ctypes.c_double)]),
        "orc_get_cells": (None, [vp, P(ctypes.c_int32), P(ctypes.c_uint32), P(ctypes.c_double),
                                 P(ctypes.c_uint32)]),
        "orc_get_hash_table": (None, [vp, P(ctypes.c_uint32), P(ctypes.c_uint32), P(ctypes.c_uint32)]),
        "orc_get_cp_list": (ctypes.c_int64, [vp, P(ctypes.c_int32), ctypes.c_int64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def num_threads(parallel: bool = False) -> int:
    """Threads one oracle step uses (OpenMP's count for the parallel build)."""
    return int(load(parallel).orc_num_threads())


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def derive(cfg, grains) -> dict:
    lib = load()
    g = np.ascontiguousarray(grains, dtype=GRAIN_IN_DTYPE)
    c = cfg.to_c()
    d = OrcDerived()
    lib.orc_derive(ctypes.byref(c), _p(g, N.DemGrainIn), len(g), ctypes.byref(d))
    return {f: getattr(d, f) for f, _ in OrcDerived._fields_}


class OracleScene:
    """Serial FP64 CPU restatement of the reference step (same method names as Scene).
    ``parallel=True`` runs the OpenMP build of the same code."""

    def __init__(self, cfg, grains, walls=None, parallel: bool = False, lib=None):
        self._lib = lib if lib is not None else load(parallel)
        self.cfg = cfg
        g = np.ascontiguousarray(grains, dtype=GRAIN_IN_DTYPE)
        self.n = len(g)
        if walls is None:
            walls = [(cfg.wall_normal, cfg.wall_distance, 2)]
        w = walls if isinstance(walls, np.ndarray) and walls.dtype == WALL_DTYPE else make_walls(walls)
        self.walls = np.ascontiguousarray(w)
        self.nwall = len(self.walls)
        c = cfg.to_c()
        self._h = self._lib.orc_create(ctypes.byref(c), _p(self.walls, N.DemWall), self.nwall,
                                       _p(g, N.DemGrainIn), self.n)
        if not self._h:
            raise RuntimeError("orc_create failed")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def initialize(self):
        rc = self._lib.orc_initialize(self._h)
        if rc:
            raise RuntimeError(f"oracle initialize failed {rc}")

    def step(self, nsteps=1):
        rc = self._lib.orc_step(self._h, int(nsteps))
        if rc:
            raise RuntimeError(f"oracle step failed {rc}")

    def contact_pass(self):
        self._lib.orc_contact_pass(self._h)

    def derived(self):
        d = OrcDerived()
        self._lib.orc_get_derived(self._h, ctypes.byref(d))
        return {f: getattr(d, f) for f, _ in OrcDerived._fields_}

    def counters(self):
        c = OrcCounters()
        self._lib.orc_get_counters(self._h, ctypes.byref(c))
        return {f: getattr(c, f) for f, _ in OrcCounters._fields_}

    def read_grains(self):
        out = np.zeros(self.n, dtype=GRAIN_DTYPE)
        self._lib.orc_get_grains(self._h, _p(out, N.DemGrain))
        return out

    def write_grains(self, g):
        a = np.ascontiguousarray(g, dtype=GRAIN_DTYPE)
        self._lib.orc_set_grains(self._h, _p(a, N.DemGrain))

    def read_contacts(self):
        m = self._lib.orc_num_active_contacts(self._h)
        out = np.zeros(m, dtype=CONTACT_DTYPE)
        bd = np.zeros(m, dtype=np.uint8)
        self._lib.orc_get_contacts(self._h, _p(out, N.DemContact), _p(bd, ctypes.c_uint8), m)
        return out, bd.astype(bool)

    def write_contacts(self, contacts, bonded):
        c = np.ascontiguousarray(contacts, dtype=CONTACT_DTYPE)
        b = np.ascontiguousarray(bonded, dtype=np.uint8)
        rc = self._lib.orc_set_contacts(self._h, _p(c, N.DemContact), _p(b, ctypes.c_uint8), len(c))
        if rc:
            raise RuntimeError("orc_set_contacts failed")

    def read_wall_shear(self):
        out = np.zeros((self.n, self.nwall, 3))
        self._lib.orc_get_wall_shear(self._h, _p(out, ctypes.c_double))
        return out

    def write_wall_shear(self, s):
        a = np.ascontiguousarray(s, dtype=np.float64).reshape(self.n, self.nwall, 3)
        self._lib.orc_set_wall_shear(self._h, _p(a, ctypes.c_double))

    def read_forces(self):
        F = np.zeros((self.n, 3))
        T = np.zeros((self.n, 3))
        self._lib.orc_get_forces(self._h, _p(F, ctypes.c_double), _p(T, ctypes.c_double))
        return F, T

    def track_force_scale(self, on: bool = True):
        """Keep sums of |contribution| per force and moment component from the next step on."""
        self._lib.orc_set_track_scale(self._h, 1 if on else 0)
        self.tracking_force_scale = bool(on)

    def read_force_scale(self):
        """Sums of |contribution| per component of the last step's forces and moments."""
        F = np.zeros((self.n, 3))
        T = np.zeros((self.n, 3))
        self._lib.orc_get_force_scale(self._h, _p(F, ctypes.c_double), _p(T, ctypes.c_double))
        return F, T

    # intermediate arrays of the most recent contact pass (call contact_pass() first)
    def cells(self):
        kid = np.zeros((self.n, 3), dtype=np.int32)
        hid = np.zeros(self.n, dtype=np.uint32)
        kc = np.zeros((self.n, 3))
        nbr = np.zeros((self.n, 8), dtype=np.uint32)
        self._lib.orc_get_cells(self._h, _p(kid, ctypes.c_int32), _p(hid, ctypes.c_uint32),
                                _p(kc, ctypes.c_double), _p(nbr, ctypes.c_uint32))
        return kid, hid, kc, nbr

    def hash_table(self):
        H = self.derived()["hash_table_size"]
        cnt = np.zeros(H, dtype=np.uint32)
        cur = np.zeros(H, dtype=np.uint32)
        pid = np.zeros(self.n, dtype=np.uint32)
        self._lib.orc_get_hash_table(self._h, _p(cnt, ctypes.c_uint32), _p(cur, ctypes.c_uint32),
                                     _p(pid, ctypes.c_uint32))
        return cnt, cur, pid

    def cp_list(self):
        m = self._lib.orc_get_cp_list(self._h, None, 0)
        out = np.zeros((m, 2), dtype=np.int32)
        self._lib.orc_get_cp_list(self._h, _p(out, ctypes.c_int32), m)
        return out
