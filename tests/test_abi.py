"""The C ABI: the library loads, exports every function include/demhip.h declares, and the
ctypes / numpy mirrors have the C layouts (checked against gcc's view of the header).
No GPU work is done here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from taichi_dem_julia_port_amd import _native as N
from taichi_dem_julia_port_amd.types import CONTACT_DTYPE, GRAIN_DTYPE, GRAIN_IN_DTYPE, WALL_DTYPE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "demhip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(dem_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_hot_path():
    fns = declared_functions()
    for f in ("dem_create", "dem_initialize", "dem_step", "dem_read_grains", "dem_read_contacts",
              "dem_destroy", "dem_last_error", "dem_abi_version"):
        assert f in fns


def test_library_exports_every_declared_symbol():
    lib = N.load_library()
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing
    assert {name for name, _, _ in N.SIGNATURES} == set(declared_functions())
    assert lib.dem_abi_version() == 2
    assert lib.dem_kernel_name(3) == b"broad_phase"


C_PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "demhip.h"
#define P(T, f) printf(#T " " #f " %zu\n", offsetof(T, f));
int main(void) {
  printf("sizes %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(dem_config_t), sizeof(dem_wall_t),
         sizeof(dem_grain_in_t), sizeof(dem_grain_t), sizeof(dem_contact_t), sizeof(dem_derived_t),
         sizeof(dem_options_t), sizeof(dem_counters_t));
  P(dem_config_t, max_coordinate_number) P(dem_config_t, tolerance) P(dem_config_t, gravity)
  P(dem_config_t, saving_interval_time) P(dem_grain_t, k) P(dem_grain_t, q) P(dem_grain_t, I)
  P(dem_contact_t, k) P(dem_contact_t, s) P(dem_derived_t, total_steps)
  return 0;
}
"""


def test_ctypes_layout_matches_c(tmp_path):
    src = tmp_path / "probe.c"
    src.write_text(C_PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = list(map(int, out[0].split()[1:]))
    assert sizes == [ctypes.sizeof(t) for t in (N.DemConfig, N.DemWall, N.DemGrainIn, N.DemGrain,
                                                N.DemContact, N.DemDerived, N.DemOptions,
                                                N.DemCounters)]
    cls = {"dem_config_t": N.DemConfig, "dem_grain_t": N.DemGrain, "dem_contact_t": N.DemContact,
           "dem_derived_t": N.DemDerived}
    for line in out[1:]:
        if not line:
            continue
        t, f, off = line.split()
        assert getattr(cls[t], f).offset == int(off), (t, f)
    # numpy mirrors
    assert GRAIN_DTYPE.itemsize == ctypes.sizeof(N.DemGrain)
    assert CONTACT_DTYPE.itemsize == ctypes.sizeof(N.DemContact)
    assert WALL_DTYPE.itemsize == ctypes.sizeof(N.DemWall)
    assert GRAIN_IN_DTYPE.itemsize == ctypes.sizeof(N.DemGrainIn)
    assert GRAIN_DTYPE.fields["q"][1] == N.DemGrain.q.offset
    assert CONTACT_DTYPE.fields["s"][1] == N.DemContact.s.offset


def test_pour_generator_is_deterministic_and_non_overlapping():
    from taichi_dem_julia_port_amd import scenes
    a = scenes.pour_grains(3000, box=(0.03, 0.03, 0.03), seed=11)
    b = scenes.pour_grains(3000, box=(0.03, 0.03, 0.03), seed=11)
    assert np.array_equal(a, b)
    r = np.cbrt(a["V"] / (4 / 3 * np.pi))
    assert r.min() >= 0.25e-3 * (1 - 1e-12) and r.max() <= 1.0e-3 * (1 + 1e-12)
    k = a["k"]
    assert (k - r[:, None] >= -1e-15).all() and (k + r[:, None] <= 0.03 + 1e-15).all()
    from scipy.spatial import cKDTree
    t = cKDTree(k)
    for i, j in t.query_pairs(2.0e-3):
        assert np.linalg.norm(k[i] - k[j]) > r[i] + r[j]


def test_create_rejects_bad_input_without_gpu_work():
    lib = N.load_library()
    h = ctypes.c_void_p()
    cfg = N.DemConfig()
    cfg.dt = 0.0  # invalid
    g = (N.DemGrainIn * 1)()
    rc = lib.dem_create(ctypes.byref(cfg), None, 0, g, 1, None, ctypes.byref(h))
    assert rc == N.DEM_ERR_INVALID
    assert b"dt" in lib.dem_last_error(None)


def _valid_cfg():
    from taichi_dem_julia_port_amd.config import DEMConfig
    return DEMConfig(domain_min=[-1.0, -1.0, -1.0], domain_max=[1.0, 1.0, 1.0], dt=1e-6).to_c()


def _grains(n=2):
    g = (N.DemGrainIn * n)()
    for i in range(n):
        g[i].id, g[i].gid, g[i].V, g[i].m = i + 1, 1, 1e-6, 2.5e-3
        g[i].k[:] = [0.1 * i, 0.0, 0.0]
    return g


@pytest.mark.parametrize("case, message", [
    ("n0", b"n out of range"),
    ("volume", b"volume and mass"),
    ("nwall", b"nwall must be in [0, 16]"),
    ("wall_mid", b"wall material id"),
    ("precision", b"unknown precision"),
    ("slots", b"contact_slots must be in [1, 256]"),
    ("out_null", b"out is NULL"),
])
def test_create_validates_before_any_gpu_work(case, message):
    """dem_create's checks all run before the device is touched, so they hold without a GPU."""
    lib = N.load_library()
    h = ctypes.c_void_p()
    cfg = _valid_cfg()
    g = _grains()
    n, nwall = 2, 1
    walls = (N.DemWall * 17)()
    for w in walls:
        w.n[:] = [0.0, 0.0, 1.0]
        w.mid = 2
    opt = N.DemOptions(N.DEM_PRECISION_FP64, 0, 0, 0)
    out = ctypes.byref(h)
    if case == "n0":
        n = 0
    elif case == "volume":
        g[1].V = 0.0
    elif case == "nwall":
        nwall = 17
    elif case == "wall_mid":
        walls[0].mid = 3
    elif case == "precision":
        opt.precision = 7
    elif case == "slots":
        opt.contact_slots = 300
    elif case == "out_null":
        out = None
    rc = lib.dem_create(ctypes.byref(cfg), walls, nwall, g, n, ctypes.byref(opt), out)
    assert rc == N.DEM_ERR_INVALID
    assert message in lib.dem_last_error(None), lib.dem_last_error(None)
    assert not h.value


def test_batch_entry_points_validate_before_any_gpu_work():
    lib = N.load_library()
    h = ctypes.c_void_p()
    cfg = (N.DemConfig * 1)(_valid_cfg())
    nw = (ctypes.c_int32 * 1)(0)
    cnt = (ctypes.c_int64 * 1)(2)
    g = _grains()
    rc = lib.dem_create_batch(0, cfg, nw, None, cnt, g, None, ctypes.byref(h))
    assert rc == N.DEM_ERR_INVALID and b"1..65535 scenes" in lib.dem_last_error(None)
    opt = N.DemOptions(N.DEM_PRECISION_FP64, 0, 0, N.DEM_FLAG_REFERENCE_GRID)
    rc = lib.dem_create_batch(1, cfg, nw, None, cnt, g, ctypes.byref(opt), ctypes.byref(h))
    assert rc == N.DEM_ERR_INVALID and b"dense sort grid" in lib.dem_last_error(None)
    rc = lib.dem_batch_repack(None, 0, None, 1, cfg, nw, None, cnt, g, None, ctypes.byref(h))
    assert rc == N.DEM_ERR_INVALID and b"src is NULL" in lib.dem_last_error(None)
    assert not h.value


def test_null_handles_are_rejected():
    lib = N.load_library()
    cnt = ctypes.c_int64(0)
    assert lib.dem_step(None, 1) == N.DEM_ERR_INVALID
    assert lib.dem_initialize(None) == N.DEM_ERR_INVALID
    assert lib.dem_get_derived(None, ctypes.byref(N.DemDerived())) == N.DEM_ERR_INVALID
    assert lib.dem_get_counters(None, ctypes.byref(N.DemCounters())) == N.DEM_ERR_INVALID
    assert lib.dem_read_grains(None, None, 0) == N.DEM_ERR_INVALID
    assert lib.dem_read_contacts(None, None, None, 0, ctypes.byref(cnt)) == N.DEM_ERR_INVALID
    assert lib.dem_write_contacts(None, None, None, 0) == N.DEM_ERR_INVALID
    assert lib.dem_batch_offsets(None, None) == N.DEM_ERR_INVALID
    assert lib.dem_batch_adopt(None, 0, None, 0) == N.DEM_ERR_INVALID
    assert lib.dem_set_memory_cache(-1) == N.DEM_ERR_INVALID
    lib.dem_destroy(None)  # a no-op
