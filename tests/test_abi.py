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
    assert lib.dem_abi_version() == 1
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
