"""The OpenMP build of the oracle (the multi-core CPU baseline) against the serial oracle.

Both builds come from one source (oracle/dem_oracle.c).  The parallel one sums contact forces
with atomic adds in whatever order the threads reach them, so positions agree to rounding, while
the contact and bond sets -- decided in pair order in both -- agree exactly.
"""
import numpy as np

from helpers import pair_keys
from oracle import oracle as O
from taichi_dem_julia_port_amd import scenes


def _run(cfg, g, w, steps, bonded, parallel):
    o = O.OracleScene(cfg, g, w, parallel=parallel)
    if bonded:
        o.initialize()
    o.step(steps)
    out = o.read_grains(), o.read_contacts(), o.counters()
    o.close()
    return out


def _compare(cfg, g, w, steps, bonded, tol):
    (ga, (ca, ba), na) = _run(cfg, g, w, steps, bonded, False)
    (gb, (cb, bb), nb) = _run(cfg, g, w, steps, bonded, True)
    ka, kb = pair_keys(ca), pair_keys(cb)
    ia, ib = np.argsort(ka), np.argsort(kb)
    assert np.array_equal(ka[ia], kb[ib]), "contact sets differ"
    assert np.array_equal(ba[ia], bb[ib]), "bond flags differ"
    for key in ("total_cps", "active_contacts", "bonded_contacts", "wall_contacts",
                "bonds_broken_total", "deactivated_total", "fatal_gap", "overflow"):
        assert na[key] == nb[key], key
    scale = np.abs(ga["k"]).max()
    assert np.abs(ga["k"] - gb["k"]).max() <= tol * scale
    return na


def test_thread_count():
    assert O.num_threads(False) == 1
    assert O.num_threads(True) >= 1


def test_bonded_cube_matches_serial():
    cfg, g, w = scenes.cube_small()
    c = _compare(cfg, g, w, 300, True, 1e-12)
    assert c["bonded_contacts"] > 0


def test_fracturing_cube_matches_serial():
    # weak bonds: about 2000 of 2420 break on impact, and Hertz contacts form and separate
    cfg, g, w = scenes.cube_small(bond_tensile_strength=3e5, bond_shear_strength=3e5)
    c = _compare(cfg, g, w, 600, True, 1e-9)
    assert c["bonds_broken_total"] > 1000 and c["deactivated_total"] > 0


def test_packed_pour_matches_serial():
    cfg, g, w = scenes.pour_packed(3000, box=(0.015, 0.015, 0.02))
    c = _compare(cfg, g, w, 40, False, 1e-12)
    assert c["active_contacts"] > 0 and c["total_cps"] > 0
