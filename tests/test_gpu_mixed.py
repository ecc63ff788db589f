"""Mixed precision (FP32 candidate filter and Hertz-Mindlin law, compensated FP32 force sums,
FP64 spring, bond and integration state) against the FP64 oracle and the FP64 engine.

Bars: contact and bond sets identical (the FP32 filter only pre-screens; the decision stays the
FP64 test); per-contact Hertz forces within 2e-5 relative; one-step displacement within 1e-4
relative; after 1000 pour steps, position error vs the FP64 engine <= 1e-6 relative.
"""
import numpy as np
import pytest

from helpers import packed_sample, pair_keys, pour_sample, reference_case, vec_rel_err
from oracle.oracle import OracleScene

pytestmark = pytest.mark.gpu


def _scene(*a, **k):
    from taichi_dem_julia_port_amd.executor import Scene
    return Scene(*a, **k)


@pytest.mark.parametrize("refgrid", [False, True])
@pytest.mark.parametrize("name", ["packed_4k", "cfg2_nobond", "cfg2_bonded"])
def test_mixed_one_step_against_oracle(name, refgrid):
    print(f"[case] mixed {name} refgrid={refgrid}", flush=True)
    if name == "packed_4k":
        cfg, g, w = packed_sample(4000)
    else:
        cfg, g, w = reference_case(2)
    o = OracleScene(cfg, g, w)
    with _scene(cfg, g, w, precision="mixed", reference_grid=refgrid) as s:
        if name == "cfg2_bonded":
            o.initialize()
            s.initialize()
        g0 = o.read_grains()
        s.step(1)
        o.step(1)
        c, b = s.read_contacts()
        co, bo = o.read_contacts()
        assert np.array_equal(pair_keys(c), pair_keys(co))
        assert np.array_equal(b, bo)
        assert vec_rel_err(c["F"], co["F"]) <= 2e-5
        gs, go = s.read_grains(), o.read_grains()
    dk_s = gs["k"] - g0["k"]
    dk_o = go["k"] - g0["k"]
    dv_s = gs["v"] - g0["v"]
    dv_o = go["v"] - g0["v"]
    assert np.abs(dv_s - dv_o).max() <= 1e-4 * np.abs(dv_o).max()
    assert np.abs(dk_s - dk_o).max() <= 1e-4 * np.abs(dk_o).max()


def test_mixed_1000_steps_position_error_vs_fp64():
    print("[case] mixed vs fp64 packed pour 1000 steps", flush=True)
    cfg, g, w = packed_sample(6000, scale=1.15, vmax=0.02)
    with _scene(cfg, g, w, precision="fp64") as a, _scene(cfg, g, w, precision="mixed") as b:
        a.step(1000)
        b.step(1000)
        ga, gb = a.read_grains(), b.read_grains()
        ca, cb = a.counters(), b.counters()
    assert ca["contacts"] > 0 and cb["contacts"] > 0
    rel = np.linalg.norm(ga["k"] - gb["k"], axis=1) / np.linalg.norm(ga["k"], axis=1)
    assert rel.max() <= 1e-6, rel.max()


def test_mixed_bonded_cube_1000_steps_vs_fp64():
    """Bonds in the mixed path (FP32 frame, stiffness and increments on FP64 state): the
    900-grain bonded cube (no bond breaks in 1000 steps) against the FP64 engine."""
    print("[case] mixed vs fp64 bonded cube 1000 steps", flush=True)
    from taichi_dem_julia_port_amd import scenes
    cfg, g, w = scenes.cube_small()
    with _scene(cfg, g, w, precision="fp64") as a, _scene(cfg, g, w, precision="mixed") as b:
        a.initialize()
        b.initialize()
        a.step(1000)
        b.step(1000)
        ga, gb = a.read_grains(), b.read_grains()
        ca, cb = a.counters(), b.counters()
        (c1, b1), (c2, b2) = a.read_contacts(), b.read_contacts()
    assert ca["bonded"] == cb["bonded"] > 0
    assert np.array_equal(pair_keys(c1), pair_keys(c2)) and np.array_equal(b1, b2)
    rel = np.linalg.norm(ga["k"] - gb["k"], axis=1) / np.linalg.norm(ga["k"], axis=1)
    assert rel.max() <= 1e-6, rel.max()
    assert vec_rel_err(c2["F"], c1["F"]) <= 1e-3
