"""Parity at the benchmark sizes: the 1M-grain pour (contact-free at first) and its contact-heavy
state (about 910k Hertz contacts), stepped on the GPU and by the OpenMP build of the FP64 oracle.
After each of two steps the contact sets are equal, contact forces, springs, positions and
velocities agree within 1e-12 relative, and the engine's count of the reference's cp_list equals
the oracle's total_cps."""
import numpy as np
import pytest

from helpers import pair_keys, vec_rel_err
from oracle.oracle import OracleScene
from taichi_dem_julia_port_amd import scenes

pytestmark = pytest.mark.gpu

REL = 1e-12


@pytest.mark.parametrize("name", ["pour", "pour_packed"])
def test_million_grain_steps_match_the_oracle(name):
    from taichi_dem_julia_port_amd.executor import Scene
    cfg, g, w = scenes.pour(1_000_000) if name == "pour" else scenes.pour_packed(1_000_000)
    print(f"[case] {name} n={len(g)}", flush=True)
    o = OracleScene(cfg, g, w, parallel=True)
    with Scene(cfg, g, w) as s:
        for step in (1, 2):
            o.step(1)
            s.step(1)
            c, b = s.read_contacts()
            co, bo = o.read_contacts()
            assert np.array_equal(pair_keys(c), pair_keys(co)), (name, step)
            assert np.array_equal(b, bo) and not b.any(), (name, step)
            for f in ("F", "s"):
                assert vec_rel_err(c[f], co[f]) <= REL, (name, step, f)
            gs, go = s.read_grains(), o.read_grains()
            for f in ("k", "v"):
                assert vec_rel_err(gs[f], go[f]) <= REL, (name, step, f)
            assert s.counters()["candidate_pairs"] == o.counters()["total_cps"], (name, step)
            print(f"[case] {name} step {step}: {len(c)} contacts equal", flush=True)
        if name == "pour_packed":
            assert len(c) > 800_000
    o.close()
