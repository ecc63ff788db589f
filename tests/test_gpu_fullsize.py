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
        s.set_diagnostics(True)
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


def test_full_batch_scenes_equal_alone():
    """The 256-scene ensemble (1.27M grains) at its benchmark shape: a sample of its scenes,
    stepped in the batch, equals the same scenes stepped alone bit for bit."""
    from taichi_dem_julia_port_amd.executor import Batch, Scene
    members = scenes.batched_cubes(256, seed=4)
    sizes = np.array([len(m[1]) for m in members])
    sample = sorted({0, int(sizes.argmax()), int(sizes.argmin()), 137})
    print(f"[case] batch of 256, n={sizes.sum()}, sample {sample}", flush=True)
    with Batch(members) as b:
        b.initialize()
        b.step(30)
        gb = b.split_grains(b.read_grains())
        cb = b.split_contacts(*b.read_contacts())
        broken = b.bonds_broken()
    for k in sample:
        cfg, g, w = members[k]
        with Scene(cfg, g, w) as s:
            s.initialize()
            s.step(30)
            ga = s.read_grains()
            ca, ba = s.read_contacts()
            assert broken[k] == s.counters()["bonds_broken_total"], k
        for f in ("k", "v", "w", "q"):
            assert np.array_equal(ga[f].view(np.uint64), gb[k][f].view(np.uint64)), (k, f)
        c, bd = cb[k]
        assert np.array_equal(pair_keys(c), pair_keys(ca)) and np.array_equal(bd, ba), k
        for f in ("F", "ti", "tj", "s"):
            assert np.array_equal(c[f].view(np.uint64), ca[f].view(np.uint64)), (k, f)
