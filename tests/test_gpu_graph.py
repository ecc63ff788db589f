"""The captured HIP graphs (sixteen-step and two-step ones) replay to exactly the state of eager
launches, with lean steps between near-list rebuilds and with the sort and broad phase on every step.

The FP64 and mixed engines are deterministic (contact rows are sorted by partner and every
particle sums its own row in that order), so graph replays must match eager stepping bit for bit
across several dem_step calls, with host synchronisations and read-backs between them.
"""
import numpy as np
import pytest

from helpers import packed_sample, pair_keys
from taichi_dem_julia_port_amd import scenes

pytestmark = pytest.mark.gpu


def _scene(*a, **k):
    from taichi_dem_julia_port_amd.executor import Scene
    return Scene(*a, **k)


@pytest.mark.parametrize("near_lists", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("case", ["lattice", "packed"])
def test_graph_replay_matches_eager(case, precision, near_lists):
    print(f"[case] graph {case} {precision} near_lists={near_lists}", flush=True)
    if case == "lattice":
        cfg, g, w = scenes.bonded_cube(6, 6, 5, speed=10.0, dt=1e-6, base_gap=2e-5)
        bonded = True
    else:
        cfg, g, w = packed_sample(3000)
        bonded = False
    with _scene(cfg, g, w, precision=precision, near_lists=near_lists) as a, \
            _scene(cfg, g, w, precision=precision, graph=False, near_lists=near_lists) as b:
        if bonded:
            a.initialize()
            b.initialize()
        # odd and even lengths: two-step replays plus eager tail steps, then calls long enough
        # for sixteen-step replays from both buffer parities
        for n in (5, 7, 4, 9, 37, 18, 35):
            a.step(n)
            b.step(n)
            ga, gb = a.read_grains(), b.read_grains()
            for f in ("k", "v", "w", "q"):
                assert np.array_equal(ga[f].view(np.uint64), gb[f].view(np.uint64)), f
            ca, ba = a.read_contacts()
            cb, bb = b.read_contacts()
            assert np.array_equal(pair_keys(ca), pair_keys(cb))
            assert np.array_equal(ba, bb)
            for f in ("F", "s", "ti", "tj"):
                assert np.array_equal(ca[f].view(np.uint64), cb[f].view(np.uint64)), f
        assert a.counters()["step"] == 115
