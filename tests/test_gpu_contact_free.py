"""The contact-free mode of a single scene.

While the last counted step left no row entry, a single scene's force kernel runs a variant that
reads and walks no rows, and a merge that would give a grain a row entry fails its step instead,
which the host replays with the full force kernel.  The mode changes no result: the same grains
stepped as a one-scene batch (which never enters it) stay bit for bit equal through calls in the
mode, the steps where contacts appear, and after; and the scene matches the oracle.

The sample is the pour sample at its own radii (no contacts) with speeds up to 0.5 m/s: no pair
touches in the first call, so the second call starts contact-free, and the first contacts form
some tens of steps later.  The floor sample moves the same grains down until the lowest press into
the floor, so the contact-free steps also carry wall contacts (whose history a one-kernel lean step
writes only once its merge blocks are done and the step has not failed).
"""
import numpy as np
import pytest

from helpers import compare_state, packed_sample, pair_keys, reseed_oracle

pytestmark = pytest.mark.gpu

GRAIN_FIELDS = ("k", "v", "w", "q", "a", "dw")
CONTACT_FIELDS = ("i", "j", "k", "F", "s", "ti", "tj")


def _sample():
    return packed_sample(2000, scale=1.0, vmax=0.5, seed=5)


def _floor_sample():
    """The grains moved down until the lowest one's centre is 0.7 radii above the floor (z = 0)."""
    cfg, g, w = _sample()
    g = g.copy()
    r = np.cbrt(3.0 * g["V"] / (4.0 * np.pi))
    low = np.argmin(g["k"][:, 2] - r)
    g["k"][:, 2] -= g["k"][low, 2] - 0.7 * r[low]
    return cfg, g, w


def _bits(x):
    x = np.ascontiguousarray(x)
    return x.view(np.uint64) if x.dtype == np.float64 else x


@pytest.mark.parametrize("precision,graph,floor", [
    ("fp64", True, False), ("fp64", False, False), ("mixed", True, False), ("mixed", False, False),
    ("fp64", True, True), ("mixed", False, True)])
def test_contact_free_mode_equals_batched_scene(precision, graph, floor):
    from taichi_dem_julia_port_amd.executor import Batch, Scene
    cfg, g, w = _floor_sample() if floor else _sample()
    print(f"[case] contact-free {precision} graph={graph} floor={floor} n={len(g)}", flush=True)
    seen_free = seen_contacts = seen_free_walls = False
    with Scene(cfg, g, w, precision=precision, graph=graph) as s, \
            Batch([(cfg, g, w)], precision=precision, graph=graph) as b:
        for k in (1, 4, 16, 40, 100, 200):
            before = s.counters()["contacts"] if k > 1 else None
            s.step(k)
            b.step(k)
            gs, gb = s.read_grains(), b.read_grains()
            for f in GRAIN_FIELDS:
                assert np.array_equal(_bits(gs[f]), _bits(gb[f])), f
            cs, bs = s.read_contacts()
            cb, bb = b.read_contacts()
            assert np.array_equal(pair_keys(cs), pair_keys(cb)), "contact set"
            assert np.array_equal(bs, bb)
            for f in CONTACT_FIELDS:
                assert np.array_equal(_bits(cs[f]), _bits(cb[f])), f
            Fs, Ts = s.read_forces()
            Fb, Tb = b.read_forces()
            assert np.array_equal(_bits(Fs), _bits(Fb)) and np.array_equal(_bits(Ts), _bits(Tb))
            c = s.counters()
            assert c["contacts"] == b.counters()["contacts"]
            seen_free |= before == 0
            seen_free_walls |= before == 0 and c["wall_contacts"] > 0
            seen_contacts |= c["contacts"] > 0
            print(f"[case] step {c['step']}: {c['contacts']} contacts, {c['wall_contacts']} at walls",
                  flush=True)
    # a call started with no contact (in the mode) and contacts formed later (it was left)
    assert seen_free and seen_contacts
    assert seen_free_walls or not floor


def test_contact_free_mode_matches_the_oracle():
    from taichi_dem_julia_port_amd.executor import Scene
    cfg, g, w = _sample()
    print(f"[case] contact-free against the oracle n={len(g)}", flush=True)
    with Scene(cfg, g, w) as s:
        s.step(1)
        for k in (4, 16, 20, 20, 20):
            o = reseed_oracle(s, cfg, g, w)
            s.step(k)
            o.step(k)
            compare_state(s, o)
            print(f"[case] step {s.counters()['step']}: {s.counters()['contacts']} contacts", flush=True)
