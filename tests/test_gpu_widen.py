"""Contact rows widen without losing state.

* A step in which a particle needs more row slots than the scene has changes nothing; the engine
  widens the rows (keeping every committed row and its state) and replays the steps that did not
  run, so a scene started with 4 slots ends bit for bit where the same scene started with 64
  slots ends, through graph replays and eager steps.
* A bond pass that overflows on a re-packed batch drops only the rows it built: the scene the
  batch kept from its source keeps its bonds, springs and later trajectory bit for bit.
"""
import numpy as np
import pytest

from helpers import packed_sample, pair_keys
from taichi_dem_julia_port_amd import scenes

pytestmark = pytest.mark.gpu

GRAIN_FIELDS = ("k", "v", "a", "q", "w", "dw")
CONTACT_FIELDS = ("i", "j", "k", "F", "ti", "tj", "s")


def _same_state(a, b, ga=None, gb=None):
    ca, ba = a.read_contacts() if ga is None else ga
    cb, bb = b.read_contacts() if gb is None else gb
    assert np.array_equal(pair_keys(ca), pair_keys(cb)) and np.array_equal(ba, bb)
    for f in CONTACT_FIELDS:
        assert np.array_equal(ca[f], cb[f]), f


@pytest.mark.parametrize("graph", [True, False])
def test_step_overflow_widens_rows_and_matches_wide_rows(graph):
    from taichi_dem_julia_port_amd.executor import Scene
    cfg, g, w = packed_sample(3000)
    print(f"[case] 3000-grain packing, 4 slots against 64, graph={graph}", flush=True)
    with Scene(cfg, g, w, contact_slots=4, graph=graph) as narrow, \
            Scene(cfg, g, w, contact_slots=64, graph=graph) as wide:
        for k in (1, 7, 20):
            narrow.step(k)
            wide.step(k)
            _same_state(narrow, wide)
            gn, gw = narrow.read_grains(), wide.read_grains()
            for f in GRAIN_FIELDS:
                assert np.array_equal(gn[f], gw[f]), f
            Fn, Tn = narrow.read_forces()
            Fw, Tw = wide.read_forces()
            assert np.array_equal(Fn, Fw) and np.array_equal(Tn, Tw)
            assert narrow.counters()["step"] == wide.counters()["step"]
            assert narrow.counters()["overflow"] == 0
        assert narrow.derived()["contact_slots"] >= 8  # the packing has grains with 8 contacts


def test_bond_pass_overflow_keeps_the_kept_scene():
    from taichi_dem_julia_port_amd.executor import Batch
    # scene A: a loaded bonded lattice (at most 6 bonds a grain); scene B: a lattice bonded to its
    # face and edge neighbours (multiplier 1.5: 18 bonds for an interior grain)
    a = scenes.bonded_cube(5, 5, 4, speed=5.0, base_gap=2e-5)
    b = scenes.bonded_cube(4, 4, 4, speed=3.0, base_gap=0.01, seed=3,
                           particle_contact_radius_multiplier=1.5)
    print("[case] re-packed batch: kept bonded lattice + a new scene with 18 bonds a grain",
          flush=True)
    with Batch([a], contact_slots=8) as b0:
        b0.initialize()
        b0.step(5)
        before = b0.read_contacts()
        gb0 = b0.read_grains()
        with Batch.repack(b0, [0], [b], contact_slots=8) as b1:
            b1.initialize()  # the bond pass over B overflows the 8-slot rows
            assert b1.derived()["contact_slots"] > 8
            c1, bd1 = b1.read_contacts()
            kept = b1.split_contacts(c1, bd1)[0]
            assert kept[1].sum() > 200  # A's bonds are still there
            _same_state(b0, b1, before, kept)
            gk = b1.split_grains(b1.read_grains())[0]
            for f in GRAIN_FIELDS:
                assert np.array_equal(gk[f], gb0[f]), f
            # and scene A goes on as it would have in its source batch
            b0.step(6)
            b1.step(6)
            c1, bd1 = b1.read_contacts()
            _same_state(b0, b1, None, b1.split_contacts(c1, bd1)[0])
            g0, gk = b0.read_grains(), b1.split_grains(b1.read_grains())[0]
            for f in GRAIN_FIELDS:
                assert np.array_equal(gk[f], g0[f]), f
            # scene B got its bonds
            assert b1.split_contacts(c1, bd1)[1][1].sum() > 300
