"""Segmented batches: every scene of a batch evolves exactly as it does alone.

A batch shares one set of arrays and one launch per kernel stage; per-particle scene ids select
each scene's grid, materials, walls and time step.  Contact rows are summed in partner order and
the contact set does not depend on the sort grid, so each scene's state in the batch must equal
its state alone bit for bit -- grains, contacts, bonds and bond breaks.
"""
import numpy as np
import pytest

from helpers import packed_sample, pair_keys, reference_case
from oracle.oracle import OracleScene
from taichi_dem_julia_port_amd import scenes

pytestmark = pytest.mark.gpu


def _engine():
    from taichi_dem_julia_port_amd.executor import Batch, Scene
    return Batch, Scene


def _members(nscenes=6):
    m = scenes.batched_cubes(nscenes, seed=9, nmin=200, nmax=3000, base_gap=2e-5)
    m.append(packed_sample(2000, seed=21))
    m.append(reference_case(2))
    return m


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_scene_equals_scene_alone(precision):
    Batch, Scene = _engine()
    members = _members()
    bond_mask = [1] * (len(members) - 2) + [0, 1]
    print(f"[case] batch of {len(members)} scenes, {precision}", flush=True)
    with Batch(members, precision=precision) as b:
        b.set_bond_init(bond_mask)
        b.initialize()
        b.step(37)
        gb = b.split_grains(b.read_grains())
        cb = b.split_contacts(*b.read_contacts())
        broken_b = b.bonds_broken()
    for k, (cfg, g, w) in enumerate(members):
        print(f"[case] scene {k} alone (n={len(g)})", flush=True)
        with Scene(cfg, g, w, precision=precision) as s:
            if bond_mask[k]:
                s.initialize()
            s.step(37)
            ga = s.read_grains()
            ca, ba = s.read_contacts()
            broken_a = s.counters()["bonds_broken_total"]
        for f in ("k", "v", "w", "q"):
            assert np.array_equal(_bits(ga[f]), _bits(gb[k][f])), (k, f)
        c, bd = cb[k]
        assert np.array_equal(pair_keys(c), pair_keys(ca)), k
        assert np.array_equal(bd, ba), k
        for f in ("F", "s", "ti", "tj", "k"):
            assert np.array_equal(_bits(c[f]), _bits(ca[f])), (k, f)
        assert broken_b[k] == broken_a, k


def test_batch_against_oracle_one_step():
    """Batched kernels read per-scene parameters: check a mixed batch against the oracle."""
    Batch, _ = _engine()
    members = _members(3)
    with Batch(members) as b:
        b.set_bond_init([1, 1, 1, 0, 1])
        b.initialize()
        b.step(1)
        gb = b.split_grains(b.read_grains())
        cb = b.split_contacts(*b.read_contacts())
    for k, (cfg, g, w) in enumerate(members):
        o = OracleScene(cfg, g, w)
        if k != 3:
            o.initialize()
        o.step(1)
        co, bo = o.read_contacts()
        c, bd = cb[k]
        assert np.array_equal(pair_keys(c), pair_keys(co)), k
        assert np.array_equal(bd, bo), k
        go = o.read_grains()
        err = np.abs(gb[k]["k"] - go["k"]).max() / max(np.abs(go["k"]).max(), 1e-30)
        assert err < 1e-12, (k, err)
