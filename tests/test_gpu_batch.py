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


def _state(b):
    g = b.split_grains(b.read_grains())
    c = b.split_contacts(*b.read_contacts())
    return g, c, b.bonds_broken()


def _assert_same(x, y, label):
    gx, cx, bx = x
    gy, cy, by = y
    assert len(gx) == len(gy), label
    for k in range(len(gx)):
        for f in ("id", "gid", "V", "m", "r", "r0", "I", "k", "v", "w", "q", "a", "dw"):
            assert np.array_equal(np.ascontiguousarray(gx[k][f]).view(np.uint8),
                                  np.ascontiguousarray(gy[k][f]).view(np.uint8)), (label, k, f)
        (c1, b1), (c2, b2) = cx[k], cy[k]
        assert np.array_equal(pair_keys(c1), pair_keys(c2)), (label, k)
        assert np.array_equal(b1, b2), (label, k)
        for f in ("F", "s", "ti", "tj", "k"):
            assert np.array_equal(_bits(c1[f]), _bits(c2[f])), (label, k, f)
    assert np.array_equal(bx, by), label


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_repack_equals_create_and_adopt(precision):
    """dem_batch_repack (kept scenes moved device to device, new ones uploaded) gives the batch
    that dem_create_batch + dem_batch_adopt gives, bit for bit, before and after more steps."""
    Batch, _ = _engine()
    old = _members(4)            # 4 cubes, a packed sample, a reference case
    extra = scenes.batched_cubes(2, seed=31, nmin=300, nmax=1500, base_gap=2e-5)
    print(f"[case] repack {precision}: {len(old)} scenes, keep 3, add 2", flush=True)
    with Batch(old, precision=precision) as src:
        src.set_bond_init([1, 1, 1, 1, 0, 1])
        src.initialize()
        src.step(23)
        keep = [5, 1, 3]
        with Batch.repack(src, keep, extra, precision=precision) as a, \
                Batch([old[k] for k in keep] + extra, precision=precision) as b:
            for j, k in enumerate(keep):
                b._check(b._lib.dem_batch_adopt(b._h, j, src._h, k))
            mask = [0, 0, 0, 1, 1]
            a.set_bond_init(mask)
            b.set_bond_init(mask)
            assert np.array_equal(a.offsets, b.offsets)
            _assert_same(_state(a), _state(b), "after repack")
            a.initialize()
            b.initialize()
            a.step(31)
            b.step(31)
            _assert_same(_state(a), _state(b), "after stepping")
        # src is untouched by the repack
        src.step(3)


def test_repack_keeps_only_and_from_single_scene():
    """No new scenes; and a single scene (not a batch) as the source."""
    Batch, Scene = _engine()
    m = _members(3)
    print("[case] repack without arrivals; repack of a single scene", flush=True)
    with Batch(m) as src:
        src.set_bond_init([1, 1, 1, 0, 1])
        src.initialize()
        src.step(11)
        before = _state(src)
        with Batch.repack(src, [4, 0]) as a:
            g, c, br = _state(a)
            _assert_same((g, c, br), ([before[0][4], before[0][0]], [before[1][4], before[1][0]],
                                      before[2][[4, 0]]), "kept only")
            a.step(5)
            src.step(5)
            after = _state(src)
            _assert_same(_state(a), ([after[0][4], after[0][0]], [after[1][4], after[1][0]],
                                     after[2][[4, 0]]), "kept only, stepped")
    cfg, g, w = m[0]
    with Scene(cfg, g, w) as s:
        s.initialize()
        s.step(9)
        with Batch.repack(s, [0], [m[1]]) as a, Batch([m[0], m[1]]) as b:
            b._check(b._lib.dem_batch_adopt(b._h, 0, s._h, 0))
            a.set_bond_init([0, 1])
            b.set_bond_init([0, 1])
            a.initialize()
            b.initialize()
            a.step(7)
            b.step(7)
            _assert_same(_state(a), _state(b), "single-scene source")


def test_repack_rejects_bad_arguments():
    Batch, _ = _engine()
    from taichi_dem_julia_port_amd.executor import DEMError
    with Batch(_members(2)) as src:
        for keep in ([0, 0], [7], [-1]):
            with pytest.raises(DEMError) as ei:
                Batch.repack(src, keep)
            assert ei.value.code == -1, keep
        with pytest.raises(DEMError):
            Batch.repack(src, [])   # no scenes at all


def test_memory_cache_reuse_gives_same_results():
    """With dem_set_memory_cache on, a batch built in a destroyed batch's memory (stale contents,
    cleared on reuse) steps exactly as one built in fresh memory."""
    Batch, _ = _engine()
    from taichi_dem_julia_port_amd.executor import set_memory_cache
    m = _members(3)
    small = m[:2] + m[3:4]

    def run(members):
        with Batch(members) as b:
            b.set_bond_init([1] * (len(members) - 1) + [0])
            b.initialize()
            b.step(17)
            return _state(b)

    print("[case] memory cache off / on", flush=True)
    fresh = run(small)
    set_memory_cache(8 << 30)
    try:
        run(m)                # leaves its (larger) blocks in the cache
        reused = run(small)   # built in them
        again = run(small)
    finally:
        set_memory_cache(0)
    _assert_same(reused, fresh, "cached blocks")
    _assert_same(again, fresh, "cached blocks, second reuse")


def test_repack_reports_rows_too_narrow_for_kept_scenes():
    """Explicit contact_slots below what a kept scene's rows hold is a capacity error, not a
    silent truncation; the source batch stays usable."""
    Batch, _ = _engine()
    from taichi_dem_julia_port_amd.executor import DEMError
    m = _members(2)
    with Batch(m) as src:
        src.set_bond_init([1, 1, 0, 1])
        src.initialize()
        src.step(3)
        with pytest.raises(DEMError) as ei:
            Batch.repack(src, [0, 1], contact_slots=2)
        assert ei.value.code == -3, ei.value
        src.step(2)
