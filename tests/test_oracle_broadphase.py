"""Pins for the oracle's counting sort (kernels.jl:7-22) and broad phase (kernels.jl:24-58).

The broad phase is checked against an O(N^2) brute force that applies the window rule directly:
(i, j), i < j, appears once for every one of i's 8 window buckets equal to j's bucket -- so
aliased window buckets yield duplicates, which a scene with a non-power-of-two hash table is
built to provoke.
"""
import numpy as np

from helpers import pour_sample, reference_case
from oracle import oracle as O
from taichi_dem_julia_port_amd import DEMConfig
from taichi_dem_julia_port_amd.types import grains_in


def brute_cp_list(hid, nbr):
    n = len(hid)
    pairs = []
    for i in range(n):
        for w in range(8):
            js = np.nonzero(hid == nbr[i, w])[0]
            for j in js:
                if i < j:
                    pairs.append((i + 1, j + 1))
    return np.array(sorted(pairs), dtype=np.int64).reshape(-1, 2)


def _sorted_pairs(a):
    a = np.asarray(a, dtype=np.int64)
    return a[np.lexsort((a[:, 1], a[:, 0]))]


def _check_scene(cfg, g, w):
    o = O.OracleScene(cfg, g, w)
    o.contact_pass()
    kid, hid, kc, nbr = o.cells()
    cnt, cur, pid = o.hash_table()
    # counting sort invariants
    assert np.array_equal(np.sort(pid), np.arange(1, len(g) + 1))
    cnt = cnt.astype(np.int64)
    assert np.array_equal(cnt, np.bincount(hid.astype(np.int64) - 1, minlength=len(cnt)))
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
    assert np.array_equal(cur, start)  # atomic_sub leaves the exclusive start
    for b in np.nonzero(cnt)[0][:500]:
        seg = pid[start[b]:start[b] + cnt[b]]
        assert np.all(hid[seg - 1] == b + 1)
    # broad phase multiset
    got = _sorted_pairs(o.cp_list())
    ref = brute_cp_list(hid, nbr)
    assert np.array_equal(got, ref)
    return o, got


def test_broad_phase_reference_cube():
    cfg, g, w = reference_case(2)
    _check_scene(cfg, g, w)


def test_broad_phase_pour_sample():
    cfg, g, w = pour_sample(1500)
    _check_scene(cfg, g, w)


def _aliasing_scene():
    """A hash table whose size is not a power of two, built so that the buckets of x-cells 128 and
    129 coincide: Morton(128 - 1) and Morton(129 - 1) differ by spread(128) - spread(127) =
    1,797,559, which lies inside the clamp window [2^20, 2^22] and becomes H itself."""
    from test_oracle_hashing import interleave_bits
    H = interleave_bits(128, 0, 0) - interleave_bits(127, 0, 0)
    assert H == 1797559 and (1 << 20) < H < (1 << 22)
    r = 0.01
    cs = 4 * r * 1.1 * (1 + 6e7 / 28e9) * 1.01
    return r, cs, (H, 1, 1), H, (128, 1)


def test_broad_phase_aliased_window_duplicates():
    r, cs, dims, H, hit = _aliasing_scene()
    cx, cy = hit
    dmin = np.array([0.0, 0.0, 0.0])
    cfg = DEMConfig(domain_min=list(dmin), domain_max=[(dims[0] - 0.5) * cs,
                                                       (dims[1] - 0.5) * cs, 0.5 * cs],
                    wall_distance=-10.0, gravity=[0.0, 0.0, 0.0])
    # particle A right of its cell centre (window includes cx+1); particle B in cell cx+1
    ka = np.array([(cx - 0.3) * cs, (cy - 0.5) * cs, 0.5 * cs])
    kb = np.array([(cx + 0.2) * cs, (cy - 0.5) * cs, 0.5 * cs])
    V = 4 / 3 * np.pi * r ** 3
    g = grains_in([1, 2], [V, V], [2500 * V] * 2, np.stack([ka, kb]), np.zeros((2, 3)))
    o, pairs = _check_scene(cfg, g, [((0.0, 0.0, -1.0), -10.0, 2)])
    assert O.derive(cfg, g)["hash_table_size"] == H
    # A's own bucket and its +x bucket coincide: B is listed in both -> duplicate pair
    assert len(pairs) >= 2 and (pairs == [1, 2]).all(axis=1).sum() >= 2
