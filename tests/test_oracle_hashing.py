"""Pins for the oracle's cell hashing (utils.jl:1-16) and sizing (executor.jl:247-267).

Checks come from outside the oracle: the README's Morton known answer (README.md:23) with the
magic-number interleave of dem.py:758-777 restated in numpy, a bit-by-bit interleave on random
triples, exact-boundary cases where ceil differs from floor + 1, and the hash-table sizes of the
reference's own cfg/1-3 including the Int32 wrap-around of cfg/3.
"""
import ctypes
import math

import numpy as np
import pytest

from helpers import reference_case
from oracle import oracle as O


def morton3d32_magic(x, y, z):
    """dem.py:758-777 restated (10-bit coordinates, x in bit 0)."""
    def spread(v):
        v &= 0x3FF
        v = (v | v << 16) & 0x30000FF
        v = (v | v << 8) & 0x300F00F
        v = (v | v << 4) & 0x30C30C3
        v = (v | v << 2) & 0x9249249
        return v
    return spread(x) | spread(y) << 1 | spread(z) << 2


def interleave_bits(x, y, z, bits=21):
    r = 0
    for b in range(bits):
        r |= ((x >> b) & 1) << (3 * b)
        r |= ((y >> b) & 1) << (3 * b + 1)
        r |= ((z >> b) & 1) << (3 * b + 2)
    return r


def _i3(v):
    return (ctypes.c_int32 * 3)(*v)


def _d3(v):
    return (ctypes.c_double * 3)(*v)


def test_morton_readme_known_answer():
    lib = O.load()
    # README.md:23 -- cartesian3morton([2,3,4]) == morton3d32(1, 2, 3) + 1
    assert lib.orc_cartesian3morton(_i3([2, 3, 4])) == morton3d32_magic(1, 2, 3) + 1
    assert morton3d32_magic(1, 2, 3) == 0b110101  # x=1,y=2,z=3 -> bits x0 y1 z2 ... by hand


def test_morton_matches_magic_numbers_on_10_bit_triples():
    lib = O.load()
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1024, size=(2000, 3)):
        assert lib.orc_morton3(int(x), int(y), int(z)) == morton3d32_magic(int(x), int(y), int(z))


def test_morton_matches_bitwise_interleave_on_21_bit_triples():
    lib = O.load()
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1 << 21, size=(500, 3)):
        assert lib.orc_morton3(int(x), int(y), int(z)) == interleave_bits(int(x), int(y), int(z))


def test_cell_uses_ceil_at_exact_boundaries():
    lib = O.load()
    out = (ctypes.c_int32 * 3)()
    # (pos - dmin) / cs is an exact integer: ceil gives k, floor + 1 would give k + 1
    lib.orc_cell(_d3([1.0, 2.5, -0.5]), _d3([0.0, 0.0, -1.0]), 0.5, out)
    assert list(out) == [2, 5, 1]
    lib.orc_cell(_d3([1.0000001, 2.4999999, -0.5]), _d3([0.0, 0.0, -1.0]), 0.5, out)
    assert list(out) == [3, 5, 1]
    # a particle exactly at domain_min lands in cell 0
    lib.orc_cell(_d3([0.0, 0.0, -1.0]), _d3([0.0, 0.0, -1.0]), 0.5, out)
    assert list(out) == [0, 0, 0]


def test_center_and_neighbors_closed_form():
    lib = O.load()
    rng = np.random.default_rng(2)
    dmin = np.array([-0.3, 0.1, -2.0])
    cs = 0.0371
    H = (1 << 20) + 12345  # not a power of two
    for _ in range(300):
        ijk = rng.integers(1, 600, size=3).astype(np.int32)
        c = (ctypes.c_double * 3)()
        lib.orc_center(_i3(ijk), _d3(dmin), cs, c)
        expect = dmin + (ijk.astype(np.float64) - 0.5) * cs
        assert np.array_equal(np.array(c[:]), expect)
        # position: random inside the cell, sometimes exactly the centre (d = -1 there)
        o = expect + rng.uniform(-0.5, 0.5, size=3) * cs
        if rng.random() < 0.2:
            o = expect.copy()
        nb = (ctypes.c_uint32 * 8)()
        lib.orc_neighbors(_d3(o), _i3(ijk), c, H, nb)
        d = np.where(o > expect, 1, -1)
        offs = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]
        for w, sel in enumerate(offs):
            cc = ijk + np.array(sel) * d
            m = interleave_bits(int(cc[0]) - 1, int(cc[1]) - 1, int(cc[2]) - 1) + 1
            assert nb[w] == m % H + 1


@pytest.mark.parametrize("k,cs,prod,H,bonds", [
    (1, 0.139173, 77760, 1 << 20, 30),
    (2, 0.00445352, 6935625, 1 << 22, 3009),
    (3, 0.012069, -1991335464, 1 << 20, 61898),
])
def test_reference_sizing_and_bond_counts(k, cs, prod, H, bonds):
    cfg, rows, walls = reference_case(k)
    d = O.derive(cfg, rows)
    assert abs(d["cell_size"] - cs) <= 5e-6 * cs  # survey values carry 6 significant digits
    assert d["grid_product_i32"] == prod
    assert d["hash_table_size"] == H
    o = O.OracleScene(cfg, rows, walls)
    o.initialize()
    assert o.counters()["bonded_contacts"] == bonds


def test_int32_wrap_rule_matches_explicit_arithmetic():
    cfg, rows, _ = reference_case(3)
    d = O.derive(cfg, rows)
    dims = [math.ceil((cfg.domain_max[a] - cfg.domain_min[a]) / d["cell_size"]) for a in range(3)]
    p = 1
    for v in dims:
        p = (p * v) & 0xFFFFFFFF
    p = p - (1 << 32) if p >= (1 << 31) else p
    assert p == d["grid_product_i32"]
