"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle on identical inputs.

Bars: cell ids / buckets / sort / cp_list / contact and bond sets bit-exact (sets compared after
sorting, since atomic order differs); FP64 per-contact quantities within 1e-12 relative;
positions and velocities after one step within 1e-12 relative of the oracle's (the only
difference is the order in which contact forces are summed).
"""
import numpy as np
import pytest

from helpers import (compare_state, packed_sample, pair_keys, pour_sample, reference_case,
                     reseed_oracle)
from oracle.oracle import OracleScene
from taichi_dem_julia_port_amd import scenes

pytestmark = pytest.mark.gpu

FP64_REL = 1e-12


def _engine():
    from taichi_dem_julia_port_amd.executor import Scene, device_scan
    return Scene, device_scan


def _cases():
    return {
        "cfg1_carom": lambda: reference_case(1),
        "cfg2_cube911": lambda: reference_case(2),
        "cfg3_cube18112": lambda: reference_case(3),
        "lattice_10x10x9": lambda: scenes.cube_small(),
        "pour_4k": lambda: pour_sample(4000),
        "packed_4k": lambda: packed_sample(4000),
        # safety factor < 1: contact reach exceeds cs/2, so the reference window rule is checked
        # per touching pair and the reference-grid path is index-owned
        "lattice_safety09": lambda: scenes.bonded_cube(8, 8, 6, speed=10.0, dt=1e-6, base_gap=2e-5,
                                                       neighboring_search_safety_factor=0.9),
        # grains beyond domain_max in x and y: the dense grid clamps them into its last cells
        "packed_outside": lambda: _shrunk_domain(packed_sample(2000)),
        # a slab narrower than two cells in x: the dense grid has one or two cells on that axis
        "packed_slab": lambda: _slab(packed_sample(1500, box=(0.004, 0.03, 0.02))),
    }


def _slab(case):
    """The domain cut to the grains' x extent (4 mm, under two contact-reach cells)."""
    cfg, g, w = case
    lo, hi = list(cfg.domain_min), list(cfg.domain_max)
    r = np.cbrt(g["V"] / (4.0 / 3.0 * np.pi))
    lo[0] = float((g["k"][:, 0] - r).min())
    hi[0] = float((g["k"][:, 0] + r).max())
    return cfg.replace(domain_min=lo, domain_max=hi), g, w


def _shrunk_domain(case):
    """The same grains in a domain whose upper x and y bounds sit at the 70th percentile of the
    grain coordinates, so about half the grains lie outside it."""
    cfg, g, w = case
    hi = list(cfg.domain_max)
    hi[0] = float(np.percentile(g["k"][:, 0], 70))
    hi[1] = float(np.percentile(g["k"][:, 1], 70))
    return cfg.replace(domain_max=hi), g, w


# cases stepped without initialize! (solve's no_bond = true)
UNBONDED = {"pour_4k", "packed_4k", "cfg2_nobond", "packed_outside", "packed_slab"}


CASES = sorted(_cases())


def _build(name):
    print(f"[case] {name}", flush=True)
    return _cases()[name]()


@pytest.mark.parametrize("n", [1, 7, 4096, 4097, 65536 * 3 + 5, (1 << 22) + 1])
def test_scan_matches_cumsum(n):
    _, device_scan = _engine()
    rng = np.random.default_rng(n)
    a = rng.integers(0, 5, size=n, dtype=np.uint32)
    print(f"[scan] n={n}", flush=True)
    out, tot = device_scan(a)
    ref = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)[:-1]]).astype(np.uint32)
    assert np.array_equal(out, ref)
    assert tot == int(a.sum(dtype=np.uint64))


def test_scan_repeated_calls_reset_state():
    _, device_scan = _engine()
    for k in range(5):
        a = np.full(10000 + k, 1, dtype=np.uint32)
        out, tot = device_scan(a)
        assert tot == len(a) and out[-1] == len(a) - 1


@pytest.mark.parametrize("name", CASES)
def test_cells_bit_exact(name):
    Scene, _ = _engine()
    cfg, g, w = _build(name)
    o = OracleScene(cfg, g, w)
    o.contact_pass()
    kid_o, hid_o, kc_o, nbr_o = o.cells()
    with Scene(cfg, g, w) as s:
        kid, hid, kc, nbr = s.debug_cells()
    assert np.array_equal(kid, kid_o)
    assert np.array_equal(hid, hid_o)
    assert np.array_equal(kc.view(np.uint64), kc_o.view(np.uint64))
    assert np.array_equal(nbr, nbr_o)


@pytest.mark.parametrize("name", CASES)
def test_counting_sort(name):
    Scene, _ = _engine()
    cfg, g, w = _build(name)
    o = OracleScene(cfg, g, w)
    o.contact_pass()
    cnt_o, start_o, pid_o = o.hash_table()
    with Scene(cfg, g, w) as s:
        cnt, start, pid = s.debug_hash_table()
    assert np.array_equal(cnt, cnt_o)
    assert np.array_equal(start, start_o)
    # same particle set in every bucket (order inside a bucket is atomic-order dependent)
    nz = np.nonzero(cnt)[0]
    for b in nz:
        a0, c = start[b], cnt[b]
        assert sorted(pid[a0:a0 + c]) == sorted(pid_o[a0:a0 + c])
    assert np.array_equal(np.sort(pid), np.arange(1, len(g) + 1))


@pytest.mark.parametrize("name", CASES)
def test_cp_list_multiset(name):
    Scene, _ = _engine()
    cfg, g, w = _build(name)
    o = OracleScene(cfg, g, w)
    o.contact_pass()
    ref = o.cp_list()
    with Scene(cfg, g, w) as s:
        got = s.debug_cp_list()
    assert len(got) == len(ref)
    key = lambda a: np.sort(a[:, 0].astype(np.int64) * (1 << 32) + a[:, 1])
    assert np.array_equal(key(got), key(ref))


@pytest.mark.parametrize("variant", ["auto", "refgrid", "index"])
@pytest.mark.parametrize("name", ["cfg1_carom", "cfg2_cube911", "cfg3_cube18112", "lattice_10x10x9"])
def test_initialize_bond_set(name, variant):
    Scene, _ = _engine()
    cfg, g, w = _build(name)
    o = OracleScene(cfg, g, w)
    o.initialize()
    co, bo = o.read_contacts()
    with Scene(cfg, g, w, **VARIANTS[variant]) as s:
        s.initialize()
        c, b = s.read_contacts()
        cnt = s.counters()
    assert np.array_equal(pair_keys(c), pair_keys(co))
    assert np.array_equal(b, bo) and b.all()
    assert cnt["bonded"] == len(co)


def _compare_state(s, o, rel=FP64_REL):
    compare_state(s, o, rel)


ONE_STEP_CASES = CASES + ["cfg2_nobond"]


# broad-phase variants: the engine's choice (dense cs/2 grid where the domain is compact), the
# reference's hashed grid with slot-owned pairs, and index-owned pairs as in kernels.jl:31
VARIANTS = {"auto": {}, "refgrid": {"reference_grid": True},
            "index": {"index_owned_broad_phase": True}}


@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("name", ONE_STEP_CASES)
def test_one_step_parity(name, variant):
    Scene, _ = _engine()
    cfg, g, w = _build("cfg2_cube911" if name == "cfg2_nobond" else name)
    bonded = name not in UNBONDED
    o = OracleScene(cfg, g, w)
    o.track_force_scale()
    with Scene(cfg, g, w, **VARIANTS[variant]) as s:
        if bonded:
            o.initialize()
            s.initialize()
        o.step(1)
        s.step(1)
        _compare_state(s, o)


@pytest.mark.parametrize("variant", ["auto", "refgrid"])
@pytest.mark.parametrize("name", ["cfg2_cube911", "lattice_10x10x9", "packed_4k", "packed_outside",
                                  "packed_slab"])
def test_stepwise_parity_from_identical_states(name, variant):
    """Run 60 steps; every 20 steps copy the engine state into the oracle and compare the next step."""
    Scene, _ = _engine()
    cfg, g, w = _build(name)
    bonded = name not in UNBONDED
    with Scene(cfg, g, w, **VARIANTS[variant]) as s:
        if bonded:
            s.initialize()
        for _ in range(3):
            s.step(20)
            o = reseed_oracle(s, cfg, g, w)
            s.step(1)
            o.step(1)
            _compare_state(s, o)


@pytest.mark.parametrize("variant", ["auto", "refgrid"])
@pytest.mark.parametrize("name", ["cfg2_cube911", "packed_4k"])
def test_candidate_pair_counter(name, variant):
    """P reported by the engine equals the length of the reference's cp_list (executor.jl:63)."""
    Scene, _ = _engine()
    cfg, g, w = _build(name)
    o = OracleScene(cfg, g, w)
    with Scene(cfg, g, w, **VARIANTS[variant]) as s:
        assert s.counters()["candidate_pairs"] == -1  # diagnostics are off by default
        s.set_diagnostics(True)
        s.step(1)
        o.step(1)
        assert s.counters()["candidate_pairs"] == o.counters()["total_cps"]
        s.set_diagnostics(False)
        s.step(1)
        assert s.counters()["candidate_pairs"] == -1 and s.counters()["examined"] == -1


@pytest.mark.parametrize("name", ["lattice_10x10x9", "cfg2_cube911"])
def test_contact_restore_round_trip(name):
    """Read the contact set, write it back, read again: the same records; a bond's s restored as
    nonzero comes back as zero (every reference step stores zero there, kernels.jl:230-238), and
    the next step still matches the oracle."""
    Scene, _ = _engine()
    cfg, g, w = _build(name)
    with Scene(cfg, g, w) as s:
        s.initialize()
        s.step(5)
        c, b = s.read_contacts()
        assert b.any()
        assert np.all(c["s"][b] == 0.0)
        c2 = c.copy()
        c2["s"][b] = 1.0
        s.write_contacts(c2, b)
        c3, b3 = s.read_contacts()
        assert np.array_equal(b3, b)
        for f in ("i", "j", "k", "F", "ti", "tj", "s"):
            assert np.array_equal(c3[f], c[f]), f
        o = reseed_oracle(s, cfg, g, w)
        s.step(1)
        o.step(1)
        _compare_state(s, o)
