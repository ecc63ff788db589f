"""Near lists (a Verlet skin) change no result.

A scene keeps, at each step that sorts and runs the broad phase, every particle's partners within
r_i + r_j + 2 delta, and later steps test only those lists until some particle has moved delta;
the same scene created with near_lists=False sorts and runs the broad phase on every step.  After
every call both are bit for bit in the same state: grain k, v, a, q, omega, d omega, the contact
and bonded sets with every contact field, the per-particle force and moment sums and the step
counters.  The cases cover graph replays and eager steps, FP64 and mixed precision, sparse and
contact-heavy states, a bonded impact, a batch, a host write between calls, the diagnostic passes, near lists that overflow and
are widened, and grains fast enough that lean steps (the merge and force kernel alone) keep
finding they must rebuild and are replayed.  Each case moves far enough that the lists are
rebuilt during the run.
"""
import math

import numpy as np
import pytest

from helpers import packed_sample, pair_keys
from taichi_dem_julia_port_amd import scenes
from taichi_dem_julia_port_amd.config import DEMConfig
from taichi_dem_julia_port_amd.types import grains_in

pytestmark = pytest.mark.gpu

GRAIN_FIELDS = ("k", "v", "a", "q", "w", "dw")
CONTACT_FIELDS = ("i", "j", "k", "F", "ti", "tj", "s")
SCHEDULE = (1, 2, 7, 40, 150)


def _Scene():
    from taichi_dem_julia_port_amd.executor import Scene
    return Scene


def bits(x):
    """The raw bits of an array (so NaN grains compare too)."""
    x = np.ascontiguousarray(x)
    return x.view(np.uint64) if x.dtype == np.float64 else x


def same_state(a, b):
    ca, ba = a.read_contacts()
    cb, bb = b.read_contacts()
    assert np.array_equal(pair_keys(ca), pair_keys(cb)), "contact set differs"
    assert np.array_equal(ba, bb), "bonded flags differ"
    for f in CONTACT_FIELDS:
        assert np.array_equal(bits(ca[f]), bits(cb[f])), f"contact field {f}"
    ga, gb = a.read_grains(), b.read_grains()
    for f in GRAIN_FIELDS:
        assert np.array_equal(bits(ga[f]), bits(gb[f])), f"grain field {f}"
    Fa, Ta = a.read_forces()
    Fb, Tb = b.read_forces()
    assert np.array_equal(bits(Fa), bits(Fb)) and np.array_equal(bits(Ta), bits(Tb)), "force sums differ"
    na, nb = a.counters(), b.counters()
    for k in na:
        # near lists need cells 2 delta wider, and their broad phase examines more candidates
        if k not in ("rebuilds", "grid_cells", "examined", "replays"):
            assert na[k] == nb[k], k
    assert nb["rebuilds"] == nb["step"]
    return na


def shaken(n=3000, vmax=2.0):
    """The contact-heavy packing with speeds up to vmax: grains cross the skin in ~100 steps."""
    return packed_sample(n, vmax=vmax)


def fast(n=3000):
    """Grains at up to 30 m/s: they cross the skin in about ten steps, so lean steps keep finding
    that they must rebuild, are replayed as full steps, and the scene falls back to full steps."""
    return packed_sample(n, vmax=30.0)


def sparse(n=3000):
    """The pour sample at its own radii (few contacts) with speeds up to 2 m/s: contacts form and
    end during lean steps, while the lean merges leave out the groups of 64 grains with no row
    and no near-list partner within twice the displacement bound."""
    return packed_sample(n, scale=1.0, vmax=2.0)


CASES = {
    "fast_fp64": dict(build=fast, precision="fp64", graph=True, bonds=False, replays=True),
    "sparse_fp64": dict(build=sparse, precision="fp64", graph=True, bonds=False),
    "sparse_mixed_eager": dict(build=sparse, precision="mixed", graph=False, bonds=False),
    "packed_fp64_graph": dict(build=shaken, precision="fp64", graph=True, bonds=False),
    "packed_fp64_eager": dict(build=shaken, precision="fp64", graph=False, bonds=False),
    "packed_mixed": dict(build=shaken, precision="mixed", graph=True, bonds=False),
    "cube900_impact": dict(build=scenes.cube_small, precision="fp64", graph=True, bonds=True),
    "cube900_mixed": dict(build=scenes.cube_small, precision="mixed", graph=True, bonds=True),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_near_lists_change_no_result(name):
    spec = CASES[name]
    cfg, g, w = spec["build"]()
    Scene = _Scene()
    print(f"[case] {name} n={len(g)}", flush=True)
    kw = dict(precision=spec["precision"], graph=spec["graph"])
    with Scene(cfg, g, w, **kw) as near, Scene(cfg, g, w, near_lists=False, **kw) as full:
        if spec["bonds"]:
            near.initialize()
            full.initialize()
        for k in SCHEDULE:
            near.step(k)
            full.step(k)
            c = same_state(near, full)
            print(f"[{name}] step {c['step']}: {c['rebuilds']} rebuilds ({c['replays']} found by "
                  f"a lean step), {c['contacts']} contacts", flush=True)
        assert 1 < c["rebuilds"] <= c["step"], c
        if spec.get("replays"):
            assert c["replays"] > 0, c


def test_batch_near_lists_change_no_result():
    from taichi_dem_julia_port_amd.executor import Batch
    members = [scenes.bonded_cube(6, 6, 5, speed=8.0, seed=11),
               scenes.bonded_cube(5, 5, 7, speed=15.0, seed=12, bond_tensile_strength=2e6,
                                  bond_shear_strength=2e6),
               scenes.bonded_cube(4, 4, 4, speed=3.0, seed=13)]
    print("[case] batch of three bonded cubes, mixed precision", flush=True)
    with Batch(members, precision="mixed") as near, \
            Batch(members, precision="mixed", near_lists=False) as full:
        near.initialize()
        full.initialize()
        for k in SCHEDULE:
            near.step(k)
            full.step(k)
            c = same_state(near, full)
            assert np.array_equal(near.bonds_broken(), full.bonds_broken())
        assert 1 < c["rebuilds"] < c["step"], c


def test_host_write_and_diagnostics_with_near_lists():
    """A grain moved by the host into another's contact zone, between calls whose lists would
    otherwise still be valid; then calls with the diagnostic passes on."""
    cfg, g, w = packed_sample(2000)
    Scene = _Scene()
    print("[case] host write + diagnostics, 2000-grain packing", flush=True)
    with Scene(cfg, g, w) as near, Scene(cfg, g, w, near_lists=False) as full:
        for s in (near, full):
            s.step(5)
        same_state(near, full)
        for s in (near, full):
            gr = s.read_grains()
            # grain 0 onto grain 1 (overlapping by a tenth of their radii)
            d = gr["k"][1] - gr["k"][0]
            r0, r1 = gr["r"][0], gr["r"][1]
            gr["k"][0] = gr["k"][1] - d / np.linalg.norm(d) * 0.9 * (r0 + r1)
            s.write_grains(gr)
        for s in (near, full):
            s.step(3)
        c = same_state(near, full)
        ca, _ = near.read_contacts()
        assert np.any((ca["i"] == 1) & (ca["j"] == 2)), "the moved grain must touch grain 2"
        for s in (near, full):
            s.set_diagnostics(True)
            s.step(4)
        c = same_state(near, full)
        assert c["candidate_pairs"] > 0 and c["examined"] > 0


def _shell_scene(m=40, gap=2e-5, speed=0.5):
    """One 1 mm grain in a shell of m 0.25 mm grains a gap away, moving inward at `speed`: the
    big grain has m partners within its near reach, more than the lists hold at first."""
    rb, rs = 1.0e-3, 0.25e-3
    R = rb + rs + gap
    idx = np.arange(m) + 0.5
    phi = np.arccos(1.0 - 2.0 * idx / m)
    theta = math.pi * (1.0 + 5 ** 0.5) * idx
    u = np.stack([np.cos(theta) * np.sin(phi), np.sin(theta) * np.sin(phi), np.cos(phi)], axis=1)
    c0 = np.array([0.01, 0.01, 0.01])
    k = np.vstack([c0, c0 + R * u])
    r = np.concatenate([[rb], np.full(m, rs)])
    V = 4.0 / 3.0 * math.pi * r ** 3
    v = np.vstack([np.zeros((1, 3)), -speed * u])
    g = grains_in(np.arange(1, m + 2), V, scenes.DENSITY * V, k, v)
    cfg = DEMConfig(domain_min=[0.0, 0.0, 0.0], domain_max=[0.02, 0.02, 0.02],
                    particle_elastic_modulus=1e7, pp_coefficient_friction=0.4,
                    gravity=[0.0, 0.0, 0.0], dt=2e-7, max_coordinate_number=16,
                    wall_normal=[0.0, 0.0, -1.0], wall_distance=0.0)
    return cfg, g, None


def test_near_list_overflow_widens_the_lists():
    cfg, g, w = _shell_scene()
    Scene = _Scene()
    print("[case] shell of 40 small grains around a big one", flush=True)
    with Scene(cfg, g, w) as near, Scene(cfg, g, w, near_lists=False) as full:
        near.step(1)
        full.step(1)
        same_state(near, full)
        near.step(300)
        full.step(300)
        c = same_state(near, full)
        ca, _ = near.read_contacts()
        assert np.count_nonzero(ca["i"] == 1) >= 30, "the shell must close in on the big grain"
        # the first step's lists overflowed; after the widening they hold the shell
        assert c["rebuilds"] < 30, c


def test_nan_grains_do_not_force_rebuilds():
    """Grains dropped straight onto the floor: an exactly normal impact makes the reference's
    NaN (its slip ratio divides by a zero tangential force).  A NaN grain touches nothing, so it
    must not make every step rebuild; the state, NaN grains included, is the full path's bit for
    bit."""
    n = 400
    r = 0.5e-3
    side = 20
    ii, jj = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    k = np.stack([(ii.ravel() + 0.5) * 3 * r, (jj.ravel() + 0.5) * 3 * r,
                  np.full(n, r + 2e-6)], axis=1)
    V = np.full(n, 4.0 / 3.0 * math.pi * r ** 3)
    v = np.zeros((n, 3))
    v[:, 2] = -0.5
    g = grains_in(np.arange(1, n + 1), V, scenes.DENSITY * V, k, v)
    cfg = DEMConfig(domain_min=[-0.01, -0.01, -0.01], domain_max=[0.04, 0.04, 0.02],
                    particle_elastic_modulus=1e7, pp_coefficient_friction=0.4,
                    gravity=[0.0, 0.0, -9.81], dt=2e-7, max_coordinate_number=16,
                    wall_normal=[0.0, 0.0, -1.0], wall_distance=0.0)
    Scene = _Scene()
    print("[case] 400 grains dropped straight onto the floor", flush=True)
    with Scene(cfg, g, None) as near, Scene(cfg, g, None, near_lists=False) as full:
        for s in (near, full):
            s.step(3000)  # an impact lasts about 1000 steps
        c = same_state(near, full)
        nan = ~np.isfinite(near.read_grains()["k"]).all(axis=1)
        print(f"[nan] {int(nan.sum())} grains not finite, {c['rebuilds']} rebuilds in {c['step']} "
              f"steps", flush=True)
        assert nan.any(), "no grain went non-finite: the case does not test what it names"
        assert c["rebuilds"] < 20, c
