"""GPU parity of loaded bonds, fracture, damping, the dropped-candidate rule and long runs.

* Loaded EBPM through fracture (kernels.jl:170-239): bonded cubes whose grains get seeded random
  velocity and spin perturbations, so every bond carries load from the first step, fall onto
  the floor and fracture.  At several points the oracle is re-seeded from the engine's state and
  both advance two steps; after each step the contact and bonded sets are equal, bond F, ti, tj
  and Hertz springs agree within 1e-12 relative, grain k, v, q, omega, a, d omega and the
  per-particle force and moment sums agree at the FP64 bar (sums against the sums of
  |contribution|), and the number of bonds broken in the window is equal.  The test also requires
  that the compared windows saw bonds break and broken pairs that still touch come back as
  Hertz contacts (kernels.jl:97-119 after a break), so the branches are exercised.
* Global damping (kernels.jl:435-441) on a packed pour and a loaded lattice.
* The dropped-candidate rule (kernels.jl:97-119, 241-245): a Hertz contact whose partner leaves
  the 8-bucket window stays active, is evaluated with gap > 0, and the fatal-gap count agrees.
* Long runs (10,000 steps through impact): kinetic energy, linear and angular momentum and the
  bond-break count of the engine against the serial oracle (deterministic, as the engine is);
  the dynamics are chaotic after fracture, so the tolerance at each checkpoint is 10x the largest
  divergence between the engine and the engine started from x, y or z coordinates moved by one
  ulp or x, y by 1e-12 relative, with a floor.  Mixed precision against FP64 on
  the same checkpoints, with the divergence of an FP64 run moved by 1e-7 relative (about one
  FP32 rounding) as its yardstick.
"""
import numpy as np
import pytest

from helpers import compare_state, pair_keys, packed_sample, reference_case, reseed_oracle
from oracle.oracle import OracleScene
from taichi_dem_julia_port_amd import scenes

pytestmark = pytest.mark.gpu


def _Scene():
    from taichi_dem_julia_port_amd.executor import Scene
    return Scene


def _perturb(s, seed, vstd, wstd):
    g = s.read_grains()
    rng = np.random.default_rng(seed)
    g["v"] = g["v"] + rng.normal(0.0, vstd, size=(len(g), 3))
    g["w"] = rng.normal(0.0, wstd, size=(len(g), 3))
    s.write_grains(g)


FRACTURE = {
    # config A's lattice with 3 MPa bonds: bonds break gradually under the perturbation and at
    # the floor impact (about step 100)
    "cube900": dict(build=lambda: scenes.cube_small(bond_tensile_strength=3e6,
                                                    bond_shear_strength=3e6,
                                                    bond_compressive_strength=2e7),
                    vstd=0.2, wstd=20.0, sched=[3, 10, 20, 30, 30, 40, 60, 100], parallel=False),
    # config B: 26x26x27, 1 MPa bonds, 20 m/s
    "cube18k": dict(build=scenes.cube_18k, vstd=0.1, wstd=10.0, sched=[3, 15, 30, 40, 20, 40],
                    parallel=True),
}


@pytest.mark.parametrize("name", sorted(FRACTURE))
def test_loaded_bonds_stepwise_through_fracture(name):
    spec = FRACTURE[name]
    cfg, g, w = spec["build"]()
    print(f"[case] {name} n={len(g)}", flush=True)
    broken_seen = recreated_seen = 0
    with _Scene()(cfg, g, w) as s:
        s.initialize()
        _perturb(s, 5, spec["vstd"], spec["wstd"])
        for k in spec["sched"]:
            s.step(k)
            o = reseed_oracle(s, cfg, g, w, parallel=spec["parallel"])
            c, b = s.read_contacts()
            bonded_before = set(pair_keys(c)[b].tolist())
            nb0 = s.counters()["bonds_broken_total"]
            for _ in range(2):
                s.step(1)
                o.step(1)
                compare_state(s, o)
            broken = s.counters()["bonds_broken_total"] - nb0
            assert broken == o.counters()["bonds_broken_total"], "bonds broken in the window"
            c, b = s.read_contacts()
            keys = pair_keys(c)
            recreated = int(sum(1 for key in keys[~b].tolist() if key in bonded_before))
            print(f"[{name}] window after {k} steps: {broken} bonds broke, {recreated} re-created "
                  f"as Hertz contacts, {int(b.sum())} bonds left", flush=True)
            broken_seen += broken
            recreated_seen += recreated
            o.close()
    assert broken_seen > 0 and recreated_seen > 0, "the windows did not exercise fracture"


DAMPED = {
    "packed_4k": lambda: packed_sample(4000),
    "lattice_loaded": lambda: scenes.bonded_cube(8, 8, 6, speed=5.0, base_gap=2e-5),
}


@pytest.mark.parametrize("name", sorted(DAMPED))
def test_global_damping_parity(name):
    cfg, g, w = DAMPED[name]()
    cfg = cfg.replace(global_damping=0.3)
    bonded = name.startswith("lattice")
    print(f"[case] {name} damping 0.3", flush=True)
    with _Scene()(cfg, g, w) as s:
        if bonded:
            s.initialize()
        _perturb(s, 9, 0.05, 10.0)
        for k in (1, 10, 30):
            s.step(k)
            o = reseed_oracle(s, cfg, g, w)
            s.step(1)
            o.step(1)
            compare_state(s, o)
            o.close()


def _far_point(cfg):
    """A point inside the domain, in a corner far from the grains of the test scenes."""
    lo, hi = np.array(cfg.domain_min), np.array(cfg.domain_max)
    return lo + 0.97 * (hi - lo)


@pytest.mark.parametrize("case", ["pair", "packed"])
def test_dropped_candidate_contact_stays_active(case):
    """Move one partner of an active Hertz contact out of the other's window: the contact stays
    active in both, is evaluated with gap > 0 in both, and the fatal-gap counts agree."""
    if case == "pair":
        from oracle_pins import two_spheres
        r = 0.01
        cfg, g = two_spheres(r, r, (0, 0, 0), (2 * r - 1e-5, 0, 0), dt=1e-7)
        w = None
    else:
        cfg, g, w = packed_sample(2000)
    print(f"[case] dropped candidate, {case}", flush=True)
    with _Scene()(cfg, g, w) as s:
        s.step(1)
        c, b = s.read_contacts()
        hertz = np.nonzero(~b)[0]
        assert len(hertz) > 0
        i, j = int(c["i"][hertz[0]]), int(c["j"][hertz[0]])
        gr = s.read_grains()
        gr["k"][j - 1] = _far_point(cfg)
        s.write_grains(gr)
        o = reseed_oracle(s, cfg, g, w)
        o.contact_pass()
        cp = o.cp_list()
        assert not np.any((cp[:, 0] == i) & (cp[:, 1] == j)), "the pair is still a candidate"
        s.step(1)
        o.step(1)
        compare_state(s, o)
        c, _ = s.read_contacts()
        assert np.any((c["i"] == i) & (c["j"] == j)), "the dropped pair must stay active"
        fatal = s.counters()["fatal_gap"]
        assert fatal >= 1 and fatal == o.counters()["fatal_gap"]
        o.close()


def _stats(gr, broken):
    m = gr["m"]
    I = 0.4 * m * gr["r"] ** 2
    c = (m[:, None] * gr["k"]).sum(0) / m.sum()
    d = gr["k"] - c
    return {
        "E": 0.5 * (m * (gr["v"] ** 2).sum(1)).sum() + 0.5 * (I * (gr["w"] ** 2).sum(1)).sum(),
        "P": (m[:, None] * gr["v"]).sum(0),
        "P_scale": (m * np.linalg.norm(gr["v"], axis=1)).sum(),
        "L": (np.cross(d, m[:, None] * gr["v"]) + I[:, None] * gr["w"]).sum(0),
        "L_scale": (m * np.linalg.norm(d, axis=1) * np.linalg.norm(gr["v"], axis=1)).sum()
        + (I * np.linalg.norm(gr["w"], axis=1)).sum(),
        "broken": broken,
    }


def _diff(a, b):
    return {"E": abs(a["E"] - b["E"]) / a["E"],
            "P": float(np.abs(a["P"] - b["P"]).max()) / a["P_scale"],
            "L": float(np.abs(a["L"] - b["L"]).max()) / a["L_scale"],
            "broken": abs(a["broken"] - b["broken"])}


LONG = {"cfg2_cube911": lambda: reference_case(2), "cube900": scenes.cube_small}


@pytest.mark.parametrize("name", sorted(LONG))
def test_long_run_statistics_match_the_oracle(name):
    Scene = _Scene()
    cfg, g, w = LONG[name]()
    checkpoints, every = 5, 2000
    print(f"[case] {name}: {checkpoints * every} steps, engine runs and the serial oracle",
          flush=True)

    def moved(rel, axis=0):
        gg = g.copy()
        x = gg["k"][:, axis]
        gg["k"][:, axis] = np.nextafter(x, np.inf) if rel == 0 else x * (1.0 + rel)
        return gg

    # the yardstick: how far the engine drifts from itself when every x (y, z) coordinate starts
    # one ulp, or 1e-12 relative, away -- the largest of these five.  The engine and the oracle
    # differ by an ulp or so in every step's sums (their summation orders differ), and which of
    # several such nudges a chaotic run amplifies most, and when, varies; a single control
    # understates the spread.  The serial oracle is deterministic, so the comparison is too (the
    # OpenMP build's atomic sums differ between runs by an ulp).
    controls = {"fp64_ulp": moved(0), "fp64_ulp_y": moved(0, 1), "fp64_ulp_z": moved(0, 2),
                "fp64_1e-12": moved(1e-12), "fp64_1e-12_y": moved(1e-12, 1)}
    runs = {"fp64": Scene(cfg, g, w), **{k: Scene(cfg, gg, w) for k, gg in controls.items()},
            "fp64_1e-7": Scene(cfg, moved(1e-7), w), "mixed": Scene(cfg, g, w, precision="mixed")}
    o = OracleScene(cfg, g, w)
    o.initialize()
    for s in runs.values():
        s.initialize()
    floor = {"E": 1e-10, "P": 1e-10, "L": 1e-10}
    floor_mixed = {"E": 1e-6, "P": 1e-6, "L": 1e-6}
    try:
        for t in range(1, checkpoints + 1):
            st = {}
            for k, s in runs.items():
                s.step(every)
                st[k] = _stats(s.read_grains(), s.counters()["bonds_broken_total"])
            o.step(every)
            so = _stats(o.read_grains(), o.counters()["bonds_broken_total"])
            ctl = [_diff(st["fp64"], st[k]) for k in controls]
            ref = {q: max(c[q] for c in ctl) for q in ctl[0]}
            got = _diff(st["fp64"], so)
            ref7 = _diff(st["fp64"], st["fp64_1e-7"])
            mix = _diff(st["fp64"], st["mixed"])
            print(f"[{name} step {t * every}] broken {st['fp64']['broken']} (oracle {so['broken']},"
                  f" mixed {st['mixed']['broken']}); engine vs oracle {got}; largest control "
                  f"{ref}; mixed vs fp64 {mix}; 1e-7 control {ref7}", flush=True)
            for q in ("E", "P", "L"):
                assert got[q] <= 10 * ref[q] + floor[q], (t, q, got, ref)
                assert mix[q] <= 10 * ref7[q] + floor_mixed[q], (t, q, mix, ref7)
            assert got["broken"] <= max(2, 10 * ref["broken"]), (t, got, ref)
            assert mix["broken"] <= max(5, 10 * ref7["broken"], 0.01 * st["fp64"]["broken"]), \
                (t, mix, ref7)
        if name == "cfg2_cube911":  # the run crosses the impact and fractures
            assert st["fp64"]["broken"] > 1000
    finally:
        for s in runs.values():
            s.close()
        o.close()
