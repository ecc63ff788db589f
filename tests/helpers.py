"""Shared helpers for the parity tests: scene builders and comparison utilities."""
import os

import numpy as np

from taichi_dem_julia_port_amd import from_toml, scenes
from taichi_dem_julia_port_amd.p4io import read_p4p_rows

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def reference_case(k: int):
    """cfg/k.toml with the reference's own particle file (k in 1..3)."""
    files = {1: "carom.p4p", 2: "cube_911_particles_impact.p4p",
             3: "cube_18112_particles_impact.p4p.gz"}
    cfg = from_toml(os.path.join(GOLDEN, f"cfg{k}.toml"))
    rows = read_p4p_rows(os.path.join(GOLDEN, files[k]))
    walls = [(cfg.wall_normal, cfg.wall_distance, 2)]
    return cfg, rows, walls


def pour_sample(n=4000, seed=7, box=(0.02, 0.02, 0.03), settle_gap=None):
    cfg, g, w = scenes.pour(n, seed=seed, box=box)
    return cfg, g, w


def pair_keys(contacts):
    return contacts["i"].astype(np.int64) * (1 << 32) + contacts["j"].astype(np.int64)


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(den > 0, np.abs(a - b) / den, 0.0)
    return float(np.max(r)) if r.size else 0.0


def vec_rel_err(a, b):
    """max over rows of |a-b| / max(|a|, |b|) using row norms (scale-aware)."""
    if len(a) == 0 and len(b) == 0:
        return 0.0
    a = np.asarray(a, dtype=np.float64).reshape(len(a), -1)
    b = np.asarray(b, dtype=np.float64).reshape(len(b), -1)
    num = np.linalg.norm(a - b, axis=1)
    den = np.maximum(np.linalg.norm(a, axis=1), np.linalg.norm(b, axis=1))
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(den > 0, num / den, 0.0)
    return float(r.max())


def packed_sample(n=4000, scale=1.25, seed=8, box=(0.02, 0.02, 0.03), vmax=0.05):
    """A pour sample with radii scaled up after placement (so many pairs and walls overlap) and
    random velocities: a Hertz-Mindlin heavy state."""
    cfg, g, w = scenes.pour(n, seed=seed, box=box)
    g = g.copy()
    g["V"] *= scale ** 3
    g["m"] *= scale ** 3
    rng = np.random.default_rng(seed)
    g["v"] = rng.uniform(-vmax, vmax, size=(n, 3))
    return cfg, g, w


def compare_state(s, o, rel=1e-12, forces=True):
    """Engine scene s against oracle scene o after the same step(s) from the same state, at the
    FP64 bar: equal contact and bonded sets; per-contact k, F, ti, tj, s and grain k, v, q within
    rel; omega within rel of its scale; per-particle force and moment sums, a = F / m and
    d omega within rel of the sums of |contribution| (the oracle must track them:
    o.track_force_scale() before the step); wall shear within 1e-10 of its scale."""
    c, b = s.read_contacts()
    co, bo = o.read_contacts()
    assert np.array_equal(pair_keys(c), pair_keys(co)), "contact set differs"
    assert np.array_equal(b, bo), "bonded flags differ"
    for f in ("k", "F", "ti", "tj", "s"):
        assert vec_rel_err(c[f], co[f]) <= rel, f"contact field {f}"
    gs, go = s.read_grains(), o.read_grains()
    for f in ("k", "v", "q"):
        assert vec_rel_err(gs[f], go[f]) <= rel, f"grain field {f}"
    # sums of contributions may cancel: compare angular velocity against a scale
    scale = max(np.abs(go["w"]).max(), 1e-300)
    assert np.abs(gs["w"] - go["w"]).max() <= 1e-10 * scale + 1e-300, "grain field w"
    if forces:
        F, T = s.read_forces()
        Fo, To = o.read_forces()
        Fs, Ts = o.read_force_scale()
        assert getattr(o, "tracking_force_scale", False), "the oracle does not track the force scale"
        tiny = 1e-300
        assert np.all(np.abs(F - Fo) <= rel * Fs + tiny), "per-particle force sums"
        assert np.all(np.abs(T - To) <= rel * Ts + tiny), "per-particle moment sums"
        m = go["m"][:, None]
        assert np.all(np.abs(gs["a"] - go["a"]) <= rel * Fs / m * (1 + 1e-12) + tiny), "a"
        # d omega = R^-1 I^-1 (R tau - ...): rotations mix the components of tau's error
        I = (0.4 * go["m"] * go["r"] ** 2)[:, None]
        bound = rel * 2.0 * np.linalg.norm(Ts, axis=1)[:, None] / I
        assert np.all(np.abs(gs["dw"] - go["dw"]) <= bound + 1e-13 * np.abs(go["dw"]) + tiny), "dw"
    ws, wo = s.read_wall_shear(), o.read_wall_shear()
    assert np.abs(ws - wo).max(initial=0.0) <= 1e-10 * max(np.abs(wo).max(initial=0.0), 1e-30)


def reseed_oracle(s, cfg, g, w, parallel=False):
    """A fresh oracle holding the engine's current state (grains, contacts, wall shear), tracking
    its force scale."""
    from oracle.oracle import OracleScene
    o = OracleScene(cfg, g, w, parallel=parallel)
    o.write_grains(s.read_grains())
    c, b = s.read_contacts()
    o.write_contacts(c, b)
    o.write_wall_shear(s.read_wall_shear())
    o.track_force_scale()
    return o
