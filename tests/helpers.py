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
