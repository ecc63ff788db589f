"""Synthetic workloads: the bonded-cube impacts and the polydisperse pour.

Each builder returns ``(cfg, grains_in, walls)`` ready for ``Scene(cfg, grains_in, walls)``.

* ``bonded_cube(nx, ny, nz, ...)`` -- simple-cubic lattice of spheres (spacing 2 r, radius r with
  +-1 % uniform jitter), launched downwards at a rigid floor z = 0.  With the default
  particle_contact_radius_multiplier 1.1 the t=0 bond rule ``|k_i - k_j| < r0_i + r0_j``
  (kernels.jl:72, executor.jl:235) bonds exactly the 6 lattice neighbours.
  ``cube_small()`` = 10x10x9, v = -10 m/s, dt = 1e-6 s;
  ``cube_18k()`` = 26x26x27, v = -20 m/s, dt = 5e-7 s, bond tensile/shear strength 1e6 Pa.
* ``pour(n)`` -- n spheres with log-normal radius (median 0.5 mm, sigma 0.25, clipped to
  0.25-1.0 mm), density 2500 kg/m^3, seeded at random non-overlapping positions in a
  0.1 x 0.1 x 0.3 m box with a floor and four side walls; Hertz-Mindlin only (E = 1e7 Pa,
  mu = 0.4), dt = 2e-7 s.
* ``pour_packed(n)`` -- the same grains with radii inflated 1.25x after placement: a
  contact-heavy Hertz-Mindlin state of the pour.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from .config import DEMConfig
from .types import GRAIN_IN_DTYPE, grains_in, make_walls

DENSITY = 2500.0


def bonded_cube(nx: int, ny: int, nz: int, r: float = 0.01, jitter: float = 0.01,
                speed: float = 10.0, dt: float = 1e-6, seed: int = 1, base_gap: float = 1e-3,
                **cfg_overrides):
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    spacing = 2.0 * r
    k = np.stack([(ii - (nx - 1) / 2.0) * spacing, (jj - (ny - 1) / 2.0) * spacing,
                  base_gap + r * (1.0 + jitter) + kk * spacing], axis=1)
    radii = r * (1.0 + rng.uniform(-jitter, jitter, size=n))
    V = 4.0 / 3.0 * math.pi * radii ** 3
    m = DENSITY * V
    v = np.zeros((n, 3))
    v[:, 2] = -speed
    g = grains_in(np.arange(1, n + 1), V, m, k, v)
    half = max(nx, ny) * spacing
    top = base_gap + (nz + 2) * spacing
    cfg = DEMConfig(
        domain_min=[-half - 0.05, -half - 0.05, -0.05],
        domain_max=[half + 0.05, half + 0.05, top + 0.05],
        particle_elastic_modulus=1e8, particle_poisson_ratio=0.3,
        pp_coefficient_friction=0.5, pp_coefficient_restitution=0.8,
        wall_normal=[0.0, 0.0, -1.0], wall_distance=0.0,
        gravity=[0.0, 0.0, -9.81], dt=dt, max_coordinate_number=16,
    )
    cfg = cfg.replace(**cfg_overrides)
    walls = make_walls([(cfg.wall_normal, cfg.wall_distance, 2)])
    return cfg, g, walls


def cube_small(**kw):
    """~900-particle bonded cube impact (10 x 10 x 9, v = -10 m/s, dt = 1e-6 s)."""
    return bonded_cube(10, 10, 9, speed=10.0, dt=1e-6, seed=1, **kw)


def cube_18k(**kw):
    """~18k-particle bonded cube impact (26 x 26 x 27, v = -20 m/s, dt = 5e-7 s)."""
    kw.setdefault("bond_tensile_strength", 1e6)
    kw.setdefault("bond_shear_strength", 1e6)
    return bonded_cube(26, 26, 27, speed=20.0, dt=5e-7, seed=2, **kw)


POUR_BOX = (0.1, 0.1, 0.3)


def pour_walls(box=POUR_BOX):
    bx, by, _ = box
    return make_walls([
        ((0.0, 0.0, -1.0), 0.0, 2),  # floor z = 0
        ((-1.0, 0.0, 0.0), 0.0, 2),  # x = 0
        ((1.0, 0.0, 0.0), bx, 2),    # x = bx
        ((0.0, -1.0, 0.0), 0.0, 2),  # y = 0
        ((0.0, 1.0, 0.0), by, 2),    # y = by
    ])


def pour_grains(n: int, box=POUR_BOX, seed: int = 3, median_r: float = 0.5e-3,
                sigma: float = 0.25, rmin: float = 0.25e-3, rmax: float = 1.0e-3) -> np.ndarray:
    from . import _native as N
    lib = N.load_library()
    out = np.zeros(n, dtype=GRAIN_IN_DTYPE)
    lo = (ctypes.c_double * 3)(0.0, 0.0, 0.0)
    hi = (ctypes.c_double * 3)(*box)
    rc = lib.dem_generate_pour(n, lo, hi, median_r, sigma, rmin, rmax, DENSITY, seed,
                               out.ctypes.data_as(ctypes.POINTER(N.DemGrainIn)))
    if rc != 0:
        raise RuntimeError(f"dem_generate_pour failed ({rc}): box too small for {n} grains?")
    return out


def pour_config(box=POUR_BOX, dt: float = 2e-7, **overrides) -> DEMConfig:
    bx, by, bz = box
    m = 0.005
    cfg = DEMConfig(
        domain_min=[-m, -m, -m], domain_max=[bx + m, by + m, bz + m],
        particle_elastic_modulus=1e7, pp_coefficient_friction=0.4, pw_coefficient_friction=0.4,
        wall_normal=[0.0, 0.0, -1.0], wall_distance=0.0,
        gravity=[0.0, 0.0, -9.81], dt=dt, max_coordinate_number=32,
    )
    return cfg.replace(**overrides)


def pour(n: int = 1_000_000, seed: int = 3, box=POUR_BOX, **overrides):
    """The polydisperse pour (no bonds: call Scene.step without Scene.initialize)."""
    return pour_config(box, **overrides), pour_grains(n, box, seed), pour_walls(box)


def pour_packed(n: int = 1_000_000, scale: float = 1.25, seed: int = 3, vmax: float = 0.05,
                box=POUR_BOX, **overrides):
    """A contact-heavy state of the pour: the seeded grains with radii scaled by `scale` after
    placement (volume and mass by scale^3), so neighbours overlap, and velocities uniform in
    [-vmax, vmax] per axis (a tangential component keeps every contact away from the
    exactly-normal case).  Hertz-Mindlin only."""
    cfg, g, w = pour(n, seed=seed, box=box, **overrides)
    g = g.copy()
    g["V"] *= scale ** 3
    g["m"] *= scale ** 3
    rng = np.random.default_rng(seed)
    g["v"] = rng.uniform(-vmax, vmax, size=(n, 3))
    return cfg, g, w


def batched_cubes(nscenes: int = 256, seed: int = 4, nmin: int = 500, nmax: int = 20000,
                  dt: float = 5e-7, base_gap: float = 1e-3):
    """The batched ensemble: bonded cubes with particle counts log-uniform in [nmin, nmax], per
    scene E in [1e7, 1e9] Pa, mu in [0.2, 0.8], bond strength in [1e5, 1e7] Pa (tensile and
    shear) and impact speed in [2, 30] m/s.  Returns a list of (cfg, grains, walls)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nscenes):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        nx = max(2, int(round(n ** (1.0 / 3.0))))
        nz = max(2, int(round(n / (nx * nx))))
        E = math.exp(rng.uniform(math.log(1e7), math.log(1e9)))
        mu = rng.uniform(0.2, 0.8)
        sb = math.exp(rng.uniform(math.log(1e5), math.log(1e7)))
        v = rng.uniform(2.0, 30.0)
        cfg, g, w = bonded_cube(nx, nx, nz, speed=v, dt=dt, seed=1000 + k, base_gap=base_gap,
                                particle_elastic_modulus=E, pp_coefficient_friction=mu,
                                bond_tensile_strength=sb, bond_shear_strength=sb)
        out.append((cfg, g, w))
    return out
