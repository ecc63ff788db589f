"""DEMConfig: the reference's TOML configuration (src/config.jl:3-43) and its loader.

``from_toml`` mirrors ``from_toml(DEMConfig{Int32,Float64}, cfg_filename)`` (src/executor.jl:182):
keys are the field names, missing keys take the defaults below, unknown keys are an error.
"""
from __future__ import annotations

import dataclasses
import os
from dataclasses import dataclass, field
from typing import List

try:  # Python 3.11+
    import tomllib as _toml
except ModuleNotFoundError:  # pragma: no cover - Python 3.10 in this image
    import tomli as _toml

from . import _native


def _v(*xs):
    return field(default_factory=lambda: list(xs))


@dataclass
class DEMConfig:
    domain_min: List[float] = _v(-200.0, -200.0, -30.0)
    domain_max: List[float] = _v(200.0, 200.0, 90.0)
    init_particles: str = os.path.join("..", "Resources", "bunny.p4p")
    particle_contact_radius_multiplier: float = 1.1
    neighboring_search_safety_factor: float = 1.01
    particle_elastic_modulus: float = 7e10
    particle_poisson_ratio: float = 0.25
    wall_normal: List[float] = _v(0.0, 0.0, -1.0)
    wall_distance: float = 25.0
    wall_density: float = 7800.0
    wall_elastic_modulus: float = 2e11
    wall_poisson_ratio: float = 0.25

    bond_radius_ratio: float = 0.5
    bond_elastic_modulus: float = 28e9
    bond_poisson_ratio: float = 0.2
    bond_compressive_strength: float = 3e8
    bond_tensile_strength: float = 6e7
    bond_shear_strength: float = 6e7

    pp_coefficient_friction: float = 0.3
    pp_coefficient_restitution: float = 0.9
    pp_coefficient_rolling_resistance: float = 0.01

    pw_coefficient_friction: float = 0.35
    pw_coefficient_restitution: float = 0.7
    pw_coefficient_rolling_resistance: float = 0.01

    max_coordinate_number: int = 64
    collision_pair_init_capacity_factor: int = 128

    tolerance: float = 1e-12
    max_particle_count: int = 1000000000

    gravity: List[float] = _v(0.0, 0.0, -9.81)
    global_damping: float = 0.0
    dt: float = 2.63e-5
    target_time: float = 0.5
    saving_interval_time: float = 0.05

    def replace(self, **kw) -> "DEMConfig":
        return dataclasses.replace(self, **kw)

    def to_c(self) -> _native.DemConfig:
        """Fill the dem_config_t mirror (every field but init_particles)."""
        c = _native.DemConfig()
        for f in dataclasses.fields(self):
            if f.name == "init_particles":
                continue
            v = getattr(self, f.name)
            if isinstance(v, list):
                if len(v) != 3:
                    raise ValueError(f"{f.name} must have 3 components")
                getattr(c, f.name)[:] = [float(x) for x in v]
            elif f.type in ("int", int):
                setattr(c, f.name, int(v))
            else:
                setattr(c, f.name, float(v))
        return c


_FIELDS = {f.name: f for f in dataclasses.fields(DEMConfig)}


def from_toml(path: str) -> DEMConfig:
    with open(path, "rb") as fh:
        data = _toml.load(fh)
    return from_dict(data)


def from_dict(data: dict) -> DEMConfig:
    kw = {}
    for k, v in data.items():
        if k not in _FIELDS:
            raise KeyError(f"unknown DEMConfig key {k!r}")
        f = _FIELDS[k]
        if isinstance(v, list):
            kw[k] = [float(x) for x in v]
        elif f.type in ("int", int):
            kw[k] = int(v)
        elif f.type in ("str", str):
            kw[k] = str(v)
        else:
            kw[k] = float(v)
    return DEMConfig(**kw)
