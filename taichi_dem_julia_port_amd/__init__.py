"""MI355X-native DEM step engine with the DEMPort (lucifer1004/taichi_dem_julia_port) host API.

The hot path -- initialize! / simulate! over device-resident state -- runs as hand-written gfx950
HIP kernels in libdemhip.so behind the C ABI of include/demhip.h.  This package is the host side:
config (DEMConfig), p4p/p4c IO, the solve() driver and synthetic scene generators.
"""
from .config import DEMConfig, from_dict, from_toml
from .types import CONTACT_DTYPE, GRAIN_DTYPE, GRAIN_IN_DTYPE, WALL_DTYPE, make_walls

__all__ = ["DEMConfig", "from_toml", "from_dict", "CONTACT_DTYPE", "GRAIN_DTYPE",
           "GRAIN_IN_DTYPE", "WALL_DTYPE", "make_walls", "Scene", "solve", "device_scan"]


def __getattr__(name):
    # the executor loads libdemhip.so; import it lazily so config/IO work without a build
    if name in ("Scene", "solve", "device_scan", "DEMError"):
        from . import executor
        return getattr(executor, name)
    raise AttributeError(name)
