"""ctypes binding of libdemhip.so (the C ABI declared in include/demhip.h).

The library is built in-tree by ``taichi_dem_julia_port_amd._build`` (``__graft_entry__.build()``).
There is no fallback: if the shared library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdemhip.so")

c_double3 = ctypes.c_double * 3


class DemConfig(ctypes.Structure):
    """dem_config_t: DEMConfig (src/config.jl:3-43) without init_particles."""

    _fields_ = [
        ("domain_min", c_double3),
        ("domain_max", c_double3),
        ("particle_contact_radius_multiplier", ctypes.c_double),
        ("neighboring_search_safety_factor", ctypes.c_double),
        ("particle_elastic_modulus", ctypes.c_double),
        ("particle_poisson_ratio", ctypes.c_double),
        ("wall_normal", c_double3),
        ("wall_distance", ctypes.c_double),
        ("wall_density", ctypes.c_double),
        ("wall_elastic_modulus", ctypes.c_double),
        ("wall_poisson_ratio", ctypes.c_double),
        ("bond_radius_ratio", ctypes.c_double),
        ("bond_elastic_modulus", ctypes.c_double),
        ("bond_poisson_ratio", ctypes.c_double),
        ("bond_compressive_strength", ctypes.c_double),
        ("bond_tensile_strength", ctypes.c_double),
        ("bond_shear_strength", ctypes.c_double),
        ("pp_coefficient_friction", ctypes.c_double),
        ("pp_coefficient_restitution", ctypes.c_double),
        ("pp_coefficient_rolling_resistance", ctypes.c_double),
        ("pw_coefficient_friction", ctypes.c_double),
        ("pw_coefficient_restitution", ctypes.c_double),
        ("pw_coefficient_rolling_resistance", ctypes.c_double),
        ("max_coordinate_number", ctypes.c_int32),
        ("collision_pair_init_capacity_factor", ctypes.c_int32),
        ("tolerance", ctypes.c_double),
        ("max_particle_count", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("gravity", c_double3),
        ("global_damping", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("target_time", ctypes.c_double),
        ("saving_interval_time", ctypes.c_double),
    ]


class DemWall(ctypes.Structure):
    _fields_ = [("n", c_double3), ("d", ctypes.c_double), ("mid", ctypes.c_int32),
                ("pad_", ctypes.c_int32)]


class DemGrainIn(ctypes.Structure):
    _fields_ = [("id", ctypes.c_int32), ("gid", ctypes.c_int32), ("V", ctypes.c_double),
                ("m", ctypes.c_double), ("k", c_double3), ("v", c_double3)]


class DemGrain(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_int32), ("gid", ctypes.c_int32), ("mid", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("V", ctypes.c_double), ("m", ctypes.c_double), ("r", ctypes.c_double),
        ("r0", ctypes.c_double),
        ("k", c_double3), ("v", c_double3), ("a", c_double3),
        ("q", ctypes.c_double * 4),
        ("w", c_double3), ("dw", c_double3),
        ("I", ctypes.c_double * 9),
    ]


class DemContact(ctypes.Structure):
    _fields_ = [("i", ctypes.c_int32), ("j", ctypes.c_int32), ("midi", ctypes.c_int32),
                ("midj", ctypes.c_int32), ("k", c_double3), ("F", c_double3), ("ti", c_double3),
                ("tj", c_double3), ("s", c_double3)]


class DemDerived(ctypes.Structure):
    _fields_ = [("r_max", ctypes.c_double), ("cell_size", ctypes.c_double),
                ("material_density", ctypes.c_double), ("hash_table_size", ctypes.c_int64),
                ("grid_product_i32", ctypes.c_int32), ("contact_slots", ctypes.c_int32),
                ("total_steps", ctypes.c_int64), ("save_per_steps", ctypes.c_int64)]


class DemOptions(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("contact_slots", ctypes.c_int32),
                ("device", ctypes.c_int32), ("flags", ctypes.c_int32)]


DEM_FLAG_INDEX_OWNED_BROAD_PHASE = 1
DEM_FLAG_REFERENCE_GRID = 2
DEM_FLAG_NO_GRAPH = 4
DEM_FLAG_NO_NEAR_LISTS = 8


class DemCounters(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "step", "n", "hash_table_size", "candidate_pairs", "contacts", "bonded", "hertz",
        "wall_contacts", "created", "deactivated", "bonds_broken", "bonds_broken_total",
        "fatal_gap", "overflow", "examined", "grid_cells", "rebuilds", "replays")]


DEM_NUM_KERNELS = 8


class DemKernelTimes(ctypes.Structure):
    _fields_ = [("ms", ctypes.c_double * DEM_NUM_KERNELS),
                ("launches", ctypes.c_int64 * DEM_NUM_KERNELS)]


# struct sizes that must equal the Julia isbits layouts (types.jl)
assert ctypes.sizeof(DemGrain) == 272
assert ctypes.sizeof(DemContact) == 136
assert ctypes.sizeof(DemWall) == 40

DEM_OK = 0
DEM_ERR_INVALID = -1
DEM_ERR_HIP = -2
DEM_ERR_CAPACITY = -3
DEM_ERR_DUPLICATE = -4
DEM_ERR_NONFINITE = -5
DEM_ERR_STATE = -6
DEM_PRECISION_FP64 = 0
DEM_PRECISION_MIXED = 1

# (name, restype, argtypes) for every entry point of include/demhip.h
P = ctypes.POINTER
SIGNATURES = [
    ("dem_abi_version", ctypes.c_int32, []),
    ("dem_create", ctypes.c_int32, [P(DemConfig), P(DemWall), ctypes.c_int32, P(DemGrainIn),
                                    ctypes.c_int64, P(DemOptions), P(ctypes.c_void_p)]),
    ("dem_destroy", None, [ctypes.c_void_p]),
    ("dem_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("dem_initialize", ctypes.c_int32, [ctypes.c_void_p]),
    ("dem_step", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int64]),
    ("dem_get_derived", ctypes.c_int32, [ctypes.c_void_p, P(DemDerived)]),
    ("dem_get_counters", ctypes.c_int32, [ctypes.c_void_p, P(DemCounters)]),
    ("dem_set_diagnostics", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    ("dem_read_grains", ctypes.c_int32, [ctypes.c_void_p, P(DemGrain), ctypes.c_int64]),
    ("dem_write_grains", ctypes.c_int32, [ctypes.c_void_p, P(DemGrain), ctypes.c_int64]),
    ("dem_read_contacts", ctypes.c_int32, [ctypes.c_void_p, P(DemContact), P(ctypes.c_uint8),
                                           ctypes.c_int64, P(ctypes.c_int64)]),
    ("dem_write_contacts", ctypes.c_int32, [ctypes.c_void_p, P(DemContact), P(ctypes.c_uint8),
                                            ctypes.c_int64]),
    ("dem_read_wall_shear", ctypes.c_int32, [ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int64]),
    ("dem_write_wall_shear", ctypes.c_int32, [ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int64]),
    ("dem_read_forces", ctypes.c_int32, [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double),
                                         ctypes.c_int64]),
    ("dem_debug_cells", ctypes.c_int32, [ctypes.c_void_p, P(ctypes.c_int32), P(ctypes.c_uint32),
                                         P(ctypes.c_double), P(ctypes.c_uint32)]),
    ("dem_debug_hash_table", ctypes.c_int32, [ctypes.c_void_p, P(ctypes.c_uint32),
                                              P(ctypes.c_uint32), P(ctypes.c_uint32)]),
    ("dem_debug_cp_list", ctypes.c_int32, [ctypes.c_void_p, P(ctypes.c_int32), ctypes.c_int64,
                                           P(ctypes.c_int64)]),
    ("dem_debug_scan", ctypes.c_int32, [P(ctypes.c_uint32), P(ctypes.c_uint32), ctypes.c_int64,
                                        P(ctypes.c_uint64)]),
    ("dem_scan_noalloc", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    ("dem_set_timing", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32]),
    ("dem_get_kernel_times", ctypes.c_int32, [ctypes.c_void_p, P(DemKernelTimes)]),
    ("dem_kernel_name", ctypes.c_char_p, [ctypes.c_int32]),
    ("dem_set_memory_cache", ctypes.c_int32, [ctypes.c_int64]),
    ("dem_debug_copy", ctypes.c_int32, [ctypes.c_int64, ctypes.c_int32, P(ctypes.c_double)]),
    ("dem_create_batch", ctypes.c_int32, [ctypes.c_int32, P(DemConfig), P(ctypes.c_int32), P(DemWall),
                                          P(ctypes.c_int64), P(DemGrainIn), P(DemOptions),
                                          P(ctypes.c_void_p)]),
    ("dem_batch_offsets", ctypes.c_int32, [ctypes.c_void_p, P(ctypes.c_int64)]),
    ("dem_batch_set_bond_init", ctypes.c_int32, [ctypes.c_void_p, P(ctypes.c_uint8)]),
    ("dem_batch_bonds_broken", ctypes.c_int32, [ctypes.c_void_p, P(ctypes.c_int64)]),
    ("dem_batch_adopt", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_int32]),
    ("dem_batch_repack", ctypes.c_int32, [ctypes.c_void_p, ctypes.c_int32, P(ctypes.c_int32),
                                          ctypes.c_int32, P(DemConfig), P(ctypes.c_int32),
                                          P(DemWall), P(ctypes.c_int64), P(DemGrainIn),
                                          P(DemOptions), P(ctypes.c_void_p)]),
    ("dem_generate_pour", ctypes.c_int32, [ctypes.c_int64, P(ctypes.c_double), P(ctypes.c_double), ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_uint64, P(DemGrainIn)]),
]

_LIB = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libdemhip.so and declare every entry point; raises if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise ImportError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.dem_abi_version() != 2:
        raise ImportError("libdemhip ABI version mismatch")
    _LIB = lib
    return lib
