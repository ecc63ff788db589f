"""numpy mirrors of the reference's isbits types (src/types.jl) and of include/demhip.h.

The byte layouts equal Julia's, so arrays of these dtypes can be handed to the C ABI (and would
be to a Julia ccall) without conversion.
"""
import numpy as np

# dem_grain_in_t: one .p4p row (executor.jl:219-220)
GRAIN_IN_DTYPE = np.dtype([
    ("id", "<i4"), ("gid", "<i4"), ("V", "<f8"), ("m", "<f8"), ("k", "<f8", 3), ("v", "<f8", 3),
])

# Grain{Int32,Float64} (types.jl:31-46)
GRAIN_DTYPE = np.dtype([
    ("id", "<i4"), ("gid", "<i4"), ("mid", "<i4"), ("pad_", "<i4"),
    ("V", "<f8"), ("m", "<f8"), ("r", "<f8"), ("r0", "<f8"),
    ("k", "<f8", 3), ("v", "<f8", 3), ("a", "<f8", 3), ("q", "<f8", 4),
    ("w", "<f8", 3), ("dw", "<f8", 3), ("I", "<f8", 9),
])

# Contact{Int32,Float64} (types.jl:68-78)
CONTACT_DTYPE = np.dtype([
    ("i", "<i4"), ("j", "<i4"), ("midi", "<i4"), ("midj", "<i4"),
    ("k", "<f8", 3), ("F", "<f8", 3), ("ti", "<f8", 3), ("tj", "<f8", 3), ("s", "<f8", 3),
])

# Wall{Int32,Float64} (types.jl:60-64)
WALL_DTYPE = np.dtype([("n", "<f8", 3), ("d", "<f8"), ("mid", "<i4"), ("pad_", "<i4")])

assert GRAIN_IN_DTYPE.itemsize == 72
assert GRAIN_DTYPE.itemsize == 272
assert CONTACT_DTYPE.itemsize == 136
assert WALL_DTYPE.itemsize == 40


def make_walls(walls) -> np.ndarray:
    """walls: iterable of (normal[3], d[, mid=2]) -> WALL_DTYPE array."""
    rows = list(walls)
    out = np.zeros(len(rows), dtype=WALL_DTYPE)
    for k, w in enumerate(rows):
        out[k]["n"] = np.asarray(w[0], dtype=np.float64)
        out[k]["d"] = float(w[1])
        out[k]["mid"] = int(w[2]) if len(w) > 2 else 2
    return out


def grains_in(ids, V, m, k, v, gid=None) -> np.ndarray:
    n = len(V)
    out = np.zeros(n, dtype=GRAIN_IN_DTYPE)
    out["id"] = np.asarray(ids, dtype=np.int32)
    out["gid"] = 0 if gid is None else np.asarray(gid, dtype=np.int32)
    out["V"] = V
    out["m"] = m
    out["k"] = k
    out["v"] = v
    return out
