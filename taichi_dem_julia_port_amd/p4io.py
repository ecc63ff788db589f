"""p4p / p4c particle and contact files (src/utils.jl:30-78, 126-209; src/executor.jl:217-243).

File semantics kept from the reference:
* p4p frames: ``TIMESTEP PARTICLES`` / ``t n`` / header / n rows ``ID GROUP VOLUME MASS PX PY PZ VX
  VY VZ``; particles are written in array (file) order, identified by ID.
* p4c frames: ``TIMESTEP CONTACTS`` / ``t n`` / header / n rows ``P1 P2 CX CY CZ FX FY FZ
  CONTACT_IS_BONDED`` where P1 < P2 are 1-based *array indices* (utils.jl:54-67), F is the contact
  force in the local contact frame, and a duplicate pair is an error (utils.jl:58-60).
* comparisons use isapprox with rtol = sqrt(eps) by default (utils.jl:176,197).
"""
from __future__ import annotations

import gzip
import math
from decimal import Decimal
from typing import Dict, List, Tuple

import numpy as np

from .types import CONTACT_DTYPE, GRAIN_IN_DTYPE

IO_GRAIN_DTYPE = np.dtype([("id", "<i4"), ("V", "<f8"), ("m", "<f8"), ("k", "<f8", 3),
                           ("v", "<f8", 3)])


def _open(path, mode="rt"):
    return gzip.open(path, mode) if str(path).endswith(".gz") else open(path, mode)


def read_p4p_rows(path: str) -> np.ndarray:
    """First frame of a p4p file in file order, as solve() reads it (executor.jl:217-220)."""
    with _open(path) as fh:
        lines = fh.read().split("\n")
    n = int(lines[1].split()[1])
    rows = np.zeros(n, dtype=GRAIN_IN_DTYPE)
    data = np.array([list(map(float, ln.split()[:10])) for ln in lines[3:3 + n]], dtype=np.float64)
    rows["id"] = data[:, 0].astype(np.int32)
    rows["gid"] = data[:, 1].astype(np.int32)
    rows["V"] = data[:, 2]
    rows["m"] = data[:, 3]
    rows["k"] = data[:, 4:7]
    rows["v"] = data[:, 7:10]
    return rows


def read_p4p(path: str) -> List[Tuple[float, np.ndarray]]:
    """All frames; grains stored at position ID (utils.jl:126-145)."""
    with _open(path) as fh:
        lines = fh.read().split("\n")
    out = []
    i = 0
    while i < len(lines) and lines[i].strip():
        assert lines[i].startswith("TIMESTEP")
        t, n = lines[i + 1].split()
        n = int(float(n))
        g = np.zeros(n, dtype=IO_GRAIN_DTYPE)
        for j in range(n):
            vals = [float(x) for x in lines[i + 3 + j].split()[:10]]
            pid = int(vals[0])
            g[pid - 1] = (pid, vals[2], vals[3], vals[4:7], vals[7:10])
        out.append((float(t), g))
        i += 3 + n
    return out


def read_p4c(path: str) -> List[Tuple[float, List[Tuple[Tuple[int, int], tuple]]]]:
    """All frames as sorted ((p1, p2), (k, F, bonded)) lists (utils.jl:147-168)."""
    with _open(path) as fh:
        lines = fh.read().split("\n")
    out = []
    i = 0
    while i < len(lines) and lines[i].strip():
        assert lines[i].startswith("TIMESTEP")
        t, n = lines[i + 1].split()
        n = int(float(n))
        cur: Dict[Tuple[int, int], tuple] = {}
        for j in range(n):
            v = [float(x) for x in lines[i + 3 + j].split()[:9]]
            cur[(int(v[0]), int(v[1]))] = (np.array(v[2:5]), np.array(v[5:8]), bool(v[8]))
        out.append((float(t), sorted(cur.items(), key=lambda kv: kv[0])))
        i += 3 + n
    return out


def julia_float(x: float) -> str:
    """The text Julia 1.9 writes for a Float64 in string interpolation ("$x"), as save_single
    does (utils.jl:43,67): the shortest digits that read back as x, in decimal notation when the
    decimal point falls at -4 < pt <= 6 (x = 0.d1d2... * 10^pt and, for integral values, the
    digits dropped to the left of the point are zeros), else d.ddd e N with an unpadded exponent;
    a whole number keeps ".0" (1.0, 100000.0, 1.0e6, 1.0e-5, 0.0001, -0.0, NaN, Inf).  Python's
    repr gives the same shortest digits."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    dec = Decimal(repr(abs(x))).normalize().as_tuple()
    digits = "".join(map(str, dec.digits))
    olength = len(digits)
    pt = dec.exponent + olength
    plain = -4 < pt <= 6
    if plain and pt >= olength:
        step = float(10 ** (pt - olength))
        r = math.fmod(x + 0.05, step)
        if r != 0.0 and (r < 0) != (step < 0):
            r += step
        plain = not abs(r - 0.05) > 0.05
    if plain:
        if pt <= 0:
            body = "0." + "0" * (-pt) + digits
        elif pt >= olength:
            body = digits + "0" * (pt - olength) + ".0"
        else:
            body = digits[:pt] + "." + digits[pt:]
    else:
        body = digits[0] + "." + (digits[1:] if olength > 1 else "0") + "e" + str(pt - 1)
    return sign + body


def _fmt(x: float) -> str:
    return julia_float(x)


def write_p4p_frame(fh, grains: np.ndarray, t: float) -> None:
    """save_single's particle half (utils.jl:36-44); grains is a GRAIN_DTYPE array."""
    n = len(grains)
    fh.write("TIMESTEP PARTICLES\n")
    fh.write(f"{_fmt(t)} {n}\n")
    fh.write("ID  GROUP  VOLUME  MASS  PX  PY  PZ  VX  VY  VZ\n")
    rows = []
    for g in grains:
        k, v = g["k"], g["v"]
        rows.append(f"{int(g['id'])} {int(g['gid'])} {_fmt(g['V'])} {_fmt(g['m'])} {_fmt(k[0])} "
                    f"{_fmt(k[1])} {_fmt(k[2])} {_fmt(v[0])} {_fmt(v[1])} {_fmt(v[2])}\n")
    fh.write("".join(rows))


def write_p4c_frame(fh, contacts: np.ndarray, bonded: np.ndarray, t: float) -> None:
    """save_single's contact half (utils.jl:46-73), with its duplicate-pair check."""
    seen = set()
    rows = []
    for c, b in zip(contacts, bonded):
        key = (int(c["i"]), int(c["j"]))
        if key in seen:
            raise RuntimeError("Duplicate contact pairs!")
        seen.add(key)
        k, F = c["k"], c["F"]
        rows.append(f"{key[0]} {key[1]} {_fmt(k[0])} {_fmt(k[1])} {_fmt(k[2])} {_fmt(F[0])} "
                    f"{_fmt(F[1])} {_fmt(F[2])} {int(bool(b))}\n")
    fh.write("TIMESTEP CONTACTS\n")
    fh.write(f"{_fmt(t)} {len(rows)}\n")
    fh.write("P1  P2  CX  CY  CZ  FX  FY  FZ  CONTACT_IS_BONDED\n")
    fh.write("".join(rows))


def _isapprox(a, b, atol, rtol) -> bool:
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    b = np.atleast_1d(np.asarray(b, dtype=np.float64))
    # Julia isapprox on vectors: norm(a - b) <= max(atol, rtol * max(norm(a), norm(b)))
    return float(np.linalg.norm(a - b)) <= max(atol, rtol * max(np.linalg.norm(a), np.linalg.norm(b)))


def compare_p4p(r1, r2, atol: float = 0.0, rtol: float | None = None) -> None:
    """utils.jl:176-189: raises AssertionError on the first mismatch."""
    if rtol is None:
        rtol = 0.0 if atol > 0 else math.sqrt(np.finfo(np.float64).eps)
    assert len(r1) == len(r2)
    for i, ((t1, g1), (t2, g2)) in enumerate(zip(r1, r2)):
        assert t1 == t2, f"[timestep] i = {i}, left = {t1}, right = {t2}"
        assert len(g1) == len(g2), f"[length] i = {i}, left = {len(g1)} right = {len(g2)}"
        for j in range(len(g1)):
            a, b = g1[j], g2[j]
            assert a["id"] == b["id"], f"[pid] i = {i}, j = {j}"
            for f in ("V", "m", "k", "v"):
                assert _isapprox(a[f], b[f], atol, rtol), \
                    f"[{f}] i = {i}, j = {j}, left = {a[f]}, right = {b[f]}"


def compare_p4c(r1, r2, atol: float = 0.0, rtol: float | None = None) -> None:
    """utils.jl:197-209."""
    if rtol is None:
        rtol = 0.0 if atol > 0 else math.sqrt(np.finfo(np.float64).eps)
    assert len(r1) == len(r2)
    for i, ((t1, c1), (t2, c2)) in enumerate(zip(r1, r2)):
        assert t1 == t2, f"[timestep] i = {i}, left = {t1}, right = {t2}"
        assert len(c1) == len(c2), f"[length] i = {i}, left = {len(c1)} right = {len(c2)}"
        for j in range(len(c1)):
            (p1, a), (p2, b) = c1[j], c2[j]
            assert p1 == p2, f"[pair] i = {i}, j = {j}, left = {p1}, right = {p2}"
            assert _isapprox(a[0], b[0], atol, rtol), f"[position] i = {i}, j = {j}"
            assert _isapprox(a[1], b[1], atol, rtol), f"[force] i = {i}, j = {j}"
            assert a[2] == b[2], f"[bonded] i = {i}, j = {j}"


def contacts_to_records(contacts: np.ndarray, bonded: np.ndarray):
    """In-memory equivalent of one read_p4c frame."""
    return sorted((((int(c["i"]), int(c["j"])), (np.array(c["k"]), np.array(c["F"]), bool(b)))
                   for c, b in zip(contacts, bonded)), key=lambda kv: kv[0])


__all__ = ["read_p4p_rows", "read_p4p", "read_p4c", "write_p4p_frame", "write_p4c_frame",
           "compare_p4p", "compare_p4c", "contacts_to_records", "CONTACT_DTYPE"]
