"""The native solve() driver (taichi_dem_julia_port_amd/host/demhip_solve.cpp), without a GPU:
its TOML loader against the Python loader (from_toml) on the reference's configurations, defaults
and malformed input; and the number format of the output files, Julia's Float64 text
(save_single, utils.jl:43,67), in both writers: hand-derived cases of Julia's printing rules and
the native writer against the Python writer on 2,000 values -- so both drivers write the
reference's text and the same bytes (checked on the GPU in test_gpu_cli.py)."""
import dataclasses
import math
import os
import random
import struct
import subprocess

import pytest

from taichi_dem_julia_port_amd import _build
from taichi_dem_julia_port_amd.config import DEMConfig, from_toml
from taichi_dem_julia_port_amd.p4io import julia_float

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CLI = _build.CLI


def _run(*args, check=True):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} is missing: build it with __graft_entry__.build()")
    return subprocess.run([CLI, *args], capture_output=True, text=True, check=check, timeout=60)


def _python_lines(cfg: DEMConfig):
    out = []
    for f in dataclasses.fields(cfg):
        v = getattr(cfg, f.name)
        if isinstance(v, list):
            out.append(f"{f.name} = {' '.join(julia_float(x) for x in v)}")
        elif f.type in ("int", int):
            out.append(f"{f.name} = {int(v)}")
        elif f.type in ("str", str):
            out.append(f"{f.name} = {v}")
        else:
            out.append(f"{f.name} = {julia_float(v)}")
    return out


@pytest.mark.parametrize("k", [1, 2, 3])
def test_loader_matches_python_on_reference_configs(k):
    path = os.path.join(GOLDEN, f"cfg{k}.toml")
    assert _run("--print-config", path).stdout.splitlines() == _python_lines(from_toml(path))


def test_loader_defaults_comments_and_separators(tmp_path):
    empty = tmp_path / "empty.toml"
    empty.write_text("# nothing set\n\n")
    assert _run("--print-config", str(empty)).stdout.splitlines() == _python_lines(DEMConfig())
    p = tmp_path / "misc.toml"
    p.write_text('dt = 1e-6  # seconds\nmax_coordinate_number = 1_6\n'
                 'init_particles = "a # b.p4p"\ngravity = [ 0.0 , -1.5e0 ,2 ]\n')
    got = _run("--print-config", str(p)).stdout.splitlines()
    assert got == _python_lines(from_toml(str(p)))
    assert "init_particles = a # b.p4p" in got and "max_coordinate_number = 16" in got


@pytest.mark.parametrize("text, message", [
    ("bogus = 1\n", "unknown DEMConfig key"),
    ("gravity = [0.0, 1.0]\n", "3 components"),
    ("dt = fast\n", "bad number"),
    ("init_particles = bunny.p4p\n", "must be a string"),
    ("dt\n", "expected key = value"),
])
def test_loader_rejects_malformed_input(tmp_path, text, message):
    p = tmp_path / "bad.toml"
    p.write_text(text)
    r = _run("--print-config", str(p), check=False)
    assert r.returncode == 1 and message in r.stderr, r.stderr


# Julia's text for a Float64: shortest round-trip digits; decimal notation when the decimal point
# position pt (x = 0.d1d2... x 10^pt) is in (-4, 6], else d.ddd e N; whole numbers keep ".0"
JULIA_TEXT = [
    (0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (-1.5, "-1.5"), (0.1, "0.1"), (25.0, "25.0"),
    (1e-5, "1.0e-5"), (1.5e-5, "1.5e-5"), (1e-4, "0.0001"), (1.234e-4, "0.0001234"),
    (0.001, "0.001"), (4.18879e-9, "4.18879e-9"), (100000.0, "100000.0"), (999999.0, "999999.0"),
    (123456.7, "123456.7"), (1e6, "1.0e6"), (1234567.0, "1.234567e6"), (-9.81, "-9.81"),
    (28e9, "2.8e10"), (1e16, "1.0e16"), (2.0 / 3.0 * 1e-5, "6.666666666666667e-6"),
    (0.30000000000000004, "0.30000000000000004"), (5e-324, "5.0e-324"),
    (1.7976931348623157e308, "1.7976931348623157e308"), (float("nan"), "NaN"),
    (float("inf"), "Inf"), (float("-inf"), "-Inf"),
]


def test_number_format_is_julias():
    xs = [x for x, _ in JULIA_TEXT]
    want = [t for _, t in JULIA_TEXT]
    assert [julia_float(x) for x in xs] == want
    assert _run("--format", *[repr(x) for x in xs]).stdout.splitlines() == want
    # every text reads back as the same double
    for x, t in JULIA_TEXT:
        assert math.isnan(x) or float(t) == x


def test_native_number_format_matches_the_python_writer():
    vals = [0.0, -0.0, 1.0, -1.5, 0.1, 1e-5, 1e-4, 1.234e-4, 1e16, 1e15, 9.999999999999999e15,
            1.2345678901234567e17, 2.56e-6, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
            0.30000000000000004, -9.81, 7e10, 28e9, 0.03125, 1234567.0, 99999.99999999999,
            100000.00000000001, 999999.9999999999, 0.00010000000000000002, 9.999999999999999e-5]
    rng = random.Random(11)
    while len(vals) < 2000:
        x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
        if math.isfinite(x):
            vals.append(x)
        vals.append(rng.uniform(-1, 1) * 10.0 ** rng.randint(-30, 30))
        vals.append(float(rng.randint(-10 ** 7, 10 ** 7)) * 10.0 ** rng.randint(-3, 2))
    out = []
    for i in range(0, len(vals), 400):
        out += _run("--format", *[repr(v) for v in vals[i:i + 400]]).stdout.splitlines()
    assert out == [julia_float(v) for v in vals]
    assert all(float(t) == v for t, v in zip(out, vals))


def test_usage_errors():
    assert _run("--bogus-flag", check=False).returncode in (1, 2)
    assert _run(os.path.join(GOLDEN, "cfg1.toml"), "--precision", "fp16", check=False).returncode == 2
