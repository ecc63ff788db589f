"""The native solve() driver (taichi_dem_julia_port_amd/host/demhip_solve.cpp), without a GPU:
its TOML loader against the Python loader (from_toml) on the reference's configurations, defaults
and malformed input, and its number format against Python's repr, which the Python p4p/p4c writer
uses -- so both drivers write the same bytes (checked on the GPU in test_gpu_cli.py)."""
import dataclasses
import math
import os
import random
import struct
import subprocess

import pytest

from taichi_dem_julia_port_amd import _build
from taichi_dem_julia_port_amd.config import DEMConfig, from_toml

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
            out.append(f"{f.name} = {' '.join(repr(float(x)) for x in v)}")
        elif f.type in ("int", int):
            out.append(f"{f.name} = {int(v)}")
        elif f.type in ("str", str):
            out.append(f"{f.name} = {v}")
        else:
            out.append(f"{f.name} = {repr(float(v))}")
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


def test_number_format_is_python_repr():
    vals = [0.0, -0.0, 1.0, -1.5, 0.1, 1e-5, 1e-4, 1.234e-4, 1e16, 1e15, 9.999999999999999e15,
            1.2345678901234567e17, 2.56e-6, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
            0.30000000000000004, -9.81, 7e10, 28e9, 0.03125, 1234567.0]
    rng = random.Random(11)
    while len(vals) < 2000:
        x = struct.unpack("<d", struct.pack("<Q", rng.getrandbits(64)))[0]
        if math.isfinite(x):
            vals.append(x)
        vals.append(rng.uniform(-1, 1) * 10.0 ** rng.randint(-30, 30))
    args = [repr(v) for v in vals]
    out = []
    for i in range(0, len(args), 400):
        out += _run("--repr", *args[i:i + 400]).stdout.splitlines()
    assert out == args


def test_usage_errors():
    assert _run("--bogus-flag", check=False).returncode in (1, 2)
    assert _run(os.path.join(GOLDEN, "cfg1.toml"), "--precision", "fp16", check=False).returncode == 2
