"""The native solve() driver and the Python one write the same output.p4p / output.p4c bytes on
the reference's configurations (carom, the bonded 911- and 18112-grain cubes): the same
configuration loader, particle reader, sizing, bond pass, step cadence and save format, over the
same C ABI.  And the native driver's files pass the reference's own comparers (compare_p4p and
compare_p4c, utils.jl:170-209, at their default tolerance) against the same loop run on the
oracle (DEMPort.solve's cadence, executor.jl:341-386, written by the same frame writers): the
driver path end to end against the reference's algorithm, not only against itself."""
import os
import shutil
import subprocess

import pytest

from oracle.oracle import OracleScene
from taichi_dem_julia_port_amd import _build, from_toml
from taichi_dem_julia_port_amd.p4io import (compare_p4c, compare_p4p, read_p4c, read_p4p,
                                            read_p4p_rows, write_p4c_frame, write_p4p_frame)

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = {1: "carom.p4p", 2: "cube_911_particles_impact.p4p", 3: "cube_18112_particles_impact.p4p.gz"}


def _layout(tmp_path, k):
    """cfg/<k>.toml next to Resources/, as in the reference repository."""
    (tmp_path / "cfg").mkdir()
    (tmp_path / "Resources").mkdir()
    shutil.copy(os.path.join(GOLDEN, FILES[k]), tmp_path / "Resources" / FILES[k])
    text = open(os.path.join(GOLDEN, f"cfg{k}.toml")).read()
    if FILES[k].endswith(".gz"):
        text = text.replace(FILES[k][:-3] + '"', FILES[k] + '"')
    cfg = tmp_path / "cfg" / f"{k}.toml"
    cfg.write_text(text)
    return str(cfg)


@pytest.mark.parametrize("k, max_steps, no_bond, precision", [
    (1, 3000, False, "fp64"), (2, 3000, False, "fp64"), (2, 2000, True, "fp64"),
    (3, 20000, False, "fp64"), (2, 3000, False, "mixed")])
def test_native_driver_writes_the_python_drivers_bytes(tmp_path, k, max_steps, no_bond, precision):
    from taichi_dem_julia_port_amd.executor import solve
    print(f"[case] demhip_solve cfg{k} max_steps={max_steps} no_bond={no_bond} {precision}", flush=True)
    cfg = _layout(tmp_path, k)
    a, b = tmp_path / "native", tmp_path / "python"
    args = [_build.CLI, cfg, "--out", str(a), "--max-steps", str(max_steps), "--precision", precision]
    if no_bond:
        args.append("--no-bond")
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Total time" in r.stderr
    solve(cfg, no_bond=no_bond, out_dir=str(b), max_steps=max_steps, precision=precision).close()
    for name in ("output.p4p", "output.p4c"):
        x, y = (a / name).read_bytes(), (b / name).read_bytes()
        assert x.count(b"TIMESTEP") >= 3, name
        assert x == y, name


def _oracle_solve(cfg_path, out_dir, max_steps, no_bond, parallel=False):
    """DEMPort.solve's loop (executor.jl:341-386) on the oracle: frame 0, then one frame after
    every save interval up to max_steps, through the same writers as the engine's drivers."""
    cfg = from_toml(cfg_path)
    rows = read_p4p_rows(os.path.join(os.path.dirname(cfg_path), cfg.init_particles))
    o = OracleScene(cfg, rows, parallel=parallel)
    d = o.derived()
    total, every = min(d["total_steps"], max_steps), max(d["save_per_steps"], 1)
    if not no_bond:
        o.initialize()
    os.makedirs(out_dir)
    with open(os.path.join(out_dir, "output.p4p"), "w") as p4p, \
            open(os.path.join(out_dir, "output.p4c"), "w") as p4c:
        write_p4p_frame(p4p, o.read_grains(), 0.0)
        write_p4c_frame(p4c, *o.read_contacts(), 0.0)
        step = 0
        while step < total:
            o.step(every)
            step += every
            write_p4p_frame(p4p, o.read_grains(), step * cfg.dt)
            write_p4c_frame(p4c, *o.read_contacts(), step * cfg.dt)
    o.close()


@pytest.mark.parametrize("k, max_steps, no_bond, parallel", [
    (1, 5000, False, False), (2, 5000, False, False), (2, 3000, True, False), (3, 10000, False, True)])
def test_native_driver_output_passes_the_reference_comparers_against_the_oracle(
        tmp_path, k, max_steps, no_bond, parallel):
    print(f"[case] demhip_solve cfg{k} max_steps={max_steps} no_bond={no_bond} vs oracle", flush=True)
    cfg = _layout(tmp_path, k)
    a, b = tmp_path / "native", tmp_path / "oracle"
    args = [_build.CLI, cfg, "--out", str(a), "--max-steps", str(max_steps)]
    if no_bond:
        args.append("--no-bond")
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    _oracle_solve(cfg, str(b), max_steps, no_bond, parallel)
    pa, pb = read_p4p(str(a / "output.p4p")), read_p4p(str(b / "output.p4p"))
    ca, cb = read_p4c(str(a / "output.p4c")), read_p4c(str(b / "output.p4c"))
    assert len(pa) >= 2 and len(ca) == len(pa)
    if not no_bond:
        assert sum(len(c) for _, c in ca) > 0, "the run must hold bonds"
    compare_p4p(pa, pb)
    compare_p4c(ca, cb)
