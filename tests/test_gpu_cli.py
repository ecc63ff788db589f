"""The native solve() driver and the Python one write the same output.p4p / output.p4c bytes on
the reference's configurations (carom, the bonded 911- and 18112-grain cubes): the same
configuration loader, particle reader, sizing, bond pass, step cadence and save format, over the
same C ABI."""
import os
import shutil
import subprocess

import pytest

from taichi_dem_julia_port_amd import _build

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
