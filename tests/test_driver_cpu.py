"""The polymps_gpu program and its case files, without a GPU.

The case writer emits the reference's JSON schema; reading it back through
the library's JSON reader must give the same parameter block.  The program
must fail with a message (not crash) when its input is missing.
"""
import subprocess

import pytest

from polymps_amd import build as B
from polymps_amd import configs
from polymps_amd import lib as L
from polymps_amd.params import MpsParams


@pytest.mark.parametrize("make", [lambda: configs.small_dam3d(l0=0.025), lambda: configs.small_dam2d(l0=0.02),
                                  lambda: configs.periodic_channel2d()])
def test_case_json_round_trip(built, tmp_path, make):
    w = make()
    path = configs.write_case(w, tmp_path, "case")
    p, grid, mesh = L.params_from_json(path)
    assert grid == "case.grid"
    assert mesh == ("case.stl" if w.mesh is not None else "oops")
    a = L.derive(w.params, w.pos, w.types)
    b = L.derive(p, w.pos, w.types)
    for name, _ in MpsParams._fields_:
        x, y = getattr(a, name), getattr(b, name)
        if hasattr(x, "__len__"):
            x, y = list(x), list(y)
        assert x == y, name


def test_driver_reports_missing_input(built, tmp_path):
    exe = B.DRIVER if B.DRIVER.exists() else B.build_driver()
    res = subprocess.run([str(exe), str(tmp_path / "missing.json")], capture_output=True, text=True, timeout=60)
    assert res.returncode == 1
    assert "missing.json" in res.stderr
