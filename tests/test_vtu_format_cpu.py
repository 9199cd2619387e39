"""The ASCII VTU writer's number formatter (polymps_amd/csrc/fmt_fixed.h)
prints exactly what printf("%f") / printf("%d") print, on the CPU."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_fmt_fixed_matches_printf(tmp_path):
    exe = tmp_path / "fmt_check"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'polymps_amd' / 'csrc'}",
                    str(ROOT / "tests" / "fmt_fixed_check.cpp"), "-o", str(exe)], check=True)
    res = subprocess.run([str(exe), "300000"], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout[-2000:]
    assert "0 mismatches" in res.stdout
