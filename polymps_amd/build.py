"""Build the gfx950 MPS library (polymps_amd/libmps_gpu.so) with hipcc.

Every translation unit in polymps_amd/csrc is compiled for gfx950 only, with
FMA contraction off (FP64 results must match the CPU path bit for bit) and
packed-FP32 instruction selection disabled, then linked into one shared
library that exports the C ABI of include/mps_gpu.h.

Usage:  python -m polymps_amd.build   (or polymps_amd.build.build())
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmps_gpu.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MPS_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
HIP_ONLY = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ("_hip" if src.suffix == ".hip" else "_cpp") + ".o")
    deps = [src] + list(CSRC.glob("*.h")) + list((ROOT / "include").glob("*.h"))
    if obj.exists() and all(obj.stat().st_mtime >= d.stat().st_mtime for d in deps):
        return obj
    if src.suffix == ".hip":
        cmd = [HIPCC, "-x", "hip", *COMMON, *HIP_ONLY, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-D__HIP_PLATFORM_AMD__", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(force: bool = False, jobs: int | None = None) -> Path:
    """Compile every TU (in parallel) and link libmps_gpu.so; returns its path."""
    if force and OBJ.exists():
        shutil.rmtree(OBJ)
    OBJ.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


DRIVER = PKG / "polymps_gpu"


def build_driver() -> Path:
    """Link the polymps_gpu program (host/main.cpp over libmps_gpu.so)."""
    host = PKG / "host"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", f"-I{host}", str(host / "main.cpp"),
           f"-L{PKG}", "-lmps_gpu", "-Wl,-rpath,$ORIGIN", "-o", str(DRIVER)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return DRIVER


def build_oracle() -> Path:
    """Compile the CPU oracle (test infrastructure) with its own Makefile."""
    odir = ROOT / "oracle"
    res = subprocess.run(["make", "-C", str(odir)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{res.stdout}\n{res.stderr}")
    return odir / "liboracle_mps.so"


if __name__ == "__main__":
    force = "--force" in sys.argv
    print(build(force=force))
    print(build_driver())
    print(build_oracle())
