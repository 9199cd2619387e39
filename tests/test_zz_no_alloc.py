"""mps_download_noalloc against mps_download through the C ABI (needs a
gfx950 GPU).

The allocation-free download stages the field in the library's per-GPU
area instead of a fresh device buffer; its output must equal
mps_download's bit for bit.  The calls run back to back on the one shared
area: a vector field of a larger context, then fields of a smaller context
(a smaller input, so stale entries of the first call would show), then the
larger context again.
"""
import ctypes

import numpy as np
import pytest

from polymps_amd import configs
from polymps_amd.lib import INT_FIELDS, VEC_FIELDS, Simulation, lib

pytestmark = pytest.mark.gpu

MPS_E_ARG = -1


def _noalloc(sim: Simulation, field: str) -> np.ndarray:
    n = sim.n * (3 if field in VEC_FIELDS else 1)
    out = np.full(n, -7, dtype=np.int32) if field in INT_FIELDS else np.full(n, np.nan)
    rc = lib().mps_download_noalloc(sim._h, field.encode(), out.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_size_t(n))
    assert rc == 0, (rc, lib().mps_last_error(sim._h).decode())
    return out


def _same(sim: Simulation, field: str):
    a = _noalloc(sim, field)
    b = sim.get(field)
    assert a.dtype == b.dtype and a.shape == b.shape
    assert np.array_equal(a, b), field


def test_download_noalloc_matches_download(built):
    big = configs.small_dam3d(l0=0.025)
    small = configs.small_dam2d(l0=0.02)
    print(f"no-alloc download: {big.n} then {small.n} particles", flush=True)
    g = Simulation(big.params, big.types, big.pos, big.vel, mesh=big.mesh, device=0)
    h = Simulation(small.params, small.types, small.pos, small.vel, mesh=small.mesh, device=0)
    assert h.n < g.n
    g.call("init")
    g.steps(3)
    h.call("init")
    h.steps(3)
    for f in ("pos", "press", "bc"):
        _same(g, f)
    # the second, smaller input on the same staging area
    for f in ("vel", "type", "pnd_small"):
        _same(h, f)
    _same(g, "npcd")
    # a too-small output buffer is refused before any device work
    out = np.zeros(3, dtype=np.float64)
    rc = lib().mps_download_noalloc(g._h, b"pos", out.ctypes.data_as(ctypes.c_void_p),
                                    ctypes.c_size_t(3))
    assert rc == MPS_E_ARG
    g.close()
    h.close()
