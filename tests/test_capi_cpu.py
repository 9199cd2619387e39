"""Host-side checks of the C ABI library (no GPU calls).

The library must load without a GPU, export every entry point that
include/mps_gpu.h declares, and its host logic (JSON defaults and parsing,
derived constants, STL / .grid files) must agree with the reference's
semantics.
"""
import ctypes
import json
import re
from pathlib import Path

import numpy as np
import pytest

from polymps_amd import configs
from polymps_amd import lib as L
from polymps_amd.params import MpsParams

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "mps_gpu.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mps_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(built):
    lib = L.lib()
    syms = declared_symbols()
    assert len(syms) >= 45
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_params_layout(built):
    assert L.lib().mps_params_size() == ctypes.sizeof(MpsParams)


def test_defaults_match_reference(built):
    # json.value(key, default) defaults of src/MpsInputOutput.cpp:136-500
    p = L.defaults()
    expect = dict(wall_type=1, density_fluid=1000.0, density_wall=1000.0, knm_vs1=1e-6, fluid_type=0,
                  dns_fl1=1000.0, dns_fl2=1540.0, dim=3.0, part_dist=0.01, time_step=0.0005,
                  time_simulation=1.0, iter_output=80, cfl_number=0.2, weight_type=0, slip_condition=0,
                  re_s=2.1, re_l=2.1, gradient_type=0, gradient_correction=0, relax_press=1.0, mps_type=1,
                  sound_speed=15.0, gamma=7.0, solver_type=0, alpha_compressibility=1e-6, relax_pnd=0.001,
                  shifting_type=2, dri=0.01, coef_a=2.0, mach_number=0.1, pnd_type=0, diffusive_coef=0.35,
                  repulsive_force_type=2, re_repulsive_force=0.5, expect_max_velocity=6.0,
                  rep_coef_mitsume=40000000.0, rep_coef_lennard_jones=2.0, rep_coef_monaghan_kajtar=1.0,
                  eps_re=0.01, free_surf_type=0, pnd_threshold=0.98, neigh_threshold=0.85, npcd_threshold=0.2,
                  theta_threshold=45.0, norm_threshold=0.1, collision_type=0, collision_ratio=0.2,
                  dist_limit_ratio=0.85, lambda_collision=0.2, ghost=-1, fluid=0, wall=2, dummy_wall=3,
                  surface=1, inner=0, other=-1)
    for k, v in expect.items():
        assert getattr(p, k) == v, k
    assert list(p.gravity) == [0.0, 0.0, -9.81]


def test_json_comments_duplicates_and_defaults(built, tmp_path):
    text = """{
      // line comment
      "description": {"_comment": "a", "_comment": "b"},
      /* block
         comment */
      "flags": {"wall_type": 0, "_comment": "x"},
      "pathNames": {"particle_grid_file": "input/grid/x.grid", "mesh_rigid_file": "input/stl/box.stl"},
      "domain": {"min": {"x": -0.1, "y": -0.2}, "max": {"x": 1.5, "y": 0.9, "z": 0.0}},
      "numerical": {"dimension": 2, "particle_dist": 0.005, "mps_type": 2,
                    "effective_radius": {"small": 2.1, "large": 3.1},
                    "semi_implicit_mps": {"solver_type": 0, "source_term": {"relax_pnd": 0.01},
                                          "tolerance": 1e-8},
                    "free_surface_threshold": {"type": 1, "pnd": 0.97}}
    }"""
    f = tmp_path / "case.json"
    f.write_text(text)
    p, grid, mesh = L.params_from_json(f)
    assert grid == "input/grid/x.grid" and mesh == "input/stl/box.stl"
    assert p.wall_type == 0 and p.dim == 2.0 and p.part_dist == 0.005 and p.mps_type == 2
    assert p.re_s == 2.1 and p.re_l == 3.1 and p.relax_pnd == 0.01 and p.solver_tol == 1e-8
    assert p.free_surf_type == 1 and p.pnd_threshold == 0.97
    assert list(p.domain_min) == [-0.1, -0.2, 0.0] and list(p.domain_max) == [1.5, 0.9, 0.0]
    # keys not given keep the defaults
    assert p.sound_speed == 15.0 and p.alpha_compressibility == 1e-6 and p.npcd_threshold == 0.2
    assert list(p.gravity) == [0.0, 0.0, -9.81]


def test_json_incompressible_needs_pnd_type_0(built, tmp_path):
    f = tmp_path / "bad.json"
    f.write_text(json.dumps({"numerical": {"mps_type": 2, "pnd": {"type": 2}}}))
    with pytest.raises(L.MpsError):
        L.params_from_json(f)


@pytest.mark.parametrize("make", [lambda: configs.small_dam3d(l0=0.025), lambda: configs.small_dam2d(l0=0.02),
                                  lambda: configs.periodic_channel2d(), lambda: configs.periodic_channel3d()])
def test_derive_matches_oracle(built, make):
    from oracle.oracle import Oracle
    w = make()
    d = L.derive(w.params, w.pos, w.types)
    po = Oracle(w.params, w.types, w.pos, mesh=w.mesh).params
    for name, _ in MpsParams._fields_:
        a, b = getattr(d, name), getattr(po, name)
        if hasattr(a, "__len__"):
            a, b = list(a), list(b)
        assert a == b, name


def test_stl_round_trip(built, tmp_path):
    V, F = configs.box_mesh((0, 0, 0), (1.61, 0.15, 0.6))
    path = tmp_path / "box.stl"
    configs.write_stl(path, V.reshape(-1), F.reshape(-1))
    V2, F2, N2 = L.read_stl(path)
    assert len(V2) == 8 * 3 and len(F2) == 12 * 3
    Vr, Fr = V2.reshape(-1, 3), F2.reshape(-1, 3)
    # every face is preserved geometrically and has a unit outward normal
    centre = np.array([0.805, 0.075, 0.3])
    for t in range(12):
        a, b, c = Vr[Fr[t]]
        n = N2[t * 3:t * 3 + 3]
        assert abs(np.linalg.norm(n) - 1.0) < 1e-15
        assert np.dot(n, (a + b + c) / 3 - centre) > 0
    assert np.allclose(sorted(map(tuple, Vr)), sorted(map(tuple, V)))


def test_grid_round_trip(built, tmp_path):
    w = configs.small_dam2d(l0=0.02)
    lib = L.lib()
    n = w.n
    press = np.linspace(0, 1, n)
    pavg = np.zeros(n)
    path = tmp_path / "s.grid"
    P = L._ptr
    assert lib.mps_write_grid(str(path).encode(), n, P(w.types), P(w.pos), P(w.vel), P(press), P(pavg)) == 0
    nn = ctypes.c_int(n)
    t2, p2, v2, pr2, pa2 = np.zeros(n, np.int32), np.zeros(3 * n), np.zeros(3 * n), np.zeros(n), np.zeros(n)
    assert lib.mps_read_grid(str(path).encode(), ctypes.byref(nn), P(t2), P(p2), P(v2), P(pr2), P(pa2)) == 0
    assert nn.value == n and np.array_equal(t2, w.types) and np.array_equal(p2, w.pos) and np.array_equal(pr2, press)


def test_workload_sizes():
    assert configs.dam3d_inc(l0=0.005, steps=1).n == 216000
    w = configs.dam2d_wc()
    assert w.info["fluid"] == 5000 and 1200 < w.info["walls"] < 1400


def test_solver_failure_status_mapping(built):
    """MPS_E_SOLVER of include/mps_gpu.h reaches Python as MpsSolverError,
    every other non-zero code as MpsError (a subclass relation the callers
    that carry on after a failed solve rely on)."""
    text = (ROOT / "include" / "mps_gpu.h").read_text()
    code = int(re.search(r"#define MPS_E_SOLVER \((-?\d+)\)", text).group(1))
    assert code == L.MPS_E_SOLVER

    class Stub:
        _h = None
        _lib = L.lib()

    with pytest.raises(L.MpsSolverError):
        L.Simulation._check(Stub(), code)
    with pytest.raises(L.MpsError) as e:
        L.Simulation._check(Stub(), -1)
    assert not isinstance(e.value, L.MpsSolverError)
    L.Simulation._check(Stub(), 0)
