"""Python host binding of the C ABI (include/mps_gpu.h).

``Simulation`` mirrors the reference driver's objects (MpsParticleSystem /
MpsParticle plus the operator classes): one method per operator, each a thin
call into libmps_gpu.so.  There is no fallback: if the library is missing the
import of this module raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

from .params import MpsParams, check_layout

_LIB_PATH = Path(__file__).resolve().parent / "libmps_gpu.so"

_OPS = [
    "bucket_update", "viscosity", "add_gravity", "prediction", "check_out_domain", "build_neighbor_list",
    "prediction_press_gradient", "prediction_wall_press_gradient", "collision", "mesh_closest_point",
    "mesh_update_near", "wall_npcd", "correction_matrix", "initial_pnd",
    "pnd_neigh_npcd", "pnd_diffusive", "mean_pnd", "mean_neigh_fluid_pnd", "update_bc", "calc_press",
    "vel_divergence", "ppe_assemble", "ppe_solve", "ppe_warm_start", "extrapolate_wall_dummy", "press_gradient", "wall_press_gradient",
    "normal_particles", "wall_normal_particles", "wall_viscosity", "correction", "shifting", "wall_dummy_velocity", "init", "step",
    "volume_fraction", "viscosity_interaction", "wall_viscosity_interaction",
]

VEC_FIELDS = {"pos", "vel", "acc", "acc_star", "npcd", "mirror", "wall_pos", "polygon_normal", "force_wall",
              "cm1", "cm2", "cm3", "grad_conc", "normal"}
MESH_FIXED, MESH_DEFORMABLE, MESH_FORCED = 0, 1, 2
INT_FIELDS = {"type", "bc", "num_neigh", "num_neigh_wall", "near_wall", "near_mesh_type", "ptype"}


class MpsError(RuntimeError):
    pass


MPS_E_SOLVER = -6


class MpsSolverError(MpsError):
    """A pressure solve did not reach its tolerance (the state is written
    and a step finishes first) or produced a non-finite residual."""


def load_library(path: str | os.PathLike | None = None) -> ctypes.CDLL:
    p = Path(path) if path else _LIB_PATH
    if not p.exists():
        raise MpsError(f"{p} is not built; run polymps_amd.build.build() (no CPU fallback exists)")
    lib = ctypes.CDLL(str(p))
    vp, ip, dp, cp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p
    P = ctypes.POINTER(MpsParams)
    lib.mps_params_size.restype = ctypes.c_size_t
    lib.mps_params_defaults.argtypes = [P]
    lib.mps_params_from_json.argtypes = [cp, P, cp, ctypes.c_size_t, cp, ctypes.c_size_t]
    lib.mps_params_derive.argtypes = [P, ip, vp, vp]
    lib.mps_create.argtypes = [ctypes.POINTER(vp), P, ip, ip]
    lib.mps_destroy.argtypes = [vp]
    lib.mps_last_error.argtypes = [vp]
    lib.mps_last_error.restype = cp
    lib.mps_get_params.argtypes = [vp, P]
    lib.mps_upload_state.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.mps_download.argtypes = [vp, cp, vp, ctypes.c_size_t]
    lib.mps_download_noalloc.argtypes = [vp, cp, vp, ctypes.c_size_t]
    lib.mps_upload_field.argtypes = [vp, cp, vp, ctypes.c_size_t]
    lib.mps_get_scalar.argtypes = [vp, cp, dp]
    lib.mps_set_scalar.argtypes = [vp, cp, ctypes.c_double]
    lib.mps_sync.argtypes = [vp]
    lib.mps_mesh_read_stl.argtypes = [cp, ctypes.POINTER(ip), vp, ctypes.POINTER(ip), vp, vp]
    lib.mps_set_mesh.argtypes = [vp, ip, vp, ip, vp, vp]
    lib.mps_set_mesh_type.argtypes = [vp, ip, ip, vp, ip, vp, vp]
    lib.mps_get_mesh.argtypes = [vp, ip, ctypes.POINTER(ip), vp, ctypes.POINTER(ip), vp, vp]
    lib.mps_forced_mesh_update.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    lib.mps_mesh_closest_point_type.argtypes = [vp, ip]
    lib.mps_mesh_write_stl.argtypes = [cp, ip, vp, ip, vp]
    lib.mps_mesh_paths_from_json.argtypes = [cp, cp, ctypes.c_size_t, cp, ctypes.c_size_t]
    lib.mps_get_wall_tables.argtypes = [vp, vp, vp, vp, vp]
    lib.mps_ppe_matrix.argtypes = [vp, vp, vp, vp, ctypes.POINTER(ip)]
    lib.mps_get_buckets.argtypes = [vp, vp, vp, vp]
    lib.mps_steps.argtypes = [vp, ip]
    lib.mps_read_grid.argtypes = [cp, ctypes.POINTER(ip), vp, vp, vp, vp, vp]
    lib.mps_write_grid.argtypes = [cp, ip, vp, vp, vp, vp, vp]
    lib.mps_checkpoint_write.argtypes = [vp, cp]
    lib.mps_checkpoint_read.argtypes = [vp, cp]
    lib.mps_write_vtu.argtypes = [vp, cp, ip]
    lib.mps_write_vtu_async.argtypes = [vp, cp, ip]
    lib.mps_output_wait.argtypes = [vp]
    ull = ctypes.c_ulonglong
    lib.mps_gnn_param_count.argtypes = []
    lib.mps_gnn_init.argtypes = [vp, ull]
    lib.mps_gnn_set_params.argtypes = [vp, vp, ip]
    lib.mps_gnn_get_params.argtypes = [vp, vp, ip]
    lib.mps_gnn_forward.argtypes = [vp]
    lib.mps_gnn_graph.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.POINTER(ip)]
    lib.mps_gnn_backward.argtypes = [vp, vp, vp, vp, ip]
    lib.mps_gnn_train_step.argtypes = [vp, ctypes.c_double, ull, dp]
    lib.mps_gnn_save.argtypes = [vp, cp]
    lib.mps_gnn_load.argtypes = [vp, cp]
    for op in _OPS:
        getattr(lib, "mps_" + ("mesh_closest_point" if op == "mesh_closest_point" else op)).argtypes = [vp]
    check_layout(lib)
    return lib


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        _lib = load_library()
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def defaults() -> MpsParams:
    p = MpsParams()
    lib().mps_params_defaults(ctypes.byref(p))
    return p


def params_from_json(path: str | os.PathLike):
    p = MpsParams()
    grid = ctypes.create_string_buffer(4096)
    mesh = ctypes.create_string_buffer(4096)
    rc = lib().mps_params_from_json(str(path).encode(), ctypes.byref(p), grid, 4096, mesh, 4096)
    if rc:
        raise MpsError(lib().mps_last_error(None).decode())
    return p, grid.value.decode(), mesh.value.decode()


def derive(p: MpsParams, pos: np.ndarray | None = None, types: np.ndarray | None = None) -> MpsParams:
    q = p.copy()
    n = 0 if pos is None else len(pos) // 3
    pp = None if pos is None else _ptr(np.ascontiguousarray(pos, dtype=np.float64))
    tp = None if types is None else _ptr(np.ascontiguousarray(types, dtype=np.int32))
    rc = lib().mps_params_derive(ctypes.byref(q), n, pp, tp)
    if rc:
        raise MpsError(lib().mps_last_error(None).decode())
    return q


def read_stl(path: str | os.PathLike):
    L = lib()
    nv, nf = ctypes.c_int(0), ctypes.c_int(0)
    rc = L.mps_mesh_read_stl(str(path).encode(), ctypes.byref(nv), None, ctypes.byref(nf), None, None)
    if rc:
        raise MpsError(L.mps_last_error(None).decode())
    V = np.zeros(3 * nv.value)
    F = np.zeros(3 * nf.value, dtype=np.int32)
    N = np.zeros(3 * nf.value)
    rc = L.mps_mesh_read_stl(str(path).encode(), ctypes.byref(nv), _ptr(V), ctypes.byref(nf), _ptr(F), _ptr(N))
    if rc:
        raise MpsError(L.mps_last_error(None).decode())
    return V, F, N


class Simulation:
    """A device-resident particle system driven through the C ABI."""

    def __init__(self, params: MpsParams, types, pos, vel=None, press=None, press_avg=None, mesh=None, device=0,
                 meshes=None):
        """mesh: the rigid mesh (V, F, N); meshes: further meshes by type
        ({MESH_DEFORMABLE: (V, F, N), MESH_FORCED: (V, F, N)})."""
        self._lib = lib()
        types = np.ascontiguousarray(types, dtype=np.int32)
        n = len(types)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        vel = np.zeros(3 * n) if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1)
        press = np.zeros(n) if press is None else np.ascontiguousarray(press, dtype=np.float64)
        press_avg = np.zeros(n) if press_avg is None else np.ascontiguousarray(press_avg, dtype=np.float64)
        self.params = params if params.derived else derive(params, pos, types)
        self.n = n
        h = ctypes.c_void_p()
        rc = self._lib.mps_create(ctypes.byref(h), ctypes.byref(self.params), n, device)
        self._h = h
        if rc:
            msg = self._lib.mps_last_error(h).decode()
            self._lib.mps_destroy(h)
            self._h = None
            raise MpsError(f"mps_create failed ({rc}): {msg}")
        self._check(self._lib.mps_upload_state(h, _ptr(types), _ptr(pos), _ptr(vel), _ptr(press), _ptr(press_avg)))
        if mesh is not None:
            V, F, N = (np.ascontiguousarray(mesh[0], dtype=np.float64), np.ascontiguousarray(mesh[1], dtype=np.int32),
                       np.ascontiguousarray(mesh[2], dtype=np.float64))
            self._check(self._lib.mps_set_mesh(h, len(V) // 3, _ptr(V), len(F) // 3, _ptr(F), _ptr(N)))
        for t, m in (meshes or {}).items():
            self.set_mesh_type(t, *m)

    def set_mesh_type(self, mesh_type: int, V, F, N):
        V, F, N = (np.ascontiguousarray(V, dtype=np.float64), np.ascontiguousarray(F, dtype=np.int32),
                   np.ascontiguousarray(N, dtype=np.float64))
        self._check(self._lib.mps_set_mesh_type(self._h, mesh_type, len(V) // 3, _ptr(V), len(F) // 3, _ptr(F),
                                                _ptr(N)))

    def get_mesh(self, mesh_type: int):
        nv, nf = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.mps_get_mesh(self._h, mesh_type, ctypes.byref(nv), None, ctypes.byref(nf), None, None))
        V, F, N = np.zeros(3 * nv.value), np.zeros(3 * nf.value, dtype=np.int32), np.zeros(3 * nf.value)
        self._check(self._lib.mps_get_mesh(self._h, mesh_type, ctypes.byref(nv), _ptr(V), ctypes.byref(nf), _ptr(F),
                                           _ptr(N)))
        return V, F, N

    def forced_mesh_update(self, dt: float, time: float):
        self._check(self._lib.mps_forced_mesh_update(self._h, dt, time))

    def closest_point_type(self, mesh_type: int):
        self._check(self._lib.mps_mesh_closest_point_type(self._h, mesh_type))

    def _check(self, rc):
        if rc == MPS_E_SOLVER:
            raise MpsSolverError(f"MPS call failed ({rc}): {self._lib.mps_last_error(self._h).decode()}")
        if rc:
            raise MpsError(f"MPS call failed ({rc}): {self._lib.mps_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            self._lib.mps_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def call(self, op: str):
        self._check(getattr(self._lib, "mps_" + op)(self._h))

    def steps(self, k: int):
        self._check(self._lib.mps_steps(self._h, k))

    def get(self, field: str) -> np.ndarray:
        if field in INT_FIELDS:
            out = np.zeros(self.n, dtype=np.int32)
        elif field in VEC_FIELDS:
            out = np.zeros(3 * self.n)
        else:
            out = np.zeros(self.n)
        self._check(self._lib.mps_download(self._h, field.encode(), _ptr(out), out.size))
        return out

    def set(self, field: str, values) -> None:
        dt = np.int32 if field in INT_FIELDS else np.float64
        arr = np.ascontiguousarray(values, dtype=dt).reshape(-1)
        self._check(self._lib.mps_upload_field(self._h, field.encode(), _ptr(arr), arr.size))

    def scalar(self, name: str) -> float:
        v = ctypes.c_double()
        self._check(self._lib.mps_get_scalar(self._h, name.encode(), ctypes.byref(v)))
        return v.value

    def set_scalar(self, name: str, value: float) -> None:
        self._check(self._lib.mps_set_scalar(self._h, name.encode(), float(value)))

    def ppe_matrix(self):
        nnz = ctypes.c_int()
        self._check(self._lib.mps_ppe_matrix(self._h, None, None, None, ctypes.byref(nnz)))
        nl = int(self.scalar("num_particles"))
        rp = np.zeros(nl + 1, dtype=np.int32)
        cols = np.zeros(nnz.value, dtype=np.int32)
        vals = np.zeros(nnz.value)
        self._check(self._lib.mps_ppe_matrix(self._h, _ptr(rp), _ptr(cols), _ptr(vals), ctypes.byref(nnz)))
        return rp, cols, vals

    def checkpoint_write(self, path) -> None:
        """Write the whole state (every field carried between steps, time,
        counters, moving meshes) for a bit-exact restart."""
        self._check(self._lib.mps_checkpoint_write(self._h, str(path).encode()))

    def checkpoint_read(self, path) -> None:
        """Restore a checkpoint into this context (same parameters,
        capacity and meshes); the run continues bit for bit."""
        self._check(self._lib.mps_checkpoint_read(self._h, str(path).encode()))

    def buckets(self):
        """(first, last, next) bucket lists of the last bucket update in
        reference indices (firstParticleInBucket, lastParticleInBucket,
        nextParticleInSameBucket of the reference)."""
        nb = self._lib.mps_get_buckets(self._h, None, None, None)
        if nb < 0:
            self._check(nb)
        first = np.zeros(nb, dtype=np.int32)
        last = np.zeros(nb, dtype=np.int32)
        nxt = np.zeros(self.n, dtype=np.int32)
        rc = self._lib.mps_get_buckets(self._h, _ptr(first), _ptr(last), _ptr(nxt))
        if rc < 0:
            self._check(rc)
        return first, last, nxt

    # ------------------------------------------------ learned PPE operator
    def gnn_init(self, seed: int = 0):
        self._check(self._lib.mps_gnn_init(self._h, seed))

    def gnn_params(self) -> np.ndarray:
        n = self._lib.mps_gnn_param_count()
        th = np.zeros(n)
        self._check(self._lib.mps_gnn_get_params(self._h, _ptr(th), n))
        return th

    def gnn_set_params(self, theta) -> None:
        th = np.ascontiguousarray(theta, dtype=np.float64)
        self._check(self._lib.mps_gnn_set_params(self._h, _ptr(th), th.size))

    def gnn_forward(self) -> None:
        self._check(self._lib.mps_gnn_forward(self._h))

    def gnn_graph(self) -> dict:
        """Graph, features, diagonal and operator of the last forward pass, in
        reference order (CSR over the live particles)."""
        ne = ctypes.c_int()
        self._check(self._lib.mps_gnn_graph(self._h, None, None, None, None, None, None, None, None, ctypes.byref(ne)))
        nl = int(self.scalar("num_particles"))
        out = dict(indptr=np.zeros(nl + 1, dtype=np.int32), cols=np.zeros(ne.value, dtype=np.int32),
                   e=np.zeros(4 * ne.value), w=np.zeros(ne.value), feat=np.zeros(4 * nl), diag=np.zeros(nl),
                   cs=np.zeros(ne.value), dsi=np.zeros(nl))
        self._check(self._lib.mps_gnn_graph(self._h, *(_ptr(out[k]) for k in ("indptr", "cols", "e", "w", "feat",
                                                                              "diag", "cs", "dsi")),
                                            ctypes.byref(ne)))
        out["e"] = out["e"].reshape(-1, 4)
        out["feat"] = out["feat"].reshape(-1, 4)
        return out

    def gnn_backward(self, r, gz) -> np.ndarray:
        n = self._lib.mps_gnn_param_count()
        g = np.zeros(n)
        r = np.ascontiguousarray(r, dtype=np.float64)
        gz = np.ascontiguousarray(gz, dtype=np.float64)
        self._check(self._lib.mps_gnn_backward(self._h, _ptr(r), _ptr(gz), _ptr(g), n))
        return g

    def gnn_train_step(self, lr: float, seed: int) -> float:
        loss = ctypes.c_double()
        self._check(self._lib.mps_gnn_train_step(self._h, lr, seed, ctypes.byref(loss)))
        return loss.value

    def gnn_save(self, path) -> None:
        self._check(self._lib.mps_gnn_save(self._h, str(path).encode()))

    def gnn_load(self, path) -> None:
        self._check(self._lib.mps_gnn_load(self._h, str(path).encode()))

    def write_vtu(self, path, flags: int = 1):
        self._check(self._lib.mps_write_vtu(self._h, str(path).encode(), flags))

    def write_vtu_async(self, path, flags: int = 1):
        """Snapshot now; the file is written in the background (output_wait)."""
        self._check(self._lib.mps_write_vtu_async(self._h, str(path).encode(), flags))

    def output_wait(self):
        self._check(self._lib.mps_output_wait(self._h))

    def sync(self):
        self._check(self._lib.mps_sync(self._h))
