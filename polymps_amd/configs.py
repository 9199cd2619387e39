"""Deterministic generators of the dam-break workloads.

The reference ships no particle files (its .grid inputs are not in the
tree), so every workload is generated: lattice particles at cell centres,
seeded uniform jitter from a counter-based splitmix64 stream (identical on
every platform), polygon walls as triangle meshes.  Parameters follow the
reference's own examples: input/InputDamINC.json for the incompressible
polygon-wall cases and input/dam1610_h300_lo0p0050_WC.json for the 2-D
weakly-compressible particle-wall case.

Workloads (particle counts at the named spacing):
  dam2d_wc   2-D WC dam break, particle walls, l0 = 0.005: 5,000 fluid + 1,338 wall/dummy
  dam3d_inc  3-D INC dam break in a closed 1.61 x 0.15 x 0.6 m box (12 triangles),
             fluid block 0.6 x 0.15 x 0.3 m: 216,000 at l0 = 0.005, 1,728,000 at l0 = 0.0025
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import lib as _lib
from .params import MpsParams

FLUID, WALL, DUMMY = 0, 2, 3


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n uniform doubles in [0, 1) from a splitmix64 counter stream."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class Workload:
    name: str
    params: MpsParams
    types: np.ndarray
    pos: np.ndarray
    vel: np.ndarray
    mesh: tuple | None = None
    steps: int = 1
    extra_meshes: dict | None = None   # further meshes by type: {1: deformable (static), 2: forced}
    info: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.types)


# ------------------------------------------------------------------ meshes
def box_mesh(lo, hi):
    """Closed axis-aligned box as 12 triangles (outward normals), vertices
    rounded to float32 as an STL file stores them."""
    x0, y0, z0 = (float(np.float32(v)) for v in lo)
    x1, y1, z1 = (float(np.float32(v)) for v in hi)
    V = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], dtype=np.float64)
    F = np.array([[0, 2, 1], [0, 3, 2],   # z0 (normal -z)
                  [4, 5, 6], [4, 6, 7],   # z1 (+z)
                  [0, 1, 5], [0, 5, 4],   # y0 (-y)
                  [3, 7, 6], [3, 6, 2],   # y1 (+y)
                  [0, 4, 7], [0, 7, 3],   # x0 (-x)
                  [1, 2, 6], [1, 6, 5]],  # x1 (+x)
                 dtype=np.int32)
    return V, F


def face_normals(V, F):
    V3 = V.reshape(-1, 3)
    F3 = F.reshape(-1, 3)
    a, b, c = V3[F3[:, 0]], V3[F3[:, 1]], V3[F3[:, 2]]
    e1, e2 = b - a, c - a
    n = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1],
                  e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                  e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1)
    r = np.sqrt(n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2])
    out = np.zeros_like(n)
    nz = r != 0
    out[nz] = n[nz] / r[nz, None]
    return out.reshape(-1)


def write_stl(path, V, F):
    V3 = np.asarray(V).reshape(-1, 3)
    F3 = np.asarray(F).reshape(-1, 3)
    N = face_normals(V3.reshape(-1), F3.reshape(-1)).reshape(-1, 3)
    with open(path, "wb") as f:
        f.write(b"mps-mi355x generated STL".ljust(80, b" "))
        f.write(struct.pack("<I", len(F3)))
        for t in range(len(F3)):
            f.write(struct.pack("<3f", *N[t]))
            for k in range(3):
                f.write(struct.pack("<3f", *V3[F3[t, k]]))
            f.write(b"\0\0")


def mesh_arrays(V, F):
    V = np.ascontiguousarray(np.asarray(V, dtype=np.float64).reshape(-1))
    F = np.ascontiguousarray(np.asarray(F, dtype=np.int32).reshape(-1))
    return V, F, face_normals(V, F)


# -------------------------------------------------------------- workloads
def dam3d_inc(l0: float = 0.005, tank=(1.61, 0.15, 0.6), block=(0.6, 0.15, 0.3), seed: int = 2,
              jitter: float = 0.01, margin: float = 0.05, steps: int = 500, dt: float = 2e-4,
              tol: float = 1e-8) -> Workload:
    """3-D incompressible dam break with a closed polygon box (12 triangles)."""
    nx, ny, nz = (int(round(block[k] / l0)) for k in range(3))
    if any(abs(c * l0 - b) > 1e-9 for c, b in zip((nx, ny, nz), block)):
        raise ValueError(f"spacing {l0} does not tile the fluid block {block}")
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    # reference order: x slowest ... z fastest is arbitrary; use z, y, x nesting like a raster
    pos = np.stack([(ix.ravel(order="F") + 0.5) * l0, (iy.ravel(order="F") + 0.5) * l0,
                    (iz.ravel(order="F") + 0.5) * l0], axis=1)
    n = len(pos)
    u = splitmix64_uniform(seed, 3 * n).reshape(n, 3)
    pos = pos + (2.0 * u - 1.0) * jitter * l0
    types = np.full(n, FLUID, dtype=np.int32)
    p = _lib.defaults()
    p.wall_type = 1
    p.dim = 3.0
    p.part_dist = l0
    p.time_step = dt
    p.time_simulation = steps * dt
    p.iter_output = 100
    p.cfl_number = 0.2
    p.weight_type = 0
    p.slip_condition = 0
    p.re_s = 2.1
    p.re_l = 3.1
    p.gradient_type = 0
    p.gradient_correction = 1
    p.relax_press = 1.0
    p.mps_type = 2
    p.solver_type = 0
    p.alpha_compressibility = 1e-6
    p.relax_pnd = 0.01
    p.shifting_type = 0
    p.pnd_type = 0
    p.repulsive_force_type = 2
    p.re_repulsive_force = 0.5
    p.expect_max_velocity = 4.0
    p.free_surf_type = 1
    p.pnd_threshold = 0.97
    p.neigh_threshold = 0.85
    p.npcd_threshold = 0.2
    p.collision_type = 0
    p.collision_ratio = 0.2
    p.dist_limit_ratio = 0.85
    p.solver_tol = tol
    for k in range(3):
        p.domain_min[k] = -margin
        p.domain_max[k] = tank[k] + margin
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, 0.0, -9.81
    V, F = box_mesh((0.0, 0.0, 0.0), tank)
    return Workload(f"dam3d_inc_l0_{l0:g}", p, types, pos.reshape(-1), np.zeros(3 * n),
                    mesh=mesh_arrays(V, F), steps=steps,
                    info={"lattice": (nx, ny, nz), "seed": seed, "jitter": jitter})


def dam2d_wc(l0: float = 0.005, tank=(1.0, 0.6), column=(0.25, 0.5), seed: int = 1, jitter: float = 1e-4,
             steps: int = 2000, dt: float = 1e-4) -> Workload:
    """2-D weakly-compressible dam break with 1 wall + 2 dummy particle layers."""
    cx, cy = int(round(column[0] / l0)), int(round(column[1] / l0))
    ix, iy = np.meshgrid(np.arange(cx), np.arange(cy), indexing="ij")
    fluid = np.stack([(ix.ravel(order="F") + 0.5) * l0, (iy.ravel(order="F") + 0.5) * l0], axis=1)
    nf = len(fluid)
    u = splitmix64_uniform(seed, 2 * nf).reshape(nf, 2)
    fluid = fluid + (2.0 * u - 1.0) * jitter * l0
    walls, wtypes = [], []
    ntx, nty = int(round(tank[0] / l0)), int(round(tank[1] / l0))
    for layer in range(3):
        t = WALL if layer == 0 else DUMMY
        off = (layer + 0.5) * l0
        # the bottom row of a layer also closes the corners of the side rows
        for k in range(-(layer + 1), ntx + layer + 1):
            walls.append(((k + 0.5) * l0, -off)); wtypes.append(t)
        for k in range(nty):
            walls.append((-off, (k + 0.5) * l0)); wtypes.append(t)
            walls.append((tank[0] + off, (k + 0.5) * l0)); wtypes.append(t)
    walls = np.array(walls)
    pos2 = np.concatenate([fluid, walls])
    types = np.concatenate([np.full(nf, FLUID, dtype=np.int32), np.array(wtypes, dtype=np.int32)])
    n = len(types)
    pos = np.zeros((n, 3))
    pos[:, :2] = pos2
    p = _lib.defaults()
    p.wall_type = 0
    p.dim = 2.0
    p.part_dist = l0
    p.time_step = dt
    p.time_simulation = steps * dt
    p.iter_output = 200
    p.cfl_number = 0.2
    p.weight_type = 0
    p.slip_condition = 0
    p.re_s = 2.1
    p.re_l = 2.1
    p.gradient_type = 0
    p.gradient_correction = 0
    p.relax_press = 1.0
    p.mps_type = 1
    p.sound_speed = 20.0
    p.gamma = 7.0
    p.shifting_type = 2
    p.pnd_type = 2
    p.diffusive_coef = 0.35
    p.free_surf_type = 0
    p.pnd_threshold = 0.97
    p.neigh_threshold = 0.85
    p.collision_type = 0
    p.density_fluid = 1000.0
    p.density_wall = 1000.0
    p.knm_vs1 = 1e-6
    for k in range(3):
        p.domain_min[k] = 0.0
        p.domain_max[k] = 0.0
    p.domain_max[0], p.domain_max[1] = tank[0], tank[1]
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.81, 0.0
    return Workload(f"dam2d_wc_l0_{l0:g}", p, types, pos.reshape(-1), np.zeros(3 * n), mesh=None, steps=steps,
                    info={"fluid": nf, "walls": n - nf, "seed": seed})


def dam2d_wc_long(l0: float = 0.005, **kw) -> Workload:
    """The 2-D weakly compressible dam break with the large radius 3.1 l0 and
    the corrected pressure gradient of the reference's own 2-D WC example
    (input/dam1610_h300_lo0p0050_WC.json), which keeps the 2,000-step run
    stable; with re_l = 2.1 l0 and no correction the run diverges near step
    1,850 (in the oracle as on the device)."""
    w = dam2d_wc(l0=l0, **kw)
    w.params.re_l = 3.1
    w.params.gradient_correction = 1
    w.name = f"dam2d_wc_long_l0_{l0:g}"
    return w


def periodic_channel2d(l0: float = 0.01, length: float = 0.4, depth: float = 0.06, u0: float = 2.0, seed: int = 3,
                       jitter: float = 1e-4, steps: int = 200, dt: float = 1e-4, mps_type: int = 1) -> Workload:
    """2-D open channel, periodic in x (the reference's domain.boundary.type 1
    with direction x), a floor of 1 wall + 2 dummy particle layers, fluid of
    the given depth moving at u0 along x so that particles cross the seam.
    mps_type 1: weakly compressible; 2: incompressible PPE."""
    nx, cy = int(round(length / l0)), int(round(depth / l0))
    ix, iy = np.meshgrid(np.arange(nx), np.arange(cy), indexing="ij")
    fluid = np.stack([(ix.ravel(order="F") + 0.5) * l0, (iy.ravel(order="F") + 0.5) * l0], axis=1)
    nf = len(fluid)
    u = splitmix64_uniform(seed, 2 * nf).reshape(nf, 2)
    fluid = fluid + (2.0 * u - 1.0) * jitter * l0
    walls, wtypes = [], []
    for layer in range(3):
        for k in range(nx):
            walls.append(((k + 0.5) * l0, -(layer + 0.5) * l0))
            wtypes.append(WALL if layer == 0 else DUMMY)
    walls = np.array(walls)
    n = nf + len(walls)
    pos = np.zeros((n, 3))
    pos[:nf, :2] = fluid
    pos[nf:, :2] = walls
    types = np.concatenate([np.full(nf, FLUID, dtype=np.int32), np.array(wtypes, dtype=np.int32)])
    vel = np.zeros((n, 3))
    vel[:nf, 0] = u0
    w = dam2d_wc(l0=l0, steps=steps, dt=dt)
    p = w.params
    p.domain_type_bc = 1
    p.limit_type_bc = 0
    p.num_bc = 1
    p.periodic_dir[0], p.periodic_dir[1], p.periodic_dir[2] = 1, 0, 0
    p.domain_min[0], p.domain_min[1] = 0.0, -3 * l0
    p.domain_max[0], p.domain_max[1] = length, 2 * depth
    p.mps_type = mps_type
    if mps_type != 1:
        p.pnd_type = 0
        p.shifting_type = 0
        p.re_l = 3.1
        p.gradient_type = 2
        p.gradient_correction = 1
        p.free_surf_type = 1
    return Workload(f"periodic_channel2d_l0_{l0:g}_mps{mps_type}", p, types, pos.reshape(-1), vel.reshape(-1),
                    mesh=None, steps=steps, info={"fluid": nf, "walls": n - nf, "seed": seed, "u0": u0})


def periodic_channel3d(l0: float = 0.02, size=(0.3, 0.4), depth: float = 0.1, u0: float = 1.0, seed: int = 4,
                       jitter: float = 1e-3, steps: int = 100, dt: float = 2e-4) -> Workload:
    """3-D incompressible channel, periodic in y, with a floor and two side
    walls of 1 wall + 2 dummy particle layers; the fluid moves at u0 along y."""
    nx, ny, cz = int(round(size[0] / l0)), int(round(size[1] / l0)), int(round(depth / l0))
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(cz), indexing="ij")
    fluid = np.stack([(ix.ravel(order="F") + 0.5) * l0, (iy.ravel(order="F") + 0.5) * l0,
                      (iz.ravel(order="F") + 0.5) * l0], axis=1)
    nf = len(fluid)
    u = splitmix64_uniform(seed, 3 * nf).reshape(nf, 3)
    fluid = fluid + (2.0 * u - 1.0) * jitter * l0
    walls, wtypes = [], []
    hz = int(round(1.6 * depth / l0))
    for layer in range(3):
        t = WALL if layer == 0 else DUMMY
        off = (layer + 0.5) * l0
        for k in range(ny):
            y = (k + 0.5) * l0
            # floor row of this layer, closing the corners under the side walls
            for m in range(-(layer + 1), nx + layer + 1):
                walls.append(((m + 0.5) * l0, y, -off)); wtypes.append(t)
            for m in range(hz):
                walls.append((-off, y, (m + 0.5) * l0)); wtypes.append(t)
                walls.append((size[0] + off, y, (m + 0.5) * l0)); wtypes.append(t)
    walls = np.array(walls)
    pos = np.concatenate([fluid, walls])
    n = len(pos)
    types = np.concatenate([np.full(nf, FLUID, dtype=np.int32), np.array(wtypes, dtype=np.int32)])
    vel = np.zeros((n, 3))
    vel[:nf, 1] = u0
    p = small_dam3d(l0=0.025, steps=steps).params
    p.part_dist = l0
    p.wall_type = 0
    p.time_step = dt
    p.time_simulation = steps * dt
    p.domain_type_bc = 1
    p.limit_type_bc = 0
    p.num_bc = 1
    p.periodic_dir[0], p.periodic_dir[1], p.periodic_dir[2] = 0, 1, 0
    p.domain_min[0], p.domain_min[1], p.domain_min[2] = -3 * l0, 0.0, -3 * l0
    p.domain_max[0], p.domain_max[1], p.domain_max[2] = size[0] + 3 * l0, size[1], 1.6 * depth
    return Workload(f"periodic_channel3d_l0_{l0:g}", p, types, pos.reshape(-1), vel.reshape(-1), mesh=None,
                    steps=steps, info={"fluid": nf, "walls": n - nf, "seed": seed, "u0": u0})


def periodic_channel3d_polygon(l0: float = 0.02, size=(0.3, 0.4), depth: float = 0.1, u0: float = 2.5,
                               seed: int = 4, jitter: float = 1e-3, steps: int = 100, dt: float = 2e-4) -> Workload:
    """3-D incompressible channel, periodic in y, bounded by a polygon floor
    (z = 0) and two polygon side walls (x = 0, x = size[0]) that extend one
    periodic length beyond each seam; the fluid moves at u0 along y.  The
    periodic axis is sized by the particle positions and the others by the
    domain of the JSON file, as calcDomainLimits sets them for a periodic
    domain with polygon walls (src/MpsBucket.cpp:45-48, 164-171)."""
    nx, ny, cz = int(round(size[0] / l0)), int(round(size[1] / l0)), int(round(depth / l0))
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(cz), indexing="ij")
    fluid = np.stack([(ix.ravel(order="F") + 0.5) * l0, (iy.ravel(order="F") + 0.5) * l0,
                      (iz.ravel(order="F") + 0.5) * l0], axis=1)
    n = len(fluid)
    u = splitmix64_uniform(seed, 3 * n).reshape(n, 3)
    pos = fluid + (2.0 * u - 1.0) * jitter * l0
    types = np.full(n, FLUID, dtype=np.int32)
    vel = np.zeros((n, 3))
    vel[:, 1] = u0
    # the box's floor and x faces (outward normals), six triangles
    V, F = box_mesh((0.0, -size[1], 0.0), (size[0], 2.0 * size[1], 2.0 * depth))
    F = F[[0, 1, 8, 9, 10, 11]]
    p = small_dam3d(l0=0.025, steps=steps).params
    p.part_dist = l0
    p.time_step = dt
    p.time_simulation = steps * dt
    p.domain_type_bc = 1
    p.limit_type_bc = 0
    p.num_bc = 1
    p.periodic_dir[0], p.periodic_dir[1], p.periodic_dir[2] = 0, 1, 0
    # the domain of the non-periodic axes (the periodic one comes from the particles)
    p.domain_min[0], p.domain_min[1], p.domain_min[2] = -0.5 * l0, 0.0, -0.5 * l0
    p.domain_max[0], p.domain_max[1], p.domain_max[2] = size[0] + 0.5 * l0, size[1], 2.0 * depth
    return Workload(f"periodic_channel3d_polygon_l0_{l0:g}", p, types, pos.reshape(-1), vel.reshape(-1),
                    mesh=mesh_arrays(V, F), steps=steps, info={"fluid": n, "seed": seed, "u0": u0})


def _lattice_values(seed: int, ix, iy, octave: int):
    """Seeded uniform [0, 1) value per integer lattice point (splitmix64 of a hash)."""
    key = (np.asarray(ix, dtype=np.int64) * 73856093) ^ (np.asarray(iy, dtype=np.int64) * 19349663) ^ (octave * 83492791)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + key.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def fractal_noise(x, y, seed: int, octaves: int = 5, base: float = 4.0):
    """Sum of bilinear value-noise octaves (frequency x2, amplitude x0.5), in [0, 1]."""
    tot = np.zeros_like(x)
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        f = base * 2.0 ** o
        u, v = x * f, y * f
        i0, j0 = np.floor(u).astype(np.int64), np.floor(v).astype(np.int64)
        fu, fv = u - i0, v - j0
        fu, fv = fu * fu * (3 - 2 * fu), fv * fv * (3 - 2 * fv)
        v00 = _lattice_values(seed, i0, j0, o)
        v10 = _lattice_values(seed, i0 + 1, j0, o)
        v01 = _lattice_values(seed, i0, j0 + 1, o)
        v11 = _lattice_values(seed, i0 + 1, j0 + 1, o)
        tot += amp * ((v00 * (1 - fu) + v10 * fu) * (1 - fv) + (v01 * (1 - fu) + v11 * fu) * fv)
        norm += amp
        amp *= 0.5
    return tot / norm


def channel_mesh(seed: int = 5, length=4.0, width=1.0, height=1.0, nhx: int = 256, nhy: int = 64, amp: float = 0.15,
                 flat_until: float = 1.05, ramp: float = 0.3, boxes: int = 3):
    """The wall mesh of the 4 x 1 x 1 m channel: a heightmap bed of nhx x nhy
    vertices (2 (nhx-1)(nhy-1) triangles, seeded fractal noise of amplitude
    amp, flat under the reservoir x < flat_until and ramped in over `ramp`),
    the channel's four sides and top (10 triangles) and `boxes` seeded
    axis-aligned obstacles of 12 triangles standing on the bed.  Normals point
    away from the fluid.  Vertices are rounded to float32 as in an STL file."""
    xs = np.linspace(0.0, length, nhx)
    ys = np.linspace(0.0, width, nhy)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    t = np.clip((X - flat_until) / ramp, 0.0, 1.0)
    Z = amp * (t * t * (3 - 2 * t)) * fractal_noise(X, Y, seed)
    V = [np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)]
    F = []
    vid = lambda i, j: i * nhy + j  # noqa: E731
    for i in range(nhx - 1):
        for j in range(nhy - 1):
            a, b, c, d = vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)
            F += [[a, c, b], [a, d, c]]   # clockwise from above: normal -z (away from the fluid)
    F = [np.array(F, dtype=np.int64)]
    nv = nhx * nhy
    # sides and top of the channel: the closed box without its bottom pair
    BV, BF = box_mesh((0.0, 0.0, 0.0), (length, width, height))
    V.append(BV)
    F.append(BF[2:] + nv)
    nv += len(BV)
    u = splitmix64_uniform(7_000 + seed, 4 * boxes)
    for k in range(boxes):
        sx, sy = 0.10 + 0.10 * u[4 * k], 0.10 + 0.10 * u[4 * k + 1]
        x0 = 1.6 + (1.8 - sx) * u[4 * k + 2] + 0.6 * k
        y0 = 0.05 + (0.9 - sy) * u[4 * k + 3]
        OV, OF = box_mesh((x0, y0, 0.0), (x0 + sx, y0 + sy, amp + 0.15))
        V.append(OV)
        F.append(OF[:, ::-1] + nv)   # reversed winding: normals into the obstacle, away from the fluid
        nv += len(OV)
    V = np.concatenate(V).astype(np.float32).astype(np.float64)
    F = np.concatenate(F).astype(np.int32)
    return V, F


def channel4m(l0: float = 0.005, seed: int = 5, jitter: float = 0.01, steps: int = 200, dt: float = 2e-4,
              tol: float = 1e-8, reservoir=(1.0, 1.0, 0.5), **mesh_kw) -> Workload:
    """3-D incompressible flow over a heightmap bed with obstacles: channel
    4 x 1 x 1 m, reservoir block 1.0 x 1.0 x 0.5 m at l0 = 0.005 m
    (200 x 200 x 100 = 4,000,000 particles), mixed-precision PPE."""
    w = dam3d_inc(l0=l0, tank=(4.0, 1.0, 1.0), block=reservoir, seed=seed, jitter=jitter, steps=steps, dt=dt,
                  tol=tol)
    V, F = channel_mesh(seed=seed, **mesh_kw)
    w.mesh = mesh_arrays(V, F)
    w.params.precision_mode = 1
    w.name = f"channel4m_l0_{l0:g}"
    w.info.update(triangles=len(F))
    return w


# -------------------------------------------------------------- case files
def case_json(p: MpsParams, grid: str, mesh: str, vtu_folder: str, vtu: dict | None = None,
              extra_mesh: dict | None = None) -> dict:
    """The parameter block as an input JSON file of the reference's schema
    (the keys read by src/MpsInputOutput.cpp:136-500, plus the solver
    extension keys under numerical.semi_implicit_mps)."""
    ax = ("x", "y", "z")
    names = {"particle_grid_file": grid, "mesh_rigid_file": mesh, "vtu_output_folder": vtu_folder}
    if extra_mesh:
        names.update(extra_mesh)
    return {
        "forced_motion": {"rotation": int(p.forced_motion), "center_x": p.forced_center[0],
                          "center_z": p.forced_center[1], "theta_max_deg": p.forced_theta_max,
                          "period": p.forced_period, "velocity": p.forced_velocity},
        "flags": {"wall_type": int(p.wall_type), "fem_MESH": bool(p.fem_on), "forced_MESH": bool(p.forced_on),
                  "output_VTU": dict({"type": 0, "only_freeSurface": False, "pnd": False, "neigh": False,
                                      "deviation": False, "concentration": False, "auxiliar": False,
                                      "non_newtonian": False},
                                     **(vtu or {}))},
        "pathNames": names,
        "domain": {"min": {a: float(p.domain_min[k]) for k, a in enumerate(ax)},
                   "max": {a: float(p.domain_max[k]) for k, a in enumerate(ax)},
                   "boundary": {"type": int(p.domain_type_bc), "limit": int(p.limit_type_bc),
                                "direction": {a: bool(p.periodic_dir[k]) for k, a in enumerate(ax)}}},
        "physical": {"fluid_density": p.density_fluid, "wall_density": p.density_wall,
                     "kinematic_visc": p.knm_vs1, "fluid_type": int(p.fluid_type),
                     "gravity": {a: float(p.gravity[k]) for k, a in enumerate(ax)},
                     "rheological": {"kinematic_visc_phase_2": p.knm_vs2, "fluid_density_phase_1": p.dns_fl1,
                                     "fluid_density_phase_2": p.dns_fl2, "sediment_density": p.dns_sdt,
                                     "power_law_index": p.power_n, "consistency_index": p.meu0,
                                     "phi": {"lower": p.phi_1, "wall": p.phi_wal, "bed": p.phi_bed,
                                             "second": p.phi_2},
                                     "cohes_coeff": p.cohes, "fraction_method": int(p.fraction_method),
                                     "grain_size": p.dg, "I0": p.i0, "mm": p.mm,
                                     "stress_calc_method": int(p.stress_calc_method), "drag_coeff": p.cd,
                                     "volume_fraction": {"min": p.vf_min, "max": p.vf_max}}},
        "numerical": {
            "dimension": p.dim, "particle_dist": p.part_dist, "time_step": p.time_step,
            "final_time": p.time_simulation, "iter_output": int(p.iter_output), "CFL_number": p.cfl_number,
            "weight_type": int(p.weight_type), "slip_condition": int(p.slip_condition),
            "effective_radius": {"small": p.re_s, "large": p.re_l},
            "gradient": {"type": int(p.gradient_type), "correction": bool(p.gradient_correction),
                         "relax_fact": p.relax_press},
            "mps_type": int(p.mps_type),
            "explicit_mps": {"equation_state": {"speed_sound": p.sound_speed, "gamma": p.gamma}},
            "semi_implicit_mps": {"solver_type": int(p.solver_type),
                                  "weak_compressibility": {"alpha": p.alpha_compressibility},
                                  "source_term": {"relax_pnd": p.relax_pnd},
                                  "tolerance": p.solver_tol, "max_iterations": int(p.solver_max_iter),
                                  "precision_mode": int(p.precision_mode),
                                  "eigen_transpose": int(p.ppe_transpose),
                                  "mixed_inner_tolerance": p.mixed_inner_tol},
            "particle_shifting": {"type": int(p.shifting_type), "DRI": p.dri, "coef_A": p.coef_a,
                                  "mach_number": p.mach_number},
            "pnd": {"type": int(p.pnd_type), "diffusive_coeff": p.diffusive_coef},
            "wall_repulsive_force": {"type": int(p.repulsive_force_type), "re": p.re_repulsive_force,
                                     "maxVel": p.expect_max_velocity,
                                     "coefficient": {"Mitsume": p.rep_coef_mitsume,
                                                     "Lennard-Jones": p.rep_coef_lennard_jones,
                                                     "Monaghan-Kajtar": p.rep_coef_monaghan_kajtar},
                                     "eps_re": p.eps_re},
            "free_surface_threshold": {"type": int(p.free_surf_type), "pnd": p.pnd_threshold,
                                       "neigh": p.neigh_threshold, "NPCD": p.npcd_threshold,
                                       "ARC": p.theta_threshold, "normal": p.norm_threshold},
            "particle_collision": {"type": int(p.collision_type), "ratio": p.collision_ratio,
                                   "dist_limit_ratio": p.dist_limit_ratio, "lambda": p.lambda_collision},
            "particle_type": {"ghost": int(p.ghost), "fluid": int(p.fluid), "wall": int(p.wall),
                              "dummyWall": int(p.dummy_wall)},
            "boundary_type": {"free_surface": int(p.surface), "inner": int(p.inner), "other": int(p.other)},
        },
    }


def write_case(w: Workload, folder, name: str = "case", vtu: dict | None = None) -> Path:
    """Write <folder>/<name>.json, <name>.grid and (polygon walls) <name>.stl
    with paths relative to <folder>; returns the JSON path."""
    import json
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    grid = f"{name}.grid"
    lib = _lib.lib()
    n = w.n
    zeros = np.zeros(n)
    rc = lib.mps_write_grid(str(folder / grid).encode(), n, _lib._ptr(w.types), _lib._ptr(w.pos),
                            _lib._ptr(w.vel), _lib._ptr(zeros), _lib._ptr(zeros))
    if rc != 0:
        raise _lib.MpsError(f"mps_write_grid failed ({rc})")
    mesh = "oops"
    if w.mesh is not None:
        mesh = f"{name}.stl"
        write_stl(folder / mesh, w.mesh[0], w.mesh[1])
    extra = {}
    for t, key in ((1, "mesh_deformable_file"), (2, "mesh_forced_file")):
        if w.extra_meshes and t in w.extra_meshes:
            extra[key] = f"{name}_mesh{t}.stl"
            write_stl(folder / extra[key], w.extra_meshes[t][0], w.extra_meshes[t][1])
    path = folder / f"{name}.json"
    path.write_text(json.dumps(case_json(w.params, grid, mesh, f"{name}_out", vtu, extra), indent=1))
    return path


def forced_paddle3d(l0: float = 0.025, dt: float = 1e-3, steps: int = 40, theta_max: float = 30.0,
                    period: float = 0.4, rotation: int = 1, velocity: float = 0.0) -> Workload:
    """Dam break with all three mesh types: the closed tank 0.8 x 0.2 x 0.4 m
    (rigid), a static slab 0.20 x 0.2 x 0.03 m 1 cm above the fluid block
    (the deformable slot, meshType 1) and a paddle plate x in [0.33, 0.35],
    z in [0.02, 0.30] (forced, meshType 2) 3 cm from the block's face,
    rotating about (0.34, 0) with amplitude theta_max degrees and the given
    period (or translating along z at `velocity` with rotation=0)."""
    w = dam3d_inc(l0=l0, tank=(0.8, 0.2, 0.4), block=(0.3, 0.2, 0.2), steps=steps, dt=dt)
    p = w.params
    p.fem_on = 1
    p.forced_on = 1
    p.forced_motion = rotation
    p.forced_center[0], p.forced_center[1] = 0.34, 0.0
    p.forced_theta_max = theta_max
    p.forced_period = period
    p.forced_velocity = velocity
    SV, SF = box_mesh((0.05, 0.0, 0.21), (0.25, 0.2, 0.24))
    PV, PF = box_mesh((0.33, 0.0, 0.02), (0.35, 0.2, 0.30))
    # reversed winding: normals into the solids, away from the fluid
    w.extra_meshes = {1: mesh_arrays(SV, SF[:, ::-1]), 2: mesh_arrays(PV, PF[:, ::-1])}
    w.name = f"forced_paddle3d_l0_{l0:g}"
    return w


def small_dam3d(l0: float = 0.025, **kw) -> Workload:
    """The 3-D dam break at a coarse spacing that tiles the tank
    (0.025 m: 24 x 6 x 12 = 1,728 particles; 0.0125 m: 13,824)."""
    kw.setdefault("steps", 20)
    return dam3d_inc(l0=l0, **kw)


def small_dam2d(l0: float = 0.02, **kw) -> Workload:
    kw.setdefault("steps", 50)
    return dam2d_wc(l0=l0, **kw)


def _granular_params(p, phi=(0.4, 0.611), mm=300.0, vf_max=0.69):
    """Rheology of the reference's granular dam-break inputs
    (input/2D_dam_fraccarollo_*.json, input/damErosion3D_WC.json)."""
    p.fluid_type = 1
    p.knm_vs2 = 1e-6
    p.dns_fl1, p.dns_fl2, p.dns_sdt = 1000.0, 1540.0, 1540.0
    p.power_n, p.meu0 = 1.2, 0.03
    p.phi_1 = p.phi_wal = p.phi_bed = phi[0]
    p.phi_2 = phi[1]
    p.cohes, p.fraction_method = 0.0, 2
    p.dg, p.i0, p.mm, p.cd = 0.0035, 0.75, mm, 0.47
    p.vf_min, p.vf_max = 0.25, vf_max


def granular_dam2d(l0: float = 0.01, tank=(0.6, 0.3), bed: float = 0.06, column=(0.15, 0.15), mps_type: int = 3,
                   seed: int = 11, steps: int = 100, dt: float = 2e-4) -> Workload:
    """2-D dam break over a mobile bed (the reference's Fraccarollo-Capart
    inputs, input/2D_dam_fraccarollo_lo02p50e-03_{INC,WC}.json, at a coarse
    spacing): a granular layer of height `bed` over the whole tank (type 1 in
    the state: fluid of phase 2) and a water column on its left, particle
    walls.  mps_type 3: incompressible PND + div u with gradient correction
    (the INC file); 1: weakly compressible with the diffusive PND and
    gradient type 3 (the WC file)."""
    w = dam2d_wc(l0=l0, tank=tank, column=(tank[0], bed + column[1]), seed=seed, steps=steps, dt=dt)
    pos = w.pos.reshape(-1, 3)
    types = w.types.copy()
    fl = types == FLUID
    keep = ~fl | (pos[:, 1] < bed) | (pos[:, 0] < column[0])
    pos, types = pos[keep], types[keep]
    fl = types == FLUID
    types[fl & (pos[:, 1] < bed)] = 1   # the granular phase (readMpsParticleFile's type 1)
    p = w.params
    _granular_params(p)
    p.mps_type = mps_type
    p.re_s, p.re_l = 2.1, 3.1
    if mps_type == 3:
        p.pnd_type, p.gradient_type, p.gradient_correction, p.shifting_type = 0, 0, 1, 0
        p.relax_pnd, p.alpha_compressibility, p.solver_tol = 0.01, 1e-6, 1e-8
    else:
        p.pnd_type, p.gradient_type, p.gradient_correction, p.shifting_type = 2, 3, 0, 2
        p.time_step = 5e-5
    n = len(types)
    return Workload(f"granular_dam2d_l0_{l0:g}", p, types.astype(np.int32), pos.reshape(-1), np.zeros(3 * n),
                    mesh=None, steps=steps, info={"granular": int((types == 1).sum()), "water": int((types == 0).sum())})


def granular_dam3d(l0: float = 0.025, bed: float = 0.1, steps: int = 40, dt: float = 2e-4, slip: int = 0) -> Workload:
    """3-D dam break over a granular bed in the closed 12-triangle box
    (polygon walls, as input/damErosion3D_WC.json): the lower `bed` metres
    of the dam3d fluid block are the granular phase."""
    w = dam3d_inc(l0=l0, steps=steps, dt=dt)
    pos = w.pos.reshape(-1, 3)
    types = w.types.copy()
    types[(types == FLUID) & (pos[:, 2] < bed)] = 1
    p = w.params
    _granular_params(p, phi=(0.541, 0.6), mm=100.0, vf_max=0.65)
    p.slip_condition = slip
    w.types = types.astype(np.int32)
    w.name = f"granular_dam3d_l0_{l0:g}"
    w.info = dict(w.info or {}, granular=int((types == 1).sum()))
    return w


WORKLOADS = {
    "dam2d_wc": lambda: dam2d_wc(),
    "dam3d_inc_216k": lambda: dam3d_inc(l0=0.005, steps=500),
    "dam3d_inc_1728k": lambda: dam3d_inc(l0=0.0025, steps=300),
    "channel4m": lambda: channel4m(),
}
