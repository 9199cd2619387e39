"""Pins the oracle's polygon-wall (mirror) operators.

Every check runs on a fluid lattice over the floor z = 0 of a wide box
(tests/pin_util.floor_case) and states the expected value either as a
closed form -- the lattice together with its ghost image is a full lattice,
on which the MPS operators are exact for low-order fields -- or as the same
operator summed over explicit ghost particles on a jittered lattice.
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from pin_util import floor_case, ghost_set, real_set, reflect_floor, w0

L0 = 0.01


def setup_floor(jitter=0.0, slip=0, **kw):
    p, t, x, mesh = floor_case(l0=L0, jitter=jitter, slip=slip, **kw)
    o = Oracle(p, t, x, mesh=mesh)
    for op in ["update_buckets", "closest_point", "update_near_mesh"]:
        o.call(op)
    return p, o, x.reshape(-1, 3)


def deep_xy(P, cells=4.0):
    """Particles at least `cells` spacings inside the block in x and y."""
    lo, hi = P[:, :2].min(0), P[:, :2].max(0)
    return np.all((P[:, :2] > lo + cells * L0 - 1e-12) & (P[:, :2] < hi - cells * L0 + 1e-12), axis=1)


def test_floor_mirror_and_near_wall(built):
    p, o, P = setup_floor(jitter=0.1)
    q = o.params
    assert np.allclose(o.get("mirror").reshape(-1, 3), reflect_floor(P), atol=1e-15)
    assert np.array_equal(o.get("near_wall") == 1, P[:, 2] < q.re_l)
    n = o.get("polygon_normal").reshape(-1, 3)
    assert np.allclose(n, [0.0, 0.0, -1.0]) or np.allclose(n, [0.0, 0.0, 1.0])


# ------------------------------------------------------------- wall NPCD
def test_wall_npcd_ghost_sum(built):
    """calcWallNPCD (src/MpsPndNeigh.cpp:105-195): sum over the ghosts g
    within reS of (x_g - x_i) w(|x_g - x_i|), the particle's own ghost
    included, for near-wall fluid particles; others unchanged."""
    p, o, P = setup_floor(jitter=0.1)
    q = o.params
    o.set("npcd", np.zeros(P.size))
    o.call("wall_npcd")
    got = o.get("npcd").reshape(-1, 3)
    near = o.get("near_wall") == 1
    assert np.all(got[~near] == 0.0)
    for i in np.flatnonzero(near)[::7]:
        _, d, r = ghost_set(P, i, q.re_s)
        expect = (d * w0(r, q.re_s)[:, None]).sum(0)
        assert np.allclose(got[i], expect, rtol=1e-12, atol=1e-12 * L0), i
    assert np.abs(got[near, 2]).max() > 1.0 * L0      # the wall term is not small


def test_wall_npcd_completes_half_space_lattice(built):
    """On a perfect half-space lattice the ghosts complete the full lattice:
    calcWallNPCD + calcPndnNeighNPCD give zero deviation for the near-wall
    particles (the fluid part alone points away from the wall)."""
    p, o, P = setup_floor()
    q = o.params
    o.set("npcd", np.zeros(P.size))
    o.call("pnd_neigh_npcd")
    fluid_only = o.get("npcd").reshape(-1, 3)
    o.set("npcd", np.zeros(P.size))
    o.call("wall_npcd")
    o.call("pnd_neigh_npcd")
    both = o.get("npcd").reshape(-1, 3)
    sel = deep_xy(P) & (P[:, 2] < q.re_s)
    assert sel.sum() >= 36
    assert np.abs(fluid_only[sel, 2]).min() > 1e-3 * L0
    assert np.abs(both[sel]).max() < 1e-13 * L0


# -------------------------------------------------------- wall viscosity
def test_wall_viscosity_no_slip_linear_shear(built):
    """No slip at a fixed wall: the ghost of j moves with -u_j, so a shear
    flow u = (g z, 0, 0) continues linearly through the wall and the
    Laplacian of the full lattice vanishes: calcViscosity + calcWallNoSlip-
    Viscosity give 0 for the near-wall particles.  The wall part is
    -coeff nu sum_g 2 g z_j w(|x_g - x_i|) and the wall force its negative
    times rho l0^3."""
    p, o, P = setup_floor(slip=0)
    q = o.params
    g = 3.0
    vel = np.zeros_like(P)
    vel[:, 0] = g * P[:, 2]
    o.set("vel", vel.reshape(-1))
    o.call("viscosity")
    fluid = o.get("acc").reshape(-1, 3).copy()
    o.set("force_wall", np.zeros(P.size))
    o.call("wall_viscosity")
    total = o.get("acc").reshape(-1, 3)
    wall = total - fluid
    sel = deep_xy(P) & (P[:, 2] < q.re_l)
    nu = p.knm_vs1
    scale = q.coeff_visc_multiphase * nu * g * L0 * q.pnd_large_zero
    assert np.abs(wall[sel, 0]).min() > 1e-3 * scale
    assert np.abs(total[sel]).max() < 1e-11 * scale
    for i in np.flatnonzero(sel)[::5]:
        idx, d, r = ghost_set(P, i, q.re_l)
        expect = q.coeff_visc_multiphase * nu * ((-vel[idx, 0] - vel[i, 0]) * w0(r, q.re_l)).sum()
        assert wall[i, 0] == pytest.approx(expect, rel=1e-11)
    fw = o.get("force_wall").reshape(-1, 3)
    assert np.allclose(fw[sel], -wall[sel] * L0 ** 3 * p.density_fluid, rtol=1e-12, atol=0)


def test_wall_viscosity_free_slip_even_field(built):
    """Free slip: the ghost of j moves with R u_j, so a tangential velocity
    even in z, u = (a z^2 + b, 0, 0), continues as the same quadratic and
    the full-lattice Laplacian is exact: 2 a nu along x."""
    p, o, P = setup_floor(slip=1)
    q = o.params
    a, b = 40.0, 0.3
    vel = np.zeros_like(P)
    vel[:, 0] = a * P[:, 2] ** 2 + b
    o.set("vel", vel.reshape(-1))
    o.call("viscosity")
    o.call("wall_viscosity")
    acc = o.get("acc").reshape(-1, 3)
    sel = deep_xy(P) & (P[:, 2] < q.re_l)
    nu = p.knm_vs1
    assert np.allclose(acc[sel, 0], 2 * a * nu, rtol=1e-9)
    assert np.abs(acc[sel, 1:]).max() < 1e-9 * 2 * a * nu


@pytest.mark.parametrize("slip", [0, 1])
def test_wall_viscosity_ghost_sum(built, slip):
    """Jittered lattice, random velocities: the wall viscosity term equals
    coeff * nu * sum over ghosts within reL of (u_g - u_i) w."""
    p, o, P = setup_floor(jitter=0.1, slip=slip)
    q = o.params
    rng = np.random.default_rng(5)
    vel = rng.normal(size=P.shape)
    o.set("vel", vel.reshape(-1))
    o.set("acc", np.zeros(P.size))
    o.call("wall_viscosity")
    got = o.get("acc").reshape(-1, 3)
    near = o.get("near_wall") == 1
    assert np.all(got[~near] == 0.0)
    gv = vel * ([1.0, 1.0, -1.0] if slip else -1.0)
    for i in np.flatnonzero(near)[::9]:
        idx, d, r = ghost_set(P, i, q.re_l)
        expect = q.coeff_visc_multiphase * p.knm_vs1 * ((gv[idx] - vel[i]) * w0(r, q.re_l)[:, None]).sum(0)
        assert np.allclose(got[i], expect, rtol=1e-11, atol=1e-12 * np.abs(expect).max()), i


# ------------------------------------------------- wall pressure gradient
@pytest.mark.parametrize("gtype", [0, 1, 2, 3])
def test_wall_press_gradient_horizontal_linear_field(built, gtype):
    """Ghosts carry the pressure of their particle; for a field linear in
    x and y the full lattice gives the exact gradient with every gradient
    form (src/MpsPressure.cpp:1284-1294 and :1430-1446): calcPressGradient +
    calcWallPressGradient = -grad p / rho for the near-wall particles."""
    p, o, P = setup_floor()
    p.gradient_type = gtype
    o.close()
    o = Oracle(p, np.zeros(len(P), np.int32), P.reshape(-1), mesh=floor_case(l0=L0)[3])
    for op in ["update_buckets", "closest_point", "update_near_mesh"]:
        o.call(op)
    q = o.params
    grad = np.array([700.0, -300.0, 0.0])
    o.set("press", 2000.0 + P @ grad)
    o.set("pndi", np.full(len(P), q.pnd_small_zero))
    o.call("press_gradient")
    fluid = o.get("acc").reshape(-1, 3).copy()
    o.call("wall_press_gradient")
    acc = o.get("acc").reshape(-1, 3)
    sel = deep_xy(P) & (P[:, 2] < q.re_s)
    assert np.abs(fluid[sel, 2]).min() > 1e-3 * np.abs(grad).max() / p.density_fluid
    assert np.allclose(acc[sel], -grad / p.density_fluid, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("gtype", [0, 1, 2, 3])
def test_wall_press_gradient_ghost_sum(built, gtype):
    """Jittered lattice, random pressure and PND: the wall term equals
    relax * coeffPressGrad / rho * sum over ghosts within reS of
    (x_g - x_i) w f(p_g, p_i) / |x_g - x_i|^2, with p_min over the real
    neighbours and the particle itself; no repulsion above 0.3 l0."""
    p, t, x, mesh = floor_case(l0=L0, jitter=0.1)
    p.gradient_type = gtype
    p.relax_press = 0.8
    o = Oracle(p, t, x, mesh=mesh)
    for op in ["update_buckets", "closest_point", "update_near_mesh"]:
        o.call(op)
    q = o.params
    P = x.reshape(-1, 3)
    rng = np.random.default_rng(9)
    press = rng.uniform(100.0, 900.0, len(P))
    pnd = q.pnd_small_zero * rng.uniform(0.9, 1.1, len(P))
    o.set("press", press)
    o.set("pndi", pnd)
    o.set("acc", np.zeros(P.size))
    o.set("force_wall", np.zeros(P.size))
    o.call("wall_press_gradient")
    got = o.get("acc").reshape(-1, 3)
    near = o.get("near_wall") == 1
    for i in np.flatnonzero(near)[::9]:
        jr, _, _ = real_set(P, i, q.re_s)
        pmin = min(press[i], press[jr].min())
        idx, d, r = ghost_set(P, i, q.re_s)
        pj, pi = press[idx], press[i]
        f = {0: pj - pmin, 1: pj + pi, 2: pj + pi - 2 * pmin,
             3: pnd[i] * pj / pnd[idx] + pnd[idx] * pi / pnd[i]}[gtype]
        s = (d * (w0(r, q.re_s) * f / r ** 2)[:, None]).sum(0)
        expect = p.relax_press * q.coeff_press_grad * s / p.density_fluid
        assert np.allclose(got[i], expect, rtol=1e-11, atol=1e-12 * np.abs(expect).max()), i
    fw = o.get("force_wall").reshape(-1, 3)
    assert np.allclose(fw[near], -got[near] * L0 ** 3 * p.density_fluid, rtol=1e-12, atol=0)


@pytest.mark.parametrize("law", [0, 1, 2, 3])
def test_repulsive_force_laws(built, law):
    """One particle at height h below the repulsion radius r over the floor,
    zero pressure: only the repulsive force acts, along the wall normal,
    acc_z = W / rho and force_wall_z = -W l0^3 with (src/MpsPressure.cpp:1554-1611)
      Harada:           W = rho / dt^2 (r - h)
      Mitsume:          W = C_M (r / h - 1)
      Lennard-Jones:    W = C_LJ v^2 / h ((r/h)^4 - (r/h)^2) rho
      Monaghan-Kajtar:  W = C_MK v^2 / h / 8 (1 + 3 h / (2 r)) (1 - h / r)^3 rho
    with v = min(vel_max, maxVel)."""
    p, t, x, mesh = floor_case(nx=1, ny=1, nz=1, l0=L0)
    h = 0.2 * L0
    p.repulsive_force_type = law
    p.re_repulsive_force = 0.5
    p.expect_max_velocity = 3.0
    P = np.array([[0.5 * L0, 0.5 * L0, h]])
    o = Oracle(p, t, P.reshape(-1), mesh=mesh)
    for op in ["update_buckets", "closest_point", "update_near_mesh"]:
        o.call(op)
    o.set_scalar("vel_max", 2.0)
    o.set("press", np.zeros(1))
    o.set("acc", np.zeros(3))
    o.set("force_wall", np.zeros(3))
    o.call("wall_press_gradient")
    q = o.params
    rho, r, v = p.density_fluid, 0.5 * L0, 2.0
    W = {0: rho / p.time_step ** 2 * (r - h),
         1: p.rep_coef_mitsume * (r / h - 1.0),
         2: p.rep_coef_lennard_jones * v * v / h * ((r / h) ** 4 - (r / h) ** 2) * rho,
         3: p.rep_coef_monaghan_kajtar * v * v / h / 8 * (1 + 1.5 * h / r) * (1 - h / r) ** 3 * rho}[law]
    assert abs(q.re_repulsive_force - r) < 1e-18
    acc = o.get("acc")
    assert acc[2] == pytest.approx(W / rho, rel=1e-12)
    assert abs(acc[0]) < 1e-12 * abs(acc[2]) and abs(acc[1]) < 1e-12 * abs(acc[2])
    assert o.get("force_wall")[2] == pytest.approx(-W * L0 ** 3, rel=1e-12)
    # above the repulsion radius nothing acts
    o.set("pos", np.array([0.5 * L0, 0.5 * L0, 0.6 * L0]))
    o.set("riw2", np.array([10e8 * L0]))     # the predictor's reset (src/MpsParticleVelPos.cpp:37)
    for op in ["update_buckets", "closest_point", "update_near_mesh"]:
        o.call(op)
    o.set("acc", np.zeros(3))
    o.call("wall_press_gradient")
    assert np.all(o.get("acc") == 0.0)


# ------------------------------------------- wall divergence and PND terms
def velocity_field(P, slip):
    """Fields that continue linearly through the wall: free slip keeps the
    tangential part and flips the normal one (u_x, u_y even in z, u_z odd);
    no slip flips all (u odd in z).  Exact divergence: 2.5 - 1.5 + 4.0 = 5.0
    and 4.0."""
    if slip:
        return np.stack([2.5 * P[:, 0], -1.5 * P[:, 1], 4.0 * P[:, 2]], 1), 5.0
    return np.stack([3.0 * P[:, 2], -2.0 * P[:, 2], 4.0 * P[:, 2]], 1), 4.0


@pytest.mark.parametrize("slip", [1, 0])
def test_wall_velocity_divergence_linear_field(built, slip):
    """calcVelDivergence + calcWall{Slip,NoSlip}VelDivergence
    (src/MpsPressure.cpp:479-788) of a field that the ghosts continue
    linearly: the full-lattice divergence is exact."""
    p, o, P = setup_floor(slip=slip)
    q = o.params
    vel, div = velocity_field(P, slip)
    o.set("vel", vel.reshape(-1))
    o.set("pndi", np.full(len(P), q.pnd_small_zero))
    o.call("vel_divergence")
    fluid = o.get("div_u").copy()
    o.call("wall_vel_divergence")
    got = o.get("div_u")
    sel = deep_xy(P) & (P[:, 2] < q.re_s)
    assert np.abs(fluid[sel] - div).min() > 1e-3 * div
    assert np.allclose(got[sel], div, rtol=1e-11)


@pytest.mark.parametrize("slip", [1, 0])
def test_wall_pnd_diffusive_linear_field(built, slip):
    """calcPndDiffusiveTerm + calcWall{Slip,NoSlip}PndDiffusiveTerm +
    updatePnd (src/MpsPndNeigh.cpp:334-774) without pressure or gravity: the
    diffusive term vanishes and n <- n (1 - dt div u) with the exact
    full-lattice divergence."""
    p, t, x, mesh = floor_case(l0=L0, slip=slip, gravity=0.0)
    p.mps_type = 1
    p.pnd_type = 2
    o = Oracle(p, t, x, mesh=mesh)
    for op in ["update_buckets", "closest_point", "update_near_mesh"]:
        o.call(op)
    q = o.params
    P = x.reshape(-1, 3)
    vel, div = velocity_field(P, slip)
    n0 = q.pnd_small_zero
    o.set("vel", vel.reshape(-1))
    o.set("pndi", np.full(len(P), n0))
    o.set("press", np.zeros(len(P)))
    o.call("pnd_diffusive")
    o.call("wall_pnd_diffusive")
    o.call("update_pnd")
    got = o.get("pndi")
    sel = deep_xy(P) & (P[:, 2] < q.re_s)
    assert np.allclose(got[sel], n0 * (1 - p.time_step * div), rtol=1e-13)
    assert np.all(o.get("diffusive")[sel] == 0.0)


def test_mean_wall_pnd_ghost_average(built):
    """meanWallPnd + meanPnd (src/MpsPndNeigh.cpp:855-994): the weighted mean
    of n_j over the real neighbours and the ghosts within reS (the
    particle's own ghost included), weights w(|x_g - x_i|)."""
    p, o, P = setup_floor(jitter=0.1)
    q = o.params
    rng = np.random.default_rng(2)
    pnd = q.pnd_small_zero * (1 + 0.2 * np.sin(P[:, 0] / L0) + 0.3 * P[:, 2] / L0 + 0.01 * rng.normal(size=len(P)))
    o.set("pndi", pnd)
    o.call("pnd_neigh_npcd")        # numNeigh (the mean needs >= 1 neighbour)
    o.set("pndi", pnd)
    o.call("mean_wall_pnd")
    o.call("mean_pnd")
    got = o.get("pndi")
    near = o.get("near_wall") == 1
    for i in range(0, len(P), 11):
        jr, _, rr = real_set(P, i, q.re_s)
        up = (pnd[jr] * w0(rr, q.re_s)).sum()
        do = w0(rr, q.re_s).sum()
        if near[i]:
            jg, _, rg = ghost_set(P, i, q.re_s)
            up += (pnd[jg] * w0(rg, q.re_s)).sum()
            do += w0(rg, q.re_s).sum()
        assert got[i] == pytest.approx(up / do, rel=1e-12), i


def test_wall_normal_completes_half_space_lattice(built):
    """calcNormalParticles + calcWallNormalParticles (src/MpsShifting.cpp:
    106-175, 319-430): with the ghosts the near-wall neighbourhood is a full
    lattice and the free-surface normal vanishes; the fluid part alone
    points down, at the wall."""
    p, o, P = setup_floor()
    q = o.params
    o.call("normal_particles")
    fluid = o.get("normal").reshape(-1, 3).copy()
    o.call("wall_normal_particles")
    both = o.get("normal").reshape(-1, 3)
    sel = deep_xy(P) & (P[:, 2] < q.re_s)
    assert fluid[sel, 2].max() < -1e-3
    assert np.abs(both[sel]).max() < 1e-14
