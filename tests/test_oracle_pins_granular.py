"""Pins the oracle's granular (non-Newtonian) phase
(src/MpsViscosity.cpp:104-1526) with checks independent of its code:

* the volume fraction against an O(N^2) sum over all particle pairs;
* the strain rate on a perfect lattice, where the MPS gradient of a linear
  velocity field is exact, and the smoothed pressure of a uniform field;
* the mu(I) law: for simple shear the shear stress 2 MEU |E| over the normal
  stress equals mu(I) = alpha(phi) + (tan phi2 - tan phi) I / (I0 + I),
  alpha the Drucker-Prager match of the friction angle;
* the bed-level grid: the submerged normal stress p_smooth - PS uses the
  pressure of the highest granular particle of each 2 l0 column, the lowest
  index among equally high ones (the defined sequential order);
* the polygon-wall strain rate against explicit ghost particles behind a
  flat floor (free slip: ghost velocity R u_j; no slip: -u_j).
"""
import math

import numpy as np
import pytest

from oracle.oracle import Oracle
from pin_util import floor_case, interior, lattice_block, reflect_floor, w0

L0 = 0.01


def granular(p, phi=(0.4, 0.611)):
    p.fluid_type = 1
    p.knm_vs2 = 1e-6
    p.dns_fl1, p.dns_fl2, p.dns_sdt = 1000.0, 1540.0, 1540.0
    p.power_n, p.meu0 = 1.2, 0.03
    p.phi_1 = p.phi_wal = p.phi_bed = phi[0]
    p.phi_2 = phi[1]
    p.cohes, p.fraction_method = 0.0, 2
    p.dg, p.i0, p.mm, p.cd = 0.0035, 0.75, 300.0, 0.47
    p.vf_min, p.vf_max = 0.25, 0.69
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.81, 0.0
    return p


def block2d(nx=16, ny=16, jitter=0.0, seed=5):
    p, t, x = lattice_block(nx, ny, 1, l0=L0, dim=2, mps_type=2, jitter=jitter, seed=seed)
    p.limit_type_bc = 1
    for k in range(2):
        p.domain_min[k] = -4 * L0
        p.domain_max[k] = (max(nx, ny) + 4) * L0
    return granular(p), t, x


@pytest.mark.parametrize("method", [1, 2])
def test_volume_fraction_brute_force(built, method):
    p, t, x = block2d(jitter=0.2)
    p.fraction_method = method
    rng = np.random.default_rng(3)
    t = np.where(rng.random(len(t)) < 0.4, 1, 0).astype(np.int32)   # type 1: the granular phase
    t[::11] = 2                                                     # some walls, never counted
    o = Oracle(p, t, x)
    o.call("update_buckets")
    o.call("volume_fraction")
    cv = o.get("cv")
    P = x.reshape(-1, 3)
    ty = o.get("type")
    phase = o.get("ptype")
    res = o.params.re_s
    for i in np.flatnonzero(ty == 0):
        d = np.sqrt(((P - P[i]) ** 2).sum(1))
        m = (d < res) & (ty == 0)
        m[i] = False
        wgt = np.ones(m.sum()) if method == 1 else w0(d[m], res)
        s1 = wgt.sum()
        s2 = wgt[phase[m] >= 2].sum()
        expect = 0.0 if s1 < 10 * np.finfo(float).eps else s2 / s1
        assert abs(cv[i] - expect) <= 1e-14, (i, cv[i], expect)
    assert 0.2 < cv[ty == 0].mean() < 0.6


def shear_block(rate, press):
    p, t, x = block2d()
    P = x.reshape(-1, 3)
    v = np.zeros_like(P)
    v[:, 0] = rate * P[:, 1]
    t = np.ones(len(t), dtype=np.int32)   # every particle granular
    o = Oracle(p, t, x, vel=v.reshape(-1), press=np.full(len(t), press))
    o.call("update_buckets")
    o.call("volume_fraction")
    o.call("viscosity_interaction")
    return p, o, interior(x, L0, 3, dim=2)


def test_strain_rate_exact_for_linear_shear(built):
    rate, p0 = 7.0, 500.0
    p, o, m = shear_block(rate, p0)
    assert np.allclose(o.get("s12")[m], 0.5 * rate, rtol=1e-12)
    for f in ("s11", "s22", "s13", "s23", "s33"):
        assert np.allclose(o.get(f)[m], 0.0, atol=1e-12 * rate)
    assert np.allclose(o.get("ii")[m], 0.25 * rate * rate, rtol=1e-12)
    assert np.allclose(o.get("p_smooth")[m], p0, rtol=1e-14)


def test_mu_of_I_law_in_simple_shear(built):
    """Dry granular shear (no phase-1 water, so the water level stays below
    the bed): normal stress = smoothed pressure p0, inertial number
    I = |E| d / sqrt(p0 / rho_s) with |E| = rate / 2, and
    2 MEU |E| / p0 = alpha(phi) + (tan phi2 - tan phi) I / (I0 + I)."""
    for rate in (2.0, 20.0, 200.0):
        p0 = 500.0
        p, o, m = shear_block(rate, p0)
        q = o.params
        E = 0.5 * rate
        inertial = E * q.dg / math.sqrt(p0 / q.dns_sdt)
        vf = max(q.vf_min, q.vf_max - (q.vf_max - q.vf_min) * inertial)
        phi = q.phi_1 * vf / q.vf_max
        alpha = 2.0 * math.sqrt(3.0) * math.sin(phi) / (3.0 - math.sin(phi))
        mu = alpha + (math.tan(q.phi_2 * 1.0) - math.tan(phi)) * inertial / (q.i0 + inertial)
        meu = o.get("meu")[m]
        assert np.allclose(2.0 * meu * E / p0, mu, rtol=1e-9), (rate, (2.0 * meu * E / p0)[:3], mu)
        assert np.allclose(o.get("inertia")[m], inertial, rtol=1e-12)
        assert np.allclose(o.get("vf")[m], vf, rtol=1e-12)
        assert np.allclose(o.get("p_rheo")[m], p0, rtol=1e-14)
        assert np.all(o.get("rho")[m] == q.dns_fl2)   # Cv > 0.5: the phase-2 density


def test_bed_level_grid_submerged(built):
    """Granular bed (y < 8 l0) under water: the submerged normal stress of a
    granular particle is p_smooth - PS, PS the pressure of the highest
    granular particle (Cv > 0.5) of its 2 l0 column, the lowest index among
    equally high ones; every pressure is distinct."""
    p, t, x = block2d(nx=12, ny=14)
    P = x.reshape(-1, 3)
    t = np.where(P[:, 1] < 8 * L0 - 1e-12, 1, 0).astype(np.int32)
    pr = 2000.0 + 3.0 * np.arange(len(t))
    o = Oracle(p, t, x, press=pr)
    o.call("update_buckets")
    o.call("volume_fraction")
    o.call("viscosity_interaction")
    q = o.params
    cv, ps, prh, phase = o.get("cv"), o.get("p_smooth"), o.get("p_rheo"), o.get("ptype")
    col = ((P[:, 0] - q.domain_min[0]) / (2.0 * q.part_dist)).astype(int) + 1
    checked = 0
    for c in np.unique(col):
        members = np.flatnonzero((col == c) & (cv > 0.5))
        if len(members) == 0:
            continue
        top = P[members, 1].max()
        k = members[P[members, 1] == top].min()
        for i in np.flatnonzero((col == c) & (phase == 2) & (cv >= 0.5)):
            ns = ps[i] - pr[k]
            expect = 1.0 if ns < 1.0 else ns
            assert prh[i] == expect, (i, prh[i], expect)
            checked += 1
    assert checked > 40


@pytest.mark.parametrize("slip", [0, 1])
def test_wall_strain_rate_ghost_sum(built, slip):
    """Fluid over the floor z = 0: the wall strain rate of a near-wall
    particle i is -coeff_grad sum_g (u_g - u_i) (x) (x_g - x_i) w / r^2 over
    the ghosts g of the real particles j != i that lie within reS of i, whose
    ghost lies within reS of i and farther than j; its symmetric part
    replaces S, its invariant adds to II and sum_g p_j w / n0 to the
    smoothed pressure."""
    p, t, x, mesh = floor_case(l0=L0, jitter=0.1, slip=slip, gravity=-9.81)
    granular(p)
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, 0.0, -9.81
    P = x.reshape(-1, 3)
    rng = np.random.default_rng(9)
    v = rng.normal(size=P.shape) * 0.1
    pr = 100.0 + rng.random(len(t)) * 50.0
    o = Oracle(p, t, x, vel=v.reshape(-1), press=pr, mesh=mesh)
    for op in ["update_buckets", "closest_point", "update_near_mesh", "volume_fraction", "viscosity_interaction"]:
        o.call(op)
    ii0, ps0 = o.get("ii"), o.get("p_smooth")
    o.call("wall_viscosity_interaction")
    q = o.params
    near = o.get("near_wall") == 1
    G = reflect_floor(P)
    R = np.diag([1.0, 1.0, -1.0])
    ug = v @ R.T if slip else -v
    S = {f: o.get(f) for f in ("s11", "s12", "s13", "s22", "s23", "s33", "ii", "p_smooth")}
    checked = 0
    for i in np.flatnonzero(near)[::5]:
        dr = np.sqrt(((P - P[i]) ** 2).sum(1))
        dg = G - P[i]
        rg = np.sqrt((dg ** 2).sum(1))
        m = (dr < q.re_s) & (rg < q.re_s) & (dr < rg)
        m[i] = False
        w = w0(rg[m], q.re_s)
        du = ug[m] - v[i]
        U = -q.coeff_press_grad * ((du[:, :, None] * dg[m][:, None, :]) * (w / rg[m] ** 2)[:, None, None]).sum(0)
        E = 0.5 * (U + U.T)
        scale = max(np.abs(E).max(), 1e-12)
        for (a, b), f in zip([(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)],
                             ("s11", "s12", "s13", "s22", "s23", "s33")):
            assert abs(S[f][i] - E[a, b]) <= 1e-11 * scale, (i, f, S[f][i], E[a, b])
        inv = 0.5 * (E[0, 0] ** 2 + 2 * E[0, 1] ** 2 + 2 * E[0, 2] ** 2 + E[1, 1] ** 2 + 2 * E[1, 2] ** 2 + E[2, 2] ** 2)
        assert abs(S["ii"][i] - (ii0[i] + inv)) <= 1e-11 * max(ii0[i] + inv, 1e-12)
        assert abs(S["p_smooth"][i] - (ps0[i] + (pr[m] * w).sum() / q.pnd_gradient_zero)) <= 1e-11 * ps0[i]
        checked += 1
    assert checked > 20
