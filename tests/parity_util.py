"""Lockstep comparison of the HIP path with the CPU oracle.

Both sides run the reference's per-step operator sequence
(src/main.cpp:252-467).  After every operator the fields it writes are
compared bit for bit.  The one place where the two may legitimately differ
is the PPE solve (its dot products reduce in a different order): the solution
is compared within a tolerance and then the oracle's solution is copied into
the device state so that the following operators again start from identical
inputs.
"""
from __future__ import annotations

import numpy as np

from polymps_amd.lib import Simulation
from oracle.oracle import Oracle

PERSISTENT = ["pos", "vel", "press", "ppe_x", "pndi", "bc", "num_neigh", "mirror", "force_wall", "press_avg",
              "pndki", "pndski", "div_u", "diffusive", "grad_conc"]
# carried as well with the granular phase (its density, viscosity and phase
# per particle, and the correction matrix its strain rate reads next step)
PERSISTENT_GRANULAR = ["rho", "meu", "ptype", "cm1", "cm2", "cm3"]


class Mismatch(AssertionError):
    pass


def compare_exact(name, g, o, where=""):
    if g.dtype.kind == "f":
        same = (g == o) | (np.isnan(g) & np.isnan(o))
    else:
        same = g == o
    if not np.all(same):
        bad = np.flatnonzero(~same)
        k = bad[0]
        raise Mismatch(f"{where}: field {name} differs at {len(bad)} entries; first index {k}: "
                       f"gpu {g[k]!r} oracle {o[k]!r}")


def compare_close(name, g, o, rtol, atol, where=""):
    if not np.allclose(g, o, rtol=rtol, atol=atol):
        d = np.abs(g - o)
        k = int(np.argmax(d))
        raise Mismatch(f"{where}: field {name} max |diff| {d[k]:.3e} at {k}: gpu {g[k]!r} oracle {o[k]!r}")


def sync_state(g: Simulation, o: Oracle, fields=PERSISTENT):
    if g.params.fluid_type == 1 and fields is PERSISTENT:
        fields = PERSISTENT + PERSISTENT_GRANULAR
    for f in fields:
        g.set(f, o.get(f))
    g.set_scalar("vel_max", o.scalar("vel_max"))
    g.set_scalar("time", o.scalar("time"))   # the granular normal stress of 3-D cases ramps with time


def new_pair(w, threads=0):
    extra = getattr(w, "extra_meshes", None)
    g = Simulation(w.params, w.types, w.pos, w.vel, mesh=w.mesh, meshes=extra)
    o = Oracle(w.params, w.types, w.pos, w.vel, mesh=w.mesh, threads=threads, meshes=extra)
    return g, o


# Fields carried with a particle through the device's bucket sort (its
# persistent state).  The per-step fields of a removed (ghost) particle keep
# their last values in the reference, which no operator or output reads
# again (every loop stops at numParticles, e.g. src/MpsInputOutput.cpp:1339);
# on the device they belong to the slot, so past numParticles only these
# fields are compared.
GHOST_DEFINED = set(PERSISTENT) | {"type"}


def check(g, o, fields, where):
    nl = int(o.scalar("num_particles"))
    for f in fields:
        a, b = g.get(f), o.get(f)
        if nl < o.n and f not in GHOST_DEFINED:
            k = nl * (len(b) // o.n)
            a, b = a[:k], b[:k]
        compare_exact(f, a, b, where)


def ppe_check(g, o, where, press_rtol):
    """Compare the assembled iteration matrix and right-hand side exactly,
    then the solution within a tolerance relative to the largest pressure."""
    p = g.params
    rpg, cg, vg = g.ppe_matrix()
    rpo, co, vo = o.ppe_matrix()
    n = len(rpo) - 1
    # oracle rows are A; the device's CG iterates A^T when ppe_transpose is
    # set, its BiCGSTAB A
    trans = bool(p.ppe_transpose) and p.solver_type == 0
    ent_o = {}
    for r in range(n):
        for k in range(rpo[r], rpo[r + 1]):
            key = (int(co[k]), r) if trans else (r, int(co[k]))
            ent_o[key] = vo[k]
    ent_g = {}
    for r in range(len(rpg) - 1):
        for k in range(rpg[r], rpg[r + 1]):
            ent_g[(r, int(cg[k]))] = vg[k]
    if ent_g.keys() != ent_o.keys():
        only_g = sorted(set(ent_g) - set(ent_o))[:5]
        only_o = sorted(set(ent_o) - set(ent_g))[:5]
        raise Mismatch(f"{where}: PPE sparsity differs; gpu-only {only_g} oracle-only {only_o}")
    for key, v in ent_o.items():
        if ent_g[key] != v:
            raise Mismatch(f"{where}: PPE entry {key} gpu {ent_g[key]!r} oracle {v!r}")
    if not trans:
        # without the transpose the rows must also come in the same order
        compare_exact("ppe_cols", cg, co, where)
    compare_exact("ppe_b", g.get("ppe_b"), o.get("ppe_b"), where)
    # the entry count of the assembly (off-diagonal entries; the download adds the diagonal)
    n_off = int(rpg[-1]) - (len(rpg) - 1)
    if g.scalar("nnz") != n_off:
        raise Mismatch(f"{where}: nnz {g.scalar('nnz')} != {n_off} off-diagonal entries")


def compare_buckets(g: Simulation, o: Oracle, where=""):
    """Bucket lists bit for bit: first / next of every bucket and particle,
    and last, which ends the chain that starts at first
    (src/MpsBucket.cpp:331-341; periodic ghost layers :356-417)."""
    fg, lg, ng = g.buckets()
    fo, no = o.buckets()
    compare_exact("first_particle_in_bucket", fg, fo, where)
    compare_exact("next_particle_in_same_bucket", ng, no, where)
    lo = np.full_like(fo, -1)
    for b in np.flatnonzero(fo >= 0):
        j = int(fo[b])
        while no[j] != -1:
            j = int(no[j])
        lo[b] = j
    compare_exact("last_particle_in_bucket", lg, lo, where)


# granular phase: fields without transcendental functions (bit for bit), and
# the viscosities, whose sin / cos / tan / exp / pow come from the device
# library on one side and glibc on the other
GRANULAR_EXACT = ["ii", "s11", "s12", "s13", "s22", "s23", "s33", "p_smooth", "p_rheo", "inertia", "vf", "rho"]
GRANULAR_LIBM = ["meu", "meu_y"]
GRANULAR_LIBM_RTOL = 1e-12


def granular_check(g: Simulation, o: Oracle, ops, where):
    """Run the granular operators on both sides, compare, then hand the
    oracle's viscosities to the device so that the next operators start
    from identical inputs."""
    g.call("volume_fraction")
    o.call("volume_fraction")
    check(g, o, ["cv"], f"{where} volume_fraction")
    for op in ops:
        g.call(op)
        o.call(op)
        check(g, o, GRANULAR_EXACT, f"{where} {op}")
        nl = int(o.scalar("num_particles"))
        for f in GRANULAR_LIBM:
            a, b = g.get(f)[:nl], o.get(f)[:nl]
            compare_close(f, a, b, GRANULAR_LIBM_RTOL, 0.0, f"{where} {op}")
        for f in GRANULAR_LIBM:
            g.set(f, o.get(f))


def lockstep_step(g: Simulation, o: Oracle, step: int, press_rtol=1e-6, log=None, time=None):
    p = g.params
    W = f"step {step}"
    poly = p.wall_type == 1

    def both(gop, oops, fields):
        g.call(gop)
        for op in oops:
            o.call(op)
        check(g, o, fields, f"{W} {gop}")
        if log is not None:
            log.append(gop)

    both("bucket_update", ["update_buckets"], [])
    compare_buckets(g, o, f"{W} bucket_update")
    if p.fluid_type == 1:   # src/main.cpp:259-262
        granular_check(g, o, ["viscosity_interaction"], f"{W} granular")
    both("viscosity", ["viscosity"], ["acc"])
    both("add_gravity", ["add_gravity"], ["acc", "acc_star"])
    if p.relax_press < 1.0:
        both("prediction_press_gradient", ["prediction_press_gradient"], ["acc"])
        if poly:
            both("prediction_wall_press_gradient", ["prediction_wall_press_gradient"], ["acc"])
    both("prediction", ["prediction"], ["pos", "vel", "acc", "riw2", "npcd"])
    both("check_out_domain", ["check_out_domain"], ["type", "pos"])
    if p.collision_type == 0:
        g.call("build_neighbor_list")
        both("collision", ["collision"], ["vel"])
    else:
        both("collision", ["collision"], ["vel", "pos"])
        g.call("build_neighbor_list")
    if poly:
        both("mesh_closest_point", ["closest_point"], ["riw2", "mirror", "wall_pos", "polygon_normal", "near_mesh_type"])
        both("mesh_update_near", ["update_near_mesh"], ["pnd_wall", "num_neigh_wall", "near_wall"])
        both("wall_npcd", ["wall_npcd"], ["npcd"])
    if p.gradient_correction:
        both("correction_matrix", ["correction_matrix"], ["cm1", "cm2", "cm3"])
    both("pnd_neigh_npcd", ["pnd_neigh_npcd"], ["num_neigh", "pnd_small", "pndi", "pndki", "npcd", "npcd2"])
    if p.pnd_type == 2:
        ops = ["pnd_diffusive"] + (["wall_pnd_diffusive"] if poly else []) + ["update_pnd", "mean_pnd_wall_dummy_surface"]
        both("pnd_diffusive", ops, ["pndi", "div_u", "diffusive"])
    elif p.pnd_type == 1:
        both("mean_pnd", (["mean_wall_pnd"] if poly else []) + ["mean_pnd"], ["pndi"])
    if p.free_surf_type == 2:
        both("normal_particles", ["normal_particles"], ["normal"])
        if poly:
            both("wall_normal_particles", ["wall_normal_particles"], ["normal"])
    both("update_bc", ["update_bc"], ["bc"])
    if p.mps_type in (0, 1):
        g.call("calc_press")
        o.call("calc_press")
        compare_close("press", g.get("press"), o.get("press"), 1e-13, 1e-9, f"{W} calc_press")
        g.set("press", o.get("press"))
    else:
        if p.mps_type == 3:
            both("vel_divergence", ["vel_divergence"] + (["wall_vel_divergence"] if poly else []), ["div_u"])
        g.call("ppe_assemble")
        o.call("assemble_ppe")
        ppe_check(g, o, f"{W} ppe_assemble", press_rtol)
        g.call("ppe_solve")
        o.call("solve_ppe")
        pg, po = g.get("press"), o.get("press")
        scale = max(np.abs(po).max(), 1e-30)
        compare_close("press", pg, po, 0.0, press_rtol * scale, f"{W} ppe_solve")
        g.set("ppe_x", o.get("ppe_x"))
        g.set("press", po)
    if not poly:
        both("extrapolate_wall_dummy", ["extrapolate_wall_dummy"], ["press"])
    if p.gradient_correction:
        both("correction_matrix", ["correction_matrix"], ["cm1", "cm2", "cm3"])
    both("press_gradient", ["press_gradient"], ["acc"])
    if poly:
        both("wall_press_gradient", ["wall_press_gradient"], ["acc", "force_wall"])
    if p.fluid_type == 1:   # src/main.cpp:387-398
        granular_check(g, o, ["viscosity_interaction"] + (["wall_viscosity_interaction"] if poly else []),
                       f"{W} granular (2)")
    if poly:
        both("wall_viscosity", ["wall_viscosity"], ["acc", "force_wall"])
    if p.forced_on:
        t = step * p.time_step if time is None else time
        g.forced_mesh_update(p.time_step, t)
        o.forced_mesh_update(p.time_step, t)
        for a, b, name in zip(g.get_mesh(2), o.get_mesh(2), ("nodes", "faces", "normals")):
            if not np.array_equal(a, b):
                raise Mismatch(f"{W}: forced mesh {name} differ")
    both("correction", ["correction"], ["pos", "vel", "acc"])
    vg, vo = g.scalar("vel_max"), o.scalar("vel_max")
    if vg != vo:
        raise Mismatch(f"{W}: vel_max gpu {vg!r} oracle {vo!r}")
    both("check_out_domain", ["check_out_domain"], ["type", "pos"])
    if p.shifting_type:
        both("shifting", ["shifting"], ["pos", "vel"])
    if not poly:
        both("wall_dummy_velocity", ["wall_dummy_velocity"], ["vel"])
    both("check_out_domain", ["check_out_domain"], ["type", "pos"])


def check_list_max(g: Simulation):
    """The longest row of a list built at the current positions (no skin)
    is the largest count of other live particles within the large radius
    (brute force over all pairs; periodic domains excluded)."""
    p = g.params
    if p.domain_type_bc == 1:
        return
    n = int(g.scalar("num_particles"))
    x = g.get("pos").reshape(-1, 3)[:n]
    best = 0
    for s0 in range(0, n, 2048):
        d2 = ((x[s0:s0 + 2048, None, :] - x[None, :, :]) ** 2).sum(-1)
        best = max(best, int(((d2 < p.re_l * p.re_l).sum(1) - 1).max()))
    got = g.scalar("list_max")
    if got != best:
        raise Mismatch(f"list_max {got} != brute force {best}")


def lockstep_init(g: Simulation, o: Oracle):
    """The initial-step pipeline of main() (src/main.cpp:119-180) on both
    sides, compared after each operator, then two lockstep steps."""
    p = g.params
    poly = p.wall_type == 1
    g.call("bucket_update"); o.call("update_buckets")
    g.call("check_out_domain"); o.call("check_out_domain")
    if poly:
        g.call("mesh_closest_point"); o.call("closest_point")
        check(g, o, ["riw2", "mirror", "wall_pos", "polygon_normal"], "init closest_point")
        g.call("mesh_update_near"); o.call("update_near_mesh")
        check(g, o, ["pnd_wall", "num_neigh_wall", "near_wall"], "init update_near")
    g.call("build_neighbor_list")
    check_list_max(g)
    if poly:
        g.call("wall_npcd"); o.call("wall_npcd")
        check(g, o, ["npcd"], "init wall_npcd")
    g.call("initial_pnd"); o.call("initial_pnd")
    check(g, o, ["pndi", "pnd_small", "pndki", "num_neigh", "npcd", "npcd2"], "init initial_pnd")
    g.call("mean_pnd")
    if poly:
        o.call("mean_wall_pnd")
    o.call("mean_pnd")
    check(g, o, ["pndi"], "init mean_pnd")
    g.call("mean_neigh_fluid_pnd"); o.call("mean_neigh_fluid_pnd")
    check(g, o, ["pndski"], "init mean_neigh_fluid_pnd")
    if p.free_surf_type == 2:
        g.call("normal_particles"); o.call("normal_particles")
        if poly:
            g.call("wall_normal_particles"); o.call("wall_normal_particles")
        check(g, o, ["normal"], "init normals")
    g.call("update_bc"); o.call("update_bc")
    check(g, o, ["bc"], "init update_bc")
    if p.mps_type >= 2:
        if p.mps_type == 3:
            g.call("vel_divergence"); o.call("vel_divergence")
            if poly:
                o.call("wall_vel_divergence")
            check(g, o, ["div_u"], "init vel_divergence")
        g.call("ppe_assemble"); o.call("assemble_ppe")
        ppe_check(g, o, "init ppe", 1e-6)
        g.call("ppe_solve"); o.call("solve_ppe")
        po = o.get("press")
        compare_close("press", g.get("press"), po, 0.0, 1e-6 * max(np.abs(po).max(), 1e-30), "init ppe_solve")
        g.set("ppe_x", o.get("ppe_x"))
        g.set("press", po)
    else:
        g.call("calc_press"); o.call("calc_press")
        compare_close("press", g.get("press"), o.get("press"), 1e-13, 1e-9, "init calc_press")
        g.set("press", o.get("press"))
