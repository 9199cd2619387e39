/*
 * mps_oracle.cpp -- CPU ORACLE (test infrastructure, not product code).
 *
 * A C++ restatement of the reference's per-step MPS algorithm, written to
 * follow the reference's arithmetic order expression by expression so that it
 * can serve as the parity checker for the HIP path.  Only tests/, the smoke
 * check of __graft_entry__ and the cpu_baseline leg of bench.py may load it.
 *
 * Data model follows the reference: AoS double pos[3N], vel[3N], ...; the
 * bucket list is the reference's first/last/next linked list built in
 * ascending particle order (src/MpsBucket.cpp:299-353).  Every neighbour loop
 * visits the 3x3(x3) stencil jz -> jy -> jx ascending and walks each bucket in
 * ascending particle index (e.g. src/MpsPndNeigh.cpp:209-258).
 *
 * Deliberate, documented differences to the reference:
 *  - stencil cells outside [0,n) on any axis are skipped; the reference
 *    indexes them linearly (wrapping to a far row, or reading before the
 *    array).  Wrapped cells lie more than reL away, so results agree whenever
 *    the reference stays inside its arrays.
 *  - periodic domains: one periodic direction with particle walls, the
 *    setting MpsBucket::calcDomainLimits handles (src/MpsBucket.cpp:141-287).
 *  - Eigen's CG is restated from its published algorithm (Jacobi
 *    preconditioner, relative residual test on the recursive residual,
 *    default tol = DBL_EPSILON, default max iterations 2N) with sequential
 *    dot products; Eigen's vectorised reductions are not reproduced, so the
 *    iterate sequence itself is not pinned.
 *  - libigl's AABB closest point is replaced by brute force over all faces
 *    using Ericson's point-triangle closest point (RTCD 5.1.5); ties keep the
 *    lowest face index.
 */
#include "../include/mps_params.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Mesh {
    int nv = 0, nf = 0;
    std::vector<double> V;    // [nv*3]
    std::vector<int> F;       // [nf*3]
    std::vector<double> Nrm;  // [nf*3]
    std::vector<double> xPND, xNeigh, Zij;
    std::vector<int> nNeighWall;
};

struct Sys {
    mps_params p;
    int N = 0, N0 = 0;
    // scalars per particle
    std::vector<int> type, bc, numNeigh, ptype, nearMeshType, elementID, numNeighWall;
    std::vector<char> nearWall;
    std::vector<double> press, pressAvg, pndi, pndki, pndski, pndSmall, npcd2, conc,
        divU, diffTerm, aux1, aux2, pndWall, devDotN, nNeighSurf, riw2, RHO, MEU;
    // non-Newtonian (granular) phase, fluid_type 1: volume fraction, second
    // invariant of the strain rate, yield viscosity, inertial number, normal
    // stress, smoothed pressure, grain volume fraction, strain-rate tensor
    std::vector<double> Cv, II, MEUY, inertia, pRheo, pSmooth, VF, S11, S12, S13, S22, S23, S33;
    double cflVisc = 0.0;
    // vectors (AoS, 3 per particle)
    std::vector<double> pos, vel, acc, accStar, npcd, gradC, cm1, cm2, cm3, normal, dvel,
        wallPos, mirror, wpf1, wpf2, polyN, forceWall;
    // PPE
    std::vector<double> ppeX, ppeB;
    std::vector<int> rowPtr, colIdx;
    std::vector<double> vals;
    int solverIter = 0;
    // CG with OpenMP dot products and vector updates (the CPU baseline's
    // parallel CSR Jacobi-PCG); off: serial, in index order (the reference's
    // Eigen CG is serial, src/MpsPressure.cpp:424-445)
    bool parallelCg = false;
    // wall time of the last CG solve: building the CSR iteration matrix, then the iterations
    double solveSetupSeconds = 0.0, solveIterSeconds = 0.0;
    // CG iteration matrix (CSR) of assembly number itAsm, transposed or not
    long asmCount = 0, itAsm = -1;
    int itTrans = -1;
    std::vector<int> itRp, itCi;
    std::vector<double> itVa, itDiag;
    double solverErr = 0.0;
    // buckets
    std::vector<int> first, last, next;
    std::vector<int> bucketPBC;   // bucketPeriodicBC: -1 / +1 on the ghost layers, 0 elsewhere
    Mesh mesh;
    bool hasMesh = false;
    Mesh extra[2];            // meshType 1 (deformable, static) and 2 (forced)
    bool extraOn[2] = {false, false};
    double velMax = 0.0, cflCurrent = 0.0, timeCurrent = 0.0;
    int numIter = 0;
    int nThreads = 0;
    std::string err;
};

const double kEps = 10.0 * DBL_EPSILON;

// ---------------------------------------------------------------- kernels
// src/MpsParticle.cpp:57-72 (weight), :75-90 (weightGradient, identical forms)
inline double weight(double dst, double re, int t) {
    switch (t) {
        case 0: return re / dst - 1.0;
        case 1: return re / dst + dst / re - 2.0;
        case 2: return re / dst - dst / re;
        case 3: return (1.0 - dst / re) * (1.0 - dst / re) * (1.0 - dst / re);
        case 4: return (1.0 - dst / re) * (1.0 - dst / re);
        default: return re / dst - 1.0;
    }
}
// src/MpsParticle.cpp:93-108
inline double delWeight(double dst, double re, int t) {
    switch (t) {
        case 0: return -re / (dst * dst);
        case 1: return -re / (dst * dst) + 1.0 / re;
        case 2: return -re / (dst * dst) - 1.0 / re;
        case 3: return -3.0 / re * (1.0 - dst / re) * (1.0 - dst / re);
        case 4: return -2.0 / re * (1.0 - dst / re);
        default: return -re / (dst * dst);
    }
}

// src/MpsBucket.cpp:420-425 (stencil centre, +eps on every axis)
inline void bucketCoords(const Sys& s, double x, double y, double z, int& bx, int& by, int& bz) {
    bx = (int)((x - s.p.domain_min[0]) * s.p.inv_bucket_side + kEps);
    by = (int)((y - s.p.domain_min[1]) * s.p.inv_bucket_side + kEps);
    bz = (int)((z - s.p.domain_min[2]) * s.p.inv_bucket_side + kEps);
}

// Periodic shift of the bucket being walked (MpsParticle::getPeriodicLengths,
// src/MpsParticle.cpp:48-54): every neighbour position inside a walk is
// pos_j + pl, as sqrDistBetweenParticles (src/MpsParticle.cpp:35-45) forms it.
thread_local double tlPl[3] = {0.0, 0.0, 0.0};
inline double PJ(const Sys& s, int j, int k) { return s.pos[j * 3 + k] + tlPl[k]; }

template <class Fn>
inline void walk(const Sys& s, double xi, double yi, double zi, Fn&& fn) {
    int ix, iy, iz;
    bucketCoords(s, xi, yi, zi, ix, iy, iz);
    const int d3 = (int)(s.p.dim - 2.0);
    const int minZ = (iz - 1) * d3, maxZ = (iz + 1) * d3;
    const int nx = s.p.num_buckets_x, ny = s.p.num_buckets_y, nz = s.p.num_buckets_z;
    for (int jz = minZ; jz <= maxZ; jz++) {
        if (jz < 0 || jz >= nz) continue;
        for (int jy = iy - 1; jy <= iy + 1; jy++) {
            if (jy < 0 || jy >= ny) continue;
            for (int jx = ix - 1; jx <= ix + 1; jx++) {
                if (jx < 0 || jx >= nx) continue;
                const int jb = jz * s.p.num_buckets_xy + jy * nx + jx;
                const int bPBC = s.bucketPBC.empty() ? 0 : s.bucketPBC[jb];
                for (int k = 0; k < 3; k++) tlPl[k] = bPBC * s.p.periodic_length[k];
                int j = s.first[jb];
                while (j != -1) {
                    fn(j);
                    j = s.next[j];
                }
            }
        }
    }
    tlPl[0] = tlPl[1] = tlPl[2] = 0.0;
}

inline bool partWall(const Sys& s) { return s.p.wall_type == MPS_WALL_PARTICLE; }

// Rref = I - 2 n n^T with n = normalised 0.5 (x - xm); shared by the wall
// operators (e.g. src/MpsPndNeigh.cpp:116-136).  Returns |n| before
// normalisation in *nsqrt.
inline void reflection(const double* xi, const double* xm, double* n, double* R, double* nsqrt) {
    n[0] = 0.5 * (xi[0] - xm[0]);
    n[1] = 0.5 * (xi[1] - xm[1]);
    n[2] = 0.5 * (xi[2] - xm[2]);
    double ns = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (ns > kEps) {
        n[0] = n[0] / ns; n[1] = n[1] / ns; n[2] = n[2] / ns;
    } else {
        n[0] = 0; n[1] = 0; n[2] = 0;
    }
    R[0] = 1.0 - 2.0 * n[0] * n[0]; R[1] = 0.0 - 2.0 * n[0] * n[1]; R[2] = 0.0 - 2.0 * n[0] * n[2];
    R[3] = 0.0 - 2.0 * n[1] * n[0]; R[4] = 1.0 - 2.0 * n[1] * n[1]; R[5] = 0.0 - 2.0 * n[1] * n[2];
    R[6] = 0.0 - 2.0 * n[2] * n[0]; R[7] = 0.0 - 2.0 * n[2] * n[1]; R[8] = 1.0 - 2.0 * n[2] * n[2];
    if (nsqrt) *nsqrt = ns;
}
// variant using |n|^2 > eps as the viscosity operators do (src/MpsViscosity.cpp:1549-1562)
inline void reflectionMod2(const double* xi, const double* xm, double* n, double* R) {
    n[0] = 0.5 * (xi[0] - xm[0]);
    n[1] = 0.5 * (xi[1] - xm[1]);
    n[2] = 0.5 * (xi[2] - xm[2]);
    double m2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (m2 > kEps) {
        double m = sqrt(m2);
        n[0] = n[0] / m; n[1] = n[1] / m; n[2] = n[2] / m;
    } else {
        n[0] = 0; n[1] = 0; n[2] = 0;
    }
    R[0] = 1.0 - 2.0 * n[0] * n[0]; R[1] = 0.0 - 2.0 * n[0] * n[1]; R[2] = 0.0 - 2.0 * n[0] * n[2];
    R[3] = 0.0 - 2.0 * n[1] * n[0]; R[4] = 1.0 - 2.0 * n[1] * n[1]; R[5] = 0.0 - 2.0 * n[1] * n[2];
    R[6] = 0.0 - 2.0 * n[2] * n[0]; R[7] = 0.0 - 2.0 * n[2] * n[1]; R[8] = 1.0 - 2.0 * n[2] * n[2];
}

// ------------------------------------------------------------ parameters
// MpsParticle::setParameters (src/MpsParticle.cpp:136-227) followed by
// MpsBucket::allocateMemory / calcDomainLimits (src/MpsBucket.cpp:20-295).
int deriveParams(Sys& s) {
    mps_params& P = s.p;
    if (P.derived) return 0;
    const double l0 = P.part_dist;
    P.eps_zero = kEps;
    for (int k = 0; k < 3; k++) {
        P.domain_min[k] -= l0 * 3.0;
        P.domain_max[k] += l0 * 3.0;
    }
    if ((int)P.dim == 2) { P.domain_min[2] = 0.0; P.domain_max[2] = 0.0; }
    P.num_of_meshs = (P.wall_type == 1 ? 1 : 0) + (P.fem_on ? 1 : 0) + (P.forced_on ? 1 : 0);
    P.re_s = l0 * P.re_s;
    P.re_l = l0 * P.re_l;
    P.re_repulsive_force = l0 * P.re_repulsive_force;
    P.re_s2 = P.re_s * P.re_s;
    P.re_l2 = P.re_l * P.re_l;
    P.eps_re = P.eps_re * P.re_s2 / 4.0;
    double n0S = 0, n0L = 0, n0G = 0, lam = 0, nn0 = 0;
    int lmin = (int)ceil(P.re_l / l0) + 1;
    int lmax = (int)ceil(P.re_l / l0) + 2;
    int f2 = 0, f3 = 1;
    if ((int)P.dim == 2) { f2 = 1; f3 = 0; }
    for (int ix = -lmin; ix < lmax; ix++)
        for (int iy = -lmin; iy < lmax; iy++)
            for (int iz = -lmin * f3; iz < lmax * f3 + f2; iz++) {
                double x = l0 * (double)ix, y = l0 * (double)iy, z = l0 * (double)iz;
                double d2 = x * x + y * y + z * z;
                if (d2 <= P.re_l2) {
                    if (d2 <= kEps) continue;
                    double d = sqrt(d2);
                    n0L += weight(d, P.re_l, P.weight_type);
                    lam += d2 * weight(d, P.re_l, P.weight_type);
                    nn0 += 1;
                    if (d2 <= P.re_s2) {
                        n0S += weight(d, P.re_s, P.weight_type);
                        n0G += weight(d, P.re_s, P.weight_type);
                    }
                }
            }
    P.pnd_small_zero = n0S; P.pnd_large_zero = n0L; P.pnd_gradient_zero = n0G;
    P.num_neigh_zero = nn0;
    P.lambda_zero = lam / n0L;
    P.coeff_viscosity = 2.0 * P.knm_vs1 * P.dim / (n0L * P.lambda_zero);
    P.coeff_visc_multiphase = 2.0 * P.dim / (n0L * P.lambda_zero);
    P.coeff_press_emps = P.sound_speed * P.sound_speed / n0S;
    P.coeff_press_grad = -P.dim / n0G;
    P.coeff_press_wcmps = P.sound_speed * P.sound_speed;
    P.coeff_shifting1 = P.dri * l0 / n0S;
    P.coeff_shifting2 = P.coef_a * l0 * l0 * P.cfl_number * P.mach_number;
    P.coeff_ppe = 2.0 * P.dim / (n0L * P.lambda_zero);
    P.coeff_ppe_source = P.relax_pnd / (P.time_step * P.time_step * n0S);
    P.dns_fluid = P.density_fluid; P.dns_wall = P.density_wall;
    P.inv_dns_fluid = 1.0 / P.density_fluid; P.inv_dns_wall = 1.0 / P.density_wall;
    P.inv_part_dist = 1.0 / l0;
    P.dist_collision_limit = l0 * P.dist_limit_ratio;
    P.dist_collision_limit2 = P.dist_collision_limit * P.dist_collision_limit;
    P.restitution_collision = 1.0 + P.collision_ratio;
    P.beta_pnd = P.pnd_threshold * n0S;
    P.beta_neigh = P.neigh_threshold * nn0;
    P.delta2 = P.npcd_threshold * P.npcd_threshold * l0 * l0;
    P.theta_arc = P.theta_threshold / 180.0 * 3.14159265;
    P.h_threshold2 = 1.33 * 1.33 * l0 * l0;
    P.dst_threshold2 = 2.0 * P.h_threshold2;
    P.norm_threshold2 = P.norm_threshold * P.norm_threshold;

    // buckets
    P.bucket_side = P.re_l * (1.0 + P.cfl_number);
    P.inv_bucket_side = 1.0 / P.bucket_side;
    P.num_buckets_x = (int)((P.domain_max[0] - P.domain_min[0]) * P.inv_bucket_side) + 3;
    P.num_buckets_y = (int)((P.domain_max[1] - P.domain_min[1]) * P.inv_bucket_side) + 3;
    P.num_buckets_z = (int)((P.domain_max[2] - P.domain_min[2]) * P.inv_bucket_side) + 3;
    if ((int)P.dim == 2) P.num_buckets_z = 1;
    const bool periodic = P.domain_type_bc == 1;
    int pdir = -1;
    for (int k = 0; k < 3; k++) P.periodic_length[k] = 0.0;
    if (periodic) {
        int nper = 0;
        for (int k = 0; k < 3; k++) if (P.periodic_dir[k]) { nper++; pdir = k; }
        if (nper != 1) { s.err = "periodic: one direction"; return -1; }
    } else if (P.domain_type_bc != 0) {
        s.err = "unknown domain boundary type";
        return -1;
    }
    if (P.wall_type == MPS_WALL_PARTICLE || periodic) {
        // calcDomainLimits (src/MpsBucket.cpp:71-295, called for particle
        // walls or a periodic domain, :45-48): limitTypeBC, then the
        // non-periodic or the periodic branch
        double lim[3][2];
        if (P.limit_type_bc == 0) {
            for (int k = 0; k < 3; k++) lim[k][0] = lim[k][1] = s.pos[k];
            for (int i = 0; i < s.N; i++)
                for (int k = 0; k < 3; k++) {
                    lim[k][0] = std::min(lim[k][0], s.pos[i * 3 + k]);
                    lim[k][1] = std::max(lim[k][1], s.pos[i * 3 + k]);
                }
            int td = 0;
            for (int k = 0; k < 3; k++) if (lim[k][0] != lim[k][1]) td++;
            if (td != (int)P.dim) { s.err = "dimension mismatch between json and particles"; return -1; }
        } else {
            for (int k = 0; k < 3; k++) { lim[k][0] = P.domain_min[k]; lim[k][1] = P.domain_max[k]; }
        }
        auto setNb = [&](int k, long nb) {
            if (k == 0) P.num_buckets_x = (int)nb;
            if (k == 1) P.num_buckets_y = (int)nb;
            if (k == 2) P.num_buckets_z = (int)nb;
        };
        if (!periodic) {
            lim[0][0] -= 0.5 * l0; lim[0][1] += 0.5 * l0;
            lim[1][0] -= 0.5 * l0; lim[1][1] += 0.5 * l0;
            if (P.dim == 3) { lim[2][0] -= 0.5 * l0; lim[2][1] += 0.5 * l0; }
            for (int k = 0; k < (int)P.dim; k++) {
                lim[k][0] -= P.bucket_side;
                lim[k][1] += P.bucket_side;
                long nb = (long)((lim[k][1] - lim[k][0]) / P.bucket_side + 1);
                setNb(k, nb);
                lim[k][1] = lim[k][0] + nb * P.bucket_side;
            }
        } else {
            // e.g. periodic in x, src/MpsBucket.cpp:148-190
            const int d = pdir;
            P.periodic_length[d] = lim[d][1] - lim[d][0] + P.part_dist;
            lim[d][0] -= P.part_dist * 0.5;
            lim[d][1] += P.part_dist * 0.5;
            P.phys_min[d] = lim[d][0];
            P.phys_max[d] = lim[d][1];
            long nbd = (long)((lim[d][1] - lim[d][0]) / P.bucket_side);
            P.bucket_side = P.periodic_length[d] / (long)(P.periodic_length[d] / P.bucket_side);
            lim[d][0] -= P.bucket_side;
            lim[d][1] += P.bucket_side;
            nbd += 2;
            setNb(d, nbd);
            for (int k = 0; k < 3; k++) {
                if (k == d || (k == 2 && P.dim != 3)) continue;
                // :164-171: particle walls one bucket beyond the particles,
                // polygon walls one bucket beyond the JSON domain
                if (P.wall_type == MPS_WALL_PARTICLE) {
                    lim[k][0] -= P.bucket_side;
                    lim[k][1] += P.bucket_side;
                } else {
                    lim[k][0] = P.domain_min[k] - P.bucket_side;
                    lim[k][1] = P.domain_max[k] + P.bucket_side;
                }
                P.phys_min[k] = lim[k][0];
                P.phys_max[k] = lim[k][1];
                long nb = (long)((lim[k][1] - lim[k][0]) / P.bucket_side + 1);
                setNb(k, nb);
                lim[k][1] = lim[k][0] + nb * P.bucket_side;
            }
        }
        for (int k = 0; k < 3; k++) { P.domain_min[k] = lim[k][0]; P.domain_max[k] = lim[k][1]; }
        P.inv_bucket_side = 1.0 / P.bucket_side;
    }
    P.num_buckets_xy = P.num_buckets_x * P.num_buckets_y;
    P.num_buckets_xyz = P.num_buckets_xy * P.num_buckets_z;
    // MpsBoundaryCondition::setBucketBC (src/MpsBoundaryCondition.cpp:20-63)
    s.bucketPBC.assign(P.num_buckets_xyz, 0);
    if (periodic)
        for (int iz = 0; iz < P.num_buckets_z; iz++)
            for (int iy = 0; iy < P.num_buckets_y; iy++)
                for (int ix = 0; ix < P.num_buckets_x; ix++) {
                    const int ib = iz * P.num_buckets_xy + iy * P.num_buckets_x + ix;
                    const int c[3] = {ix, iy, iz};
                    const int n[3] = {P.num_buckets_x, P.num_buckets_y, P.num_buckets_z};
                    if (c[pdir] == 0) s.bucketPBC[ib] = -1;
                    if (c[pdir] == n[pdir] - 1) s.bucketPBC[ib] = 1;
                }
    P.derived = 1;
    return 0;
}

// ---------------------------------------------------------- MpsBucket
// MpsBucket::updateParticlesID (src/MpsBucket.cpp:299-353)
void updateParticlesID(Sys& s) {
    const mps_params& P = s.p;
    s.first.assign(P.num_buckets_xyz, -1);
    s.last.assign(P.num_buckets_xyz, -1);
    s.next.assign(s.N0, -1);
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] == P.ghost) continue;
        int ib;
        if ((int)P.dim == 2) {
            int ix = (int)((s.pos[i * 3] - P.domain_min[0]) * P.inv_bucket_side + kEps);
            int iy = (int)((s.pos[i * 3 + 1] - P.domain_min[1]) * P.inv_bucket_side + kEps);
            ib = iy * P.num_buckets_x + ix;
        } else {
            int ix = (int)((s.pos[i * 3] - P.domain_min[0]) * P.inv_bucket_side);
            int iy = (int)((s.pos[i * 3 + 1] - P.domain_min[1]) * P.inv_bucket_side);
            int iz = (int)((s.pos[i * 3 + 2] - P.domain_min[2]) * P.inv_bucket_side);
            ib = iz * P.num_buckets_xy + iy * P.num_buckets_x + ix;
        }
        if (ib < 0 || ib >= P.num_buckets_xyz) continue;  // outside the grid (UB in the reference)
        int j = s.last[ib];
        s.last[ib] = i;
        if (j == -1) s.first[ib] = i; else s.next[j] = i;
    }
    // MpsBucket::copyDataBetweenBuckets (src/MpsBucket.cpp:356-417): the ghost
    // layer of the periodic direction lists the particles of the opposite
    // interior layer
    if (P.domain_type_bc == 1) {
        const int off = P.periodic_dir[0] ? P.num_buckets_x - 2
                      : P.periodic_dir[1] ? P.num_buckets_x * (P.num_buckets_y - 2)
                                          : P.num_buckets_xy * (P.num_buckets_z - 2);
        for (int ib = 0; ib < P.num_buckets_xyz; ib++) {
            if (s.bucketPBC[ib] == -1) { s.first[ib] = s.first[ib + off]; s.last[ib] = s.last[ib + off]; }
            if (s.bucketPBC[ib] == 1) { s.first[ib] = s.first[ib - off]; s.last[ib] = s.last[ib - off]; }
        }
    }
}

// --------------------------------------------------------- MpsViscosity
// MpsViscosity::calcViscosity (src/MpsViscosity.cpp:20-102)
void calcViscosity(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        double meu_i = s.MEU[i];
        double ax = 0, ay = 0, az = 0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double ux = s.vel[i * 3], uy = s.vel[i * 3 + 1], uz = s.vel[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_l2 && (d2 < dm2 || pw)) {
                if (j != i) {
                    double dst = sqrt(d2);
                    double wL = weight(dst, P.re_l, P.weight_type);
                    double nu;
                    if ((meu_i + s.MEU[j]) > kEps) nu = 2.0 * meu_i * s.MEU[j] / (meu_i + s.MEU[j]);
                    else nu = 0.0;
                    if (s.type[j] == P.wall) nu = 2.0 * meu_i;
                    nu = nu / s.RHO[i];
                    ax += (s.vel[j * 3] - ux) * wL * nu;
                    ay += (s.vel[j * 3 + 1] - uy) * wL * nu;
                    az += (s.vel[j * 3 + 2] - uz) * wL * nu;
                }
            }
        });
        s.acc[i * 3] = P.coeff_visc_multiphase * ax;
        s.acc[i * 3 + 1] = P.coeff_visc_multiphase * ay;
        s.acc[i * 3 + 2] = P.coeff_visc_multiphase * az;
    }
}

// MpsViscosity::calcWallSlipViscosity / calcWallNoSlipViscosity
// (src/MpsViscosity.cpp:1527-1686, 1687-1910)
void calcWallViscosity(Sys& s) {
    const mps_params& P = s.p;
    const double vol = pow(P.part_dist, P.dim);
    const bool slip = P.slip_condition == MPS_FREE_SLIP;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (!(s.type[i] == P.fluid && s.nearWall[i])) continue;
        double meu_i = s.MEU[i];
        double ax = 0, ay = 0, az = 0;
        const double* xi = &s.pos[i * 3];
        const double* xm = &s.mirror[i * 3];
        const double ux = s.vel[i * 3], uy = s.vel[i * 3 + 1], uz = s.vel[i * 3 + 2];
        double n[3], R[9];
        reflectionMod2(xi, xm, n, R);
        double umx, umy, umz;
        double Rm[9];
        if (slip) {
            umx = R[0] * ux + R[1] * uy + R[2] * uz;
            umy = R[3] * ux + R[4] * uy + R[5] * uz;
            umz = R[6] * ux + R[7] * uy + R[8] * uz;
            for (int k = 0; k < 9; k++) Rm[k] = R[k];
        } else {
            double Ri[9] = {-1.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0, -1.0};
            double vw[3] = {0.0, 0.0, 0.0};   // wall velocity: uniformVelWall near the forced mesh
            if (s.nearMeshType[i] == 2) for (int k = 0; k < 3; k++) vw[k] = P.uniform_vel_wall[k];
            double dotnv = n[0] * vw[0] + n[1] * vw[1] + n[2] * vw[2];
            double vt[3];
            vt[0] = ux - 2.0 * (vw[0] - dotnv * n[0]);
            vt[1] = uy - 2.0 * (vw[1] - dotnv * n[1]);
            vt[2] = uz - 2.0 * (vw[2] - dotnv * n[2]);
            umx = Ri[0] * vt[0] + Ri[1] * vt[1] + Ri[2] * vt[2];
            umy = Ri[3] * vt[0] + Ri[4] * vt[1] + Ri[5] * vt[2];
            umz = Ri[6] * vt[0] + Ri[7] * vt[1] + Ri[8] * vt[2];
            for (int k = 0; k < 9; k++) Rm[k] = Ri[k];
        }
        const double dns_i = (s.ptype[i] == 1) ? P.dns_fl1 : P.dns_fl2;
        walk(s, xi[0], xi[1], xi[2], [&](int j) {
            double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_l2 && dm2 < P.re_l2 && d2 < dm2) {
                if (j != i) {
                    double dst = sqrt(dm2);
                    double wL = weight(dst, P.re_l, P.weight_type);
                    double nu;
                    if ((meu_i + s.MEU[j]) > kEps) nu = 2.0 * meu_i * s.MEU[j] / (meu_i + s.MEU[j]);
                    else nu = 0.0;
                    nu = nu / dns_i;
                    ax += (s.vel[j * 3] - umx) * wL * nu;
                    ay += (s.vel[j * 3 + 1] - umy) * wL * nu;
                    az += (s.vel[j * 3 + 2] - umz) * wL * nu;
                }
            }
        });
        {
            double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (dm2 < P.re_l2) {
                double dst = sqrt(dm2);
                double wL = weight(dst, P.re_l, P.weight_type);
                double nu;
                if (meu_i > kEps) nu = 2.0 * meu_i * meu_i / (meu_i + meu_i);
                else nu = 0.0;
                nu = nu / dns_i;
                ax += (ux - umx) * wL * nu;
                ay += (uy - umy) * wL * nu;
                az += (uz - umz) * wL * nu;
            }
        }
        double rx = (Rm[0] * ax + Rm[1] * ay + Rm[2] * az) * P.coeff_visc_multiphase;
        double ry = (Rm[3] * ax + Rm[4] * ay + Rm[5] * az) * P.coeff_visc_multiphase;
        double rz = (Rm[6] * ax + Rm[7] * ay + Rm[8] * az) * P.coeff_visc_multiphase;
        s.acc[i * 3] += rx;
        s.acc[i * 3 + 1] += ry;
        s.acc[i * 3 + 2] += rz;
        s.forceWall[i * 3] += -(Rm[0] * ax + Rm[1] * ay + Rm[2] * az) * P.coeff_visc_multiphase * vol * s.RHO[i];
        s.forceWall[i * 3 + 1] += -(Rm[3] * ax + Rm[4] * ay + Rm[5] * az) * P.coeff_visc_multiphase * vol * s.RHO[i];
        s.forceWall[i * 3 + 2] += -(Rm[6] * ax + Rm[7] * ay + Rm[8] * az) * P.coeff_visc_multiphase * vol * s.RHO[i];
    }
}

// ------------------------------------------- non-Newtonian (granular) phase
// MpsViscosity::calcVolumeFraction (src/MpsViscosity.cpp:104-191): share of
// phase-2 (PTYPE >= 2) fluid neighbours within re_S, counted (method 1) or
// kernel weighted (method 2).
void volumeFraction(Sys& s) {
    const mps_params& P = s.p;
    if (P.fraction_method != 1 && P.fraction_method != 2) return;
    const bool weighted = P.fraction_method == 2;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] != P.fluid) continue;
        double sum1 = 0.0, sum2 = 0.0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            if (d2 < P.re_s2) {
                if (j != i && s.type[j] == P.fluid) {
                    if (weighted) {
                        double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                        sum1 += wS;
                        if (s.ptype[j] >= 2) sum2 += wS;
                    } else {
                        sum1 += 1.0;
                        if (s.ptype[j] >= 2) sum2 += 1.0;
                    }
                }
            }
        });
        s.Cv[i] = (sum1 < kEps) ? 0.0 : sum2 / sum1;
    }
}

// Bed level BL, water level WL and the pressure PS of the bed-level particle
// on a horizontal grid of 2 l0 columns (src/MpsViscosity.cpp:236-305).  The
// reference fills it from a parallel loop without synchronisation; the
// defined result is that of the loop in index order: BL is the highest
// phase-2-rich (Cv > 0.5) fluid particle of a column, PS the pressure of the
// lowest index among the particles at that height, WL the highest phase-1
// particle; a column nobody raises keeps its initial level and PS = 0.
// Particles outside the grid (an out-of-bounds write in the reference) are
// left out.  wallForm: the polygon-wall variants' grid (3-D only, :766-783).
struct BedGrid {
    int kxMax = 0, kyMax = 0;
    double xmin = 0, ymin = 0, aa = 0, level0 = 0;
    std::vector<double> BL, WL, PS;
    // column index (1-based as the reference's arrays), -1 outside
    int cell(int kx, int ky) const {
        if (kx < 1 || kx > kxMax || ky < 1 || ky > kyMax) return -1;
        return (kx - 1) * kyMax + (ky - 1);
    }
};

static BedGrid bedGrid(const Sys& s, bool wallForm) {
    const mps_params& P = s.p;
    BedGrid g;
    const bool d3 = (int)P.dim == 3;
    const double Xmin = P.domain_min[0], Xmax = P.domain_max[0], Ymin = P.domain_min[1];
    const double Ymax = d3 ? P.domain_max[1] : Ymin;
    const double Zmin = d3 ? P.domain_min[2] : 0.0;
    g.aa = 2.0 * P.part_dist;
    g.kxMax = int((Xmax - Xmin) / g.aa) + 1;
    g.kyMax = int((Ymax - Ymin) / g.aa) + 1;
    g.xmin = Xmin;
    g.ymin = Ymin;
    g.level0 = (d3 || wallForm) ? Zmin : Ymin;
    const size_t nc = (size_t)g.kxMax * g.kyMax;
    g.BL.assign(nc, g.level0);
    g.WL.assign(nc, g.level0);
    g.PS.assign(nc, 0.0);
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] != P.fluid) continue;
        const double x = s.pos[i * 3], y = s.pos[i * 3 + 1], z = s.pos[i * 3 + 2];
        const int kx = int((x - Xmin) / g.aa) + 1;
        const int ky = d3 ? int((y - Ymin) / g.aa) + 1 : 1;
        const int c = g.cell(kx, ky);
        if (c < 0) continue;
        const double h = d3 ? z : y;
        if (h > g.BL[c] && s.Cv[i] > 0.5) { g.BL[c] = h; g.PS[c] = s.press[i]; }
        if (h > g.WL[c] && s.ptype[i] == 1) g.WL[c] = h;
    }
    return g;
}

// Strain-rate tensor, its second invariant and the smoothed pressure of the
// fluid particles (src/MpsViscosity.cpp:307-440).
static void strainRate(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] != P.fluid) continue;
        double sum1 = 0.0, sum2 = 0.0, sum3 = 0.0, sum4 = 0.0, sum5 = 0.0, sum6 = 0.0, sum7 = 0.0, sum8 = 0.0,
               sum9 = 0.0, sum10 = 0.0, sumWs = 0.0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double ux = s.vel[i * 3], uy = s.vel[i * 3 + 1], uz = s.vel[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        const double MC[9] = {s.cm1[i * 3], s.cm1[i * 3 + 1], s.cm1[i * 3 + 2], s.cm2[i * 3], s.cm2[i * 3 + 1],
                              s.cm2[i * 3 + 2], s.cm3[i * 3], s.cm3[i * 3 + 1], s.cm3[i * 3 + 2]};
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
                if (j != i) {
                    double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                    double ex = s.vel[j * 3] - ux, ey = s.vel[j * 3 + 1] - uy, ez = s.vel[j * 3 + 2] - uz;
                    double inv = 1.0 / d2;
                    if (!P.gradient_correction) {
                        sum1 += ex * v0 * wS * inv; sum2 += ex * v1 * wS * inv; sum3 += ex * v2 * wS * inv;
                        sum4 += ey * v0 * wS * inv; sum5 += ey * v1 * wS * inv; sum6 += ey * v2 * wS * inv;
                        sum7 += ez * v0 * wS * inv; sum8 += ez * v1 * wS * inv; sum9 += ez * v2 * wS * inv;
                    } else {
                        double c0 = (v0 * MC[0] + v1 * MC[1] + v2 * MC[2]);
                        double c1 = (v0 * MC[3] + v1 * MC[4] + v2 * MC[5]);
                        double c2 = (v0 * MC[6] + v1 * MC[7] + v2 * MC[8]);
                        sum1 += ex * c0 * wS * inv; sum2 += ex * c1 * wS * inv; sum3 += ex * c2 * wS * inv;
                        sum4 += ey * c0 * wS * inv; sum5 += ey * c1 * wS * inv; sum6 += ey * c2 * wS * inv;
                        sum7 += ez * c0 * wS * inv; sum8 += ez * c1 * wS * inv; sum9 += ez * c2 * wS * inv;
                    }
                    sum10 += s.press[j] * wS;
                    sumWs += wS;
                }
            }
        });
        const double c = P.coeff_press_grad;
        const double Uxx = -c * sum1, Uxy = -c * sum2, Uxz = -c * sum3;
        const double Uyx = -c * sum4, Uyy = -c * sum5, Uyz = -c * sum6;
        const double Uzx = -c * sum7, Uzy = -c * sum8, Uzz = -c * sum9;
        s.pSmooth[i] = (sumWs > kEps) ? sum10 / sumWs : s.press[i];
        if (s.pSmooth[i] < kEps) s.pSmooth[i] = 0.0;
        s.S11[i] = 0.5 * (Uxx + Uxx);
        s.S12[i] = 0.5 * (Uxy + Uyx);
        s.S13[i] = 0.5 * (Uxz + Uzx);
        s.S22[i] = 0.5 * (Uyy + Uyy);
        s.S23[i] = 0.5 * (Uyz + Uzy);
        s.S33[i] = 0.5 * (Uzz + Uzz);
        double ii = 0.5 * (s.S11[i] * s.S11[i] + 2.0 * s.S12[i] * s.S12[i] + 2.0 * s.S13[i] * s.S13[i] +
                           s.S22[i] * s.S22[i] + 2.0 * s.S23[i] * s.S23[i] + s.S33[i] * s.S33[i]);
        if (ii < kEps || ii * 0 != 0) ii = 0.0;
        s.II[i] = ii;
    }
}

// Granular rheology of the fluid particles (src/MpsViscosity.cpp:461-691):
// mu(I) with a Drucker-Prager yield stress; sets MEU, MEU_Y, inertia, VF, RHO,
// the normal stress and CFLvisc.  The reference runs this loop serially.
static void rheologyFluid(Sys& s, const BedGrid& g) {
    const mps_params& P = s.p;
    const double gmod = sqrt(P.gravity[0] * P.gravity[0] + P.gravity[1] * P.gravity[1] + P.gravity[2] * P.gravity[2]);
    const bool d3 = (int)P.dim == 3;
    const double f21 = P.dns_fl2 - P.dns_fl1;
    double miMax = 0.0;
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] != P.fluid) continue;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double vx = s.vel[i * 3], vy = s.vel[i * 3 + 1], vz = s.vel[i * 3 + 2];
        const double vel2i = vx * vx + vy * vy + vz * vz;
        if (s.ptype[i] == 1) {
            s.MEU[i] = P.knm_vs1 * P.dns_fl1;
        } else if (s.ptype[i] == 2) {
            double phi = (s.Cv[i] - 0.25) * P.phi_1 / (1.0 - 0.25);
            double phi2 = (s.Cv[i] - 0.25) * P.phi_2 / (1.0 - 0.25);
            if (s.Cv[i] <= 0.25) { phi = kEps; phi2 = kEps; }
            s.pRheo[i] = s.pSmooth[i];
            const int kx = int((xi - g.xmin) / g.aa) + 1;
            const int ky = d3 ? int((yi - g.ymin) / g.aa) + 1 : 1;
            const int c = g.cell(kx, ky);
            const double BL = c < 0 ? g.level0 : g.BL[c];
            const double WL = c < 0 ? g.level0 : g.WL[c];
            const double PS = c < 0 ? 0.0 : g.PS[c];
            double ns;
            if (!d3) {
                ns = (BL - yi + P.part_dist * 0.5) * f21 * gmod - vel2i * f21 * 0.5;
                if (WL < BL) {
                    ns = s.pSmooth[i];
                } else {
                    // the reference first forms (p_smooth - PS) f21 VF / RHO and overwrites it
                    ns = s.pSmooth[i] - PS;
                }
            } else {
                ns = (BL - zi + P.part_dist * 0.5) * f21 * gmod - vel2i * f21 * 0.5;
                if (s.pSmooth[i] - (WL - zi) * P.dns_fl1 * gmod < kEps) s.pSmooth[i] = (WL - zi) * P.dns_fl1 * gmod;
                if (s.timeCurrent <= 1.0)
                    ns = (1.0 - s.timeCurrent) * (s.pSmooth[i] - (WL - zi) * P.dns_fl1 * gmod) + s.timeCurrent * ns;
            }
            if (ns < 1.0 || s.Cv[i] < 0.5) ns = 1.0;
            s.pRheo[i] = ns;
            const double modE = sqrt(s.II[i]);
            if (WL < BL) s.inertia[i] = modE * P.dg / sqrt(ns / P.dns_sdt);
            else s.inertia[i] = modE * P.dg / sqrt(ns / (P.dns_fl1 * P.cd));
            s.VF[i] = P.vf_max - (P.vf_max - P.vf_min) * s.inertia[i];
            if (s.VF[i] < P.vf_min) s.VF[i] = P.vf_min;
            s.RHO[i] = P.dns_sdt * s.VF[i] + (1.0 - s.VF[i]) * P.dns_fl1;
            phi *= (s.VF[i] / P.vf_max);
            const double alpha1 = 2.0 * sqrt(3.0) * sin(phi) / (3.0 - sin(phi));
            const double beta1 = 2.0 * sqrt(3.0) * cos(phi) / (3.0 - sin(phi));
            double yield = P.cohes * beta1 + ns * alpha1;
            if (yield < kEps) yield = 0.0;
            double viscMax = (P.meu0 + yield * P.mm * 0.5);
            s.MEUY[i] = yield / (2.0 * sqrt(modE * modE + kEps));
            double meu0;
            if (WL < BL)
                meu0 = 0.5 * (tan(phi2) - tan(phi)) * ns * P.dg / (P.i0 * sqrt(ns / P.dns_fl2) + modE * P.dg) * modE /
                       (sqrt(modE * modE + kEps));
            else
                meu0 = 0.5 * (tan(phi2) - tan(phi)) * ns * P.dg / (P.i0 * sqrt(ns / (P.dns_fl1 * P.cd)) + modE * P.dg) *
                       modE / (sqrt(modE * modE + kEps));
            if (modE <= kEps || (meu0 * 0) != 0) meu0 = P.meu0;
            viscMax = (meu0 + yield * P.mm * 0.5);
            s.MEU[i] = s.MEUY[i] + meu0;
            if (s.II[i] <= kEps) s.MEU[i] = viscMax;
            if (s.MEU[i] > miMax) miMax = s.MEU[i];
        }
        if (s.ptype[i] >= 2) {
            if (s.Cv[i] > 0.5) s.RHO[i] = P.dns_fl2;
            else s.RHO[i] = s.Cv[i] * P.dns_fl2 + (1.0 - s.Cv[i]) * P.dns_fl1;
        }
    }
    s.cflVisc = P.time_step * miMax /
                ((P.dns_sdt * P.vf_max + (1.0 - P.vf_max) * P.dns_fl1) * P.part_dist * P.part_dist);
}

// MpsViscosity::calcViscosityInteractionVal (src/MpsViscosity.cpp:194-700)
void viscosityInteractionVal(Sys& s) {
    BedGrid g = bedGrid(s, false);
    strainRate(s);
    rheologyFluid(s, g);
}

// MpsViscosity::calcWall{Slip,NoSlip}ViscosityInteractionVal
// (src/MpsViscosity.cpp:703-1110, 1112-1526), polygon walls in 3-D: the
// mirror-neighbourhood strain rate of the fluid particles near the wall
// replaces S (II and the smoothed pressure add to the fluid values), then
// the rheology of those particles with a Papanastasiou-regularised yield
// viscosity.
void wallViscosityInteractionVal(Sys& s) {
    const mps_params& P = s.p;
    const bool noSlip = P.slip_condition == 0;
    BedGrid g = bedGrid(s, true);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (!(s.type[i] == P.fluid && s.nearWall[i])) continue;
        double sum1 = 0.0, sum2 = 0.0, sum3 = 0.0, sum4 = 0.0, sum5 = 0.0, sum6 = 0.0, sum7 = 0.0, sum8 = 0.0,
               sum9 = 0.0, sum10 = 0.0;
        const double xi[3] = {s.pos[i * 3], s.pos[i * 3 + 1], s.pos[i * 3 + 2]};
        const double ux = s.vel[i * 3], uy = s.vel[i * 3 + 1], uz = s.vel[i * 3 + 2];
        const double xm[3] = {s.mirror[i * 3], s.mirror[i * 3 + 1], s.mirror[i * 3 + 2]};
        double n[3], R[9];
        reflectionMod2(xi, xm, n, R);
        double vmx, vmy, vmz;
        if (!noSlip) {   // mirror velocity R v (:826-829)
            vmx = (R[0] * ux + R[1] * uy + R[2] * uz);
            vmy = (R[3] * ux + R[4] * uy + R[5] * uz);
            vmz = (R[6] * ux + R[7] * uy + R[8] * uz);
        } else {         // -(v - 2 (v_wall - (n.v_wall) n)) (:1239-1259)
            double vw[3] = {0.0, 0.0, 0.0};
            if (s.nearMeshType[i] == 2)
                for (int k = 0; k < 3; k++) vw[k] = P.uniform_vel_wall[k];
            const double dotnv = n[0] * vw[0] + n[1] * vw[1] + n[2] * vw[2];
            const double t0 = ux - 2.0 * (vw[0] - dotnv * n[0]);
            const double t1 = uy - 2.0 * (vw[1] - dotnv * n[1]);
            const double t2 = uz - 2.0 * (vw[2] - dotnv * n[2]);
            vmx = (-1.0 * t0 + 0.0 * t1 + 0.0 * t2);
            vmy = (0.0 * t0 + -1.0 * t1 + 0.0 * t2);
            vmz = (0.0 * t0 + 0.0 * t1 + -1.0 * t2);
        }
        walk(s, xi[0], xi[1], xi[2], [&](int j) {
            double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2) {
                if (j != i) {
                    double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
                    double ex = s.vel[j * 3] - vmx, ey = s.vel[j * 3 + 1] - vmy, ez = s.vel[j * 3 + 2] - vmz;
                    sum1 += ex * w0 * wS / dm2; sum2 += ex * w1 * wS / dm2; sum3 += ex * w2 * wS / dm2;
                    sum4 += ey * w0 * wS / dm2; sum5 += ey * w1 * wS / dm2; sum6 += ey * w2 * wS / dm2;
                    sum7 += ez * w0 * wS / dm2; sum8 += ez * w1 * wS / dm2; sum9 += ez * w2 * wS / dm2;
                    sum10 += s.press[j] * wS;
                }
            }
        });
        const double c = P.coeff_press_grad;
        double Uxx, Uxy, Uxz, Uyx, Uyy, Uyz, Uzx, Uzy, Uzz;
        if (!noSlip) {   // R grad(u) R (:878-905)
            const double aUxx = -c * (R[0] * sum1 + R[1] * sum4 + R[2] * sum7);
            const double aUxy = -c * (R[0] * sum2 + R[1] * sum5 + R[2] * sum8);
            const double aUxz = -c * (R[0] * sum3 + R[1] * sum6 + R[2] * sum9);
            const double aUyx = -c * (R[3] * sum1 + R[4] * sum4 + R[5] * sum7);
            const double aUyy = -c * (R[3] * sum2 + R[4] * sum5 + R[5] * sum8);
            const double aUyz = -c * (R[3] * sum3 + R[4] * sum6 + R[5] * sum9);
            const double aUzx = -c * (R[6] * sum1 + R[7] * sum4 + R[8] * sum7);
            const double aUzy = -c * (R[6] * sum2 + R[7] * sum5 + R[8] * sum8);
            const double aUzz = -c * (R[6] * sum3 + R[7] * sum6 + R[8] * sum9);
            Uxx = aUxx * R[0] + aUxy * R[3] + aUxz * R[6];
            Uxy = aUxx * R[1] + aUxy * R[4] + aUxz * R[7];
            Uxz = aUxx * R[2] + aUxy * R[5] + aUxz * R[8];
            Uyx = aUyx * R[0] + aUyy * R[3] + aUyz * R[6];
            Uyy = aUyx * R[1] + aUyy * R[4] + aUyz * R[7];
            Uyz = aUyx * R[2] + aUyy * R[5] + aUyz * R[8];
            Uzx = aUzx * R[0] + aUzy * R[3] + aUzz * R[6];
            Uzy = aUzx * R[1] + aUzy * R[4] + aUzz * R[7];
            Uzz = aUzx * R[2] + aUzy * R[5] + aUzz * R[8];
        } else {         // -grad(u) R (:1310-1320)
            Uxx = c * (sum1 * R[0] + sum2 * R[3] + sum3 * R[6]);
            Uxy = c * (sum1 * R[1] + sum2 * R[4] + sum3 * R[7]);
            Uxz = c * (sum1 * R[2] + sum2 * R[5] + sum3 * R[8]);
            Uyx = c * (sum4 * R[0] + sum5 * R[3] + sum6 * R[6]);
            Uyy = c * (sum4 * R[1] + sum5 * R[4] + sum6 * R[7]);
            Uyz = c * (sum4 * R[2] + sum5 * R[5] + sum6 * R[8]);
            Uzx = c * (sum7 * R[0] + sum8 * R[3] + sum9 * R[6]);
            Uzy = c * (sum7 * R[1] + sum8 * R[4] + sum9 * R[7]);
            Uzz = c * (sum7 * R[2] + sum8 * R[5] + sum9 * R[8]);
        }
        s.pSmooth[i] += sum10 / P.pnd_gradient_zero;
        if (s.pSmooth[i] < kEps) s.pSmooth[i] = 0.0;
        s.S11[i] = 0.5 * (Uxx + Uxx);
        s.S12[i] = 0.5 * (Uxy + Uyx);
        s.S13[i] = 0.5 * (Uxz + Uzx);
        s.S22[i] = 0.5 * (Uyy + Uyy);
        s.S23[i] = 0.5 * (Uyz + Uzy);
        s.S33[i] = 0.5 * (Uzz + Uzz);
        s.II[i] += 0.5 * (s.S11[i] * s.S11[i] + 2.0 * s.S12[i] * s.S12[i] + 2.0 * s.S13[i] * s.S13[i] +
                          s.S22[i] * s.S22[i] + 2.0 * s.S23[i] * s.S23[i] + s.S33[i] * s.S33[i]);
        if (s.II[i] < kEps || s.II[i] * 0 != 0) s.II[i] = 0.0;
    }
    const double gmod = sqrt(P.gravity[0] * P.gravity[0] + P.gravity[1] * P.gravity[1] + P.gravity[2] * P.gravity[2]);
    const double f21 = P.dns_fl2 - P.dns_fl1;
    for (int i = 0; i < s.N; i++) {
        if (!(s.type[i] == P.fluid && s.nearWall[i])) continue;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double vx = s.vel[i * 3], vy = s.vel[i * 3 + 1], vz = s.vel[i * 3 + 2];
        const double vel2i = vx * vx + vy * vy + vz * vz;
        if (s.ptype[i] == 1) {
            s.MEU[i] = P.knm_vs1 * P.dns_fl1;
        } else if (s.ptype[i] == 2) {
            double phi = (s.Cv[i] - 0.25) * P.phi_1 / (1.0 - 0.25);
            double phi2 = (s.Cv[i] - 0.25) * P.phi_2 / (1.0 - 0.25);
            if (s.Cv[i] <= 0.25) { phi = kEps; phi2 = kEps; }
            s.pRheo[i] = s.pSmooth[i];
            const int kx = int((xi - g.xmin) / g.aa) + 1;
            const int ky = int((yi - g.ymin) / g.aa) + 1;
            const int c = g.cell(kx, ky);
            const double BL = c < 0 ? g.level0 : g.BL[c];
            const double WL = c < 0 ? g.level0 : g.WL[c];
            double ns = (BL - zi + P.part_dist * 0.5) * f21 * gmod - vel2i * f21 * 0.5;
            if (s.pSmooth[i] < (WL - zi) * P.dns_fl1 * gmod) s.pSmooth[i] = (WL - zi) * P.dns_fl1 * gmod;
            if (s.timeCurrent <= 1.0)
                ns = (1.0 - s.timeCurrent) * (s.pSmooth[i] - (WL - zi) * P.dns_fl1 * gmod) + s.timeCurrent * ns;
            if (ns < 1.0 || s.Cv[i] < 0.5) ns = 1.0;
            s.pRheo[i] = ns;
            s.inertia[i] = sqrt(s.II[i]) * P.dg / sqrt(ns / (P.dns_fl1 * P.cd));
            s.VF[i] = P.vf_max - (P.vf_max - P.vf_min) * s.inertia[i];
            if (s.VF[i] < P.vf_min) s.VF[i] = P.vf_min;
            s.RHO[i] = P.dns_sdt * s.VF[i] + (1.0 - s.VF[i]) * P.dns_fl1;
            phi = phi * s.VF[i] / P.vf_max;
            double yield = P.cohes * cos(phi) + ns * sin(phi);
            if (noSlip ? yield < 0.0 : yield < kEps) yield = 0.0;
            double viscMax = (yield * P.mm * 0.5 + P.meu0);
            if (s.II[i] > kEps)
                s.MEUY[i] = yield * (1.0 - exp(-P.mm * sqrt(s.II[i]))) * 0.5 / sqrt(s.II[i]);
            else
                s.MEUY[i] = viscMax;
            double meu0 = 0.5 * (tan(phi2) - tan(phi)) * ns * P.dg /
                          (P.i0 * sqrt(ns / (P.dns_fl1 * P.cd)) + sqrt(s.II[i]) * P.dg);
            if (s.II[i] <= kEps || (meu0 * 0) != 0) meu0 = P.meu0;
            viscMax = (yield * P.mm * 0.5 + meu0);
            s.MEU[i] = s.MEUY[i] + P.meu0 * pow(4.0 * s.II[i], (P.power_n - 1.0) * 0.5);
            if (s.II[i] == 0 || s.MEU[i] > viscMax) s.MEU[i] = viscMax;
        }
        if (s.ptype[i] >= 2) {
            if (s.Cv[i] > 0.5) s.RHO[i] = P.dns_fl2;
            else s.RHO[i] = s.Cv[i] * P.dns_fl2 + (1.0 - s.Cv[i]) * P.dns_fl1;
        }
    }
}

// main.cpp's non-Newtonian calls after the bucket update (src/main.cpp:259-262)
// and before the wall viscosity (:387-398)
void nonNewtonian(Sys& s, bool withWall) {
    volumeFraction(s);
    viscosityInteractionVal(s);
    if (withWall && s.p.wall_type == MPS_WALL_POLYGON) wallViscosityInteractionVal(s);
}

// ------------------------------------------------ MpsParticle / VelPos
// MpsParticle::addGravity (src/MpsParticle.cpp:111-130)
void addGravity(Sys& s) {
    const mps_params& P = s.p;
#pragma omp parallel for
    for (int i = 0; i < s.N; i++) {
        s.acc[i * 3] += P.gravity[0];
        s.acc[i * 3 + 1] += P.gravity[1];
        s.acc[i * 3 + 2] += P.gravity[2];
        s.accStar[i * 3] = s.acc[i * 3];
        s.accStar[i * 3 + 1] = s.acc[i * 3 + 1];
        s.accStar[i * 3 + 2] = s.acc[i * 3 + 2];
    }
}

// MpsParticleVelPos::predictionVelocityPosition (src/MpsParticleVelPos.cpp:20-50)
void prediction(Sys& s) {
    const mps_params& P = s.p;
    const double far = 10e8 * P.part_dist;
#pragma omp parallel for
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] == P.fluid) {
            for (int k = 0; k < 3; k++) s.vel[i * 3 + k] += s.acc[i * 3 + k] * P.time_step;
            for (int k = 0; k < 3; k++) s.pos[i * 3 + k] += s.vel[i * 3 + k] * P.time_step;
        }
        for (int k = 0; k < 3; k++) {
            s.acc[i * 3 + k] = 0.0; s.dvel[i * 3 + k] = 0.0;
            s.wpf1[i * 3 + k] = 0.0; s.wpf2[i * 3 + k] = 0.0; s.npcd[i * 3 + k] = 0.0;
        }
        s.numNeighWall[i] = 0;
        s.nearWall[i] = 0;
        s.riw2[i] = far;
        s.nNeighSurf[i] = 0.0;
        if (P.wall_type == MPS_WALL_POLYGON)
            for (int k = 0; k < 3; k++) s.mirror[i * 3 + k] = far;
    }
}

// MpsParticleVelPos::correctionVelocityPosition (src/MpsParticleVelPos.cpp:53-97)
void correction(Sys& s) {
    const mps_params& P = s.p;
    double vmax = 0.0;
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] != P.fluid) continue;
        for (int k = 0; k < 3; k++) s.vel[i * 3 + k] += s.acc[i * 3 + k] * P.time_step;
        for (int k = 0; k < 3; k++) s.pos[i * 3 + k] += s.acc[i * 3 + k] * P.time_step * P.time_step;
        for (int k = 0; k < 3; k++) s.acc[i * 3 + k] = 0.0;
        double v2 = s.vel[i * 3] * s.vel[i * 3] + s.vel[i * 3 + 1] * s.vel[i * 3 + 1] + s.vel[i * 3 + 2] * s.vel[i * 3 + 2];
        if (v2 > vmax * vmax) vmax = sqrt(v2);
    }
    s.velMax = vmax;
    s.cflCurrent = P.time_step * s.velMax / P.part_dist;
}

// MpsParticleVelPos::updateVelocityParticlesWallDummy (src/MpsParticleVelPos.cpp:100-158)
void wallDummyVelocity(Sys& s) {
    const mps_params& P = s.p;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] != P.dummy_wall) continue;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        double dx = 0, dy = 0, dz = 0, ni = 0;
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            if (d2 < P.re_l2 && j != i && s.type[j] == P.fluid) {
                double wL = weight(sqrt(d2), P.re_l, P.weight_type);
                ni += wL;
                dx += s.vel[j * 3] * wL;
                dy += s.vel[j * 3 + 1] * wL;
                dz += s.vel[j * 3 + 2] * wL;
            }
        });
        if (ni > kEps) {
            s.vel[i * 3] = 2.0 * 0.0 - dx / ni;
            s.vel[i * 3 + 1] = 2.0 * 0.0 - dy / ni;
            s.vel[i * 3 + 2] = 2.0 * 0.0 - dz / ni;
        } else {
            s.vel[i * 3] = 2.0 * 0.0 - dx;
            s.vel[i * 3 + 1] = 2.0 * 0.0 - dy;
            s.vel[i * 3 + 2] = 2.0 * 0.0 - dz;
        }
    }
}

// MpsParticle::checkParticleOutDomain (src/MpsParticle.cpp:234-402): the
// periodic wrap (:236-269), then the removal of fluid outside the domain
void checkOutDomain(Sys& s) {
    const mps_params& P = s.p;
    if (P.domain_type_bc == 1)
        for (int i = 0; i < s.N; i++)
            for (int k = 0; k < 3; k++) {
                if (!P.periodic_dir[k]) continue;
                if (s.pos[i * 3 + k] > P.phys_max[k]) s.pos[i * 3 + k] -= P.periodic_length[k];
                if (s.pos[i * 3 + k] < P.phys_min[k]) s.pos[i * 3 + k] += P.periodic_length[k];
            }
    int newN = s.N;
    double lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = P.domain_min[k] + P.bucket_side;
        hi[k] = P.domain_max[k] - P.bucket_side;
    }
    for (int i = 0; i < s.N; i++) {
        const double* x = &s.pos[i * 3];
        bool out = x[0] > hi[0] || x[0] < lo[0] || x[1] > hi[1] || x[1] < lo[1] ||
                   (P.dim == 3 && (x[2] > hi[2] || x[2] < lo[2]));
        if (s.type[i] == P.fluid && i < newN && out) {
            int L = newN - 1;
            s.type[i] = s.type[L]; s.bc[i] = s.bc[L]; s.numNeigh[i] = s.numNeigh[L];
            s.press[i] = s.press[L]; s.pressAvg[i] = s.pressAvg[L];
            s.pndi[i] = s.pndi[L]; s.pndki[i] = s.pndki[L]; s.pndski[i] = s.pndski[L];
            s.pndSmall[i] = s.pndSmall[L]; s.npcd2[i] = s.npcd2[L]; s.conc[i] = s.conc[L];
            s.divU[i] = s.divU[L]; s.diffTerm[i] = s.diffTerm[L]; s.nearMeshType[i] = s.nearMeshType[L];
            s.nearWall[i] = s.nearWall[L]; s.numNeighWall[i] = s.numNeighWall[L];
            s.pndWall[i] = s.pndWall[L]; s.devDotN[i] = s.devDotN[L]; s.nNeighSurf[i] = s.nNeighSurf[L];
            s.riw2[i] = s.riw2[L]; s.ptype[i] = s.ptype[L]; s.RHO[i] = s.RHO[L]; s.MEU[i] = s.MEU[L];
            s.elementID[i] = s.elementID[L];
            if (P.fluid_type == 1) {   // src/MpsParticle.cpp:322-337
                s.Cv[i] = s.Cv[L]; s.II[i] = s.II[L]; s.MEUY[i] = s.MEUY[L]; s.inertia[i] = s.inertia[L];
                s.pRheo[i] = s.pRheo[L]; s.pSmooth[i] = s.pSmooth[L]; s.VF[i] = s.VF[L];
                s.S12[i] = s.S12[L]; s.S13[i] = s.S13[L]; s.S23[i] = s.S23[L];
                s.S11[i] = s.S11[L]; s.S22[i] = s.S22[L]; s.S33[i] = s.S33[L];
            }
            if (P.mps_type == MPS_IMPLICIT_PND || P.mps_type == MPS_IMPLICIT_PND_DIVU) {
                s.ppeX[i] = s.ppeX[L];
                s.ppeB[i] = s.ppeB[L];
            }
            for (int k = 0; k < 3; k++) {
                s.acc[i * 3 + k] = s.acc[L * 3 + k]; s.accStar[i * 3 + k] = s.accStar[L * 3 + k];
                s.pos[i * 3 + k] = s.pos[L * 3 + k]; s.vel[i * 3 + k] = s.vel[L * 3 + k];
                s.npcd[i * 3 + k] = s.npcd[L * 3 + k]; s.gradC[i * 3 + k] = s.gradC[L * 3 + k];
                s.cm1[i * 3 + k] = s.cm1[L * 3 + k]; s.cm2[i * 3 + k] = s.cm2[L * 3 + k];
                s.cm3[i * 3 + k] = s.cm3[L * 3 + k]; s.normal[i * 3 + k] = s.normal[L * 3 + k];
                s.dvel[i * 3 + k] = s.dvel[L * 3 + k]; s.wallPos[i * 3 + k] = s.wallPos[L * 3 + k];
                s.mirror[i * 3 + k] = s.mirror[L * 3 + k]; s.wpf1[i * 3 + k] = s.wpf1[L * 3 + k];
                s.wpf2[i * 3 + k] = s.wpf2[L * 3 + k]; s.polyN[i * 3 + k] = s.polyN[L * 3 + k];
                s.forceWall[i * 3 + k] = s.forceWall[L * 3 + k];
            }
            s.type[L] = P.ghost; s.bc[L] = P.other; s.nearWall[L] = 0; s.nearMeshType[L] = 0;
            s.riw2[L] = 10e8 * P.part_dist;
            for (int k = 0; k < 3; k++) s.vel[L * 3 + k] = 0.0;
            s.press[L] = 0.0;
            s.pos[L * 3] = P.domain_max[0] - P.part_dist;
            s.pos[L * 3 + 1] = P.domain_max[1] - P.part_dist;
            if (P.dim == 2) s.pos[L * 3 + 2] = 0.0;
            else s.pos[L * 3 + 2] = P.domain_max[2] - P.part_dist;
            newN--;
        }
        s.N = newN;
    }
}

// ------------------------------------------------- MpsParticleCollision
// checkParticleCollisions (src/MpsParticleCollision.cpp:22-119)
void collisionPC(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] != P.fluid) continue;
        double mi = (s.ptype[i] == 1) ? P.dns_fl1 : P.dns_fl2;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double ux = s.vel[i * 3], uy = s.vel[i * 3 + 1], uz = s.vel[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        double dux = 0, duy = 0, duz = 0;
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.dist_collision_limit2 && (d2 < dm2 || pw) && j != i) {
                double fDT = (ux - s.vel[j * 3]) * v0 + (uy - s.vel[j * 3 + 1]) * v1 + (uz - s.vel[j * 3 + 2]) * v2;
                if (fDT > 0.0) {
                    double mj;
                    if (s.type[j] == P.fluid) mj = (s.ptype[j] == 1) ? P.dns_fl1 : P.dns_fl2;
                    else mj = P.dns_wall;
                    fDT *= P.restitution_collision * mj / (mi + mj) / d2;
                    if (s.type[j] == P.fluid) {
                        dux -= v0 * fDT; duy -= v1 * fDT; duz -= v2 * fDT;
                    } else {
                        dux -= 2.0 * v0 * fDT; duy -= 2.0 * v1 * fDT; duz -= 2.0 * v2 * fDT;
                    }
                }
            }
        });
        s.dvel[i * 3] = dux; s.dvel[i * 3 + 1] = duy; s.dvel[i * 3 + 2] = duz;
    }
#pragma omp parallel for
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] == P.fluid)
            for (int k = 0; k < 3; k++) s.vel[i * 3 + k] += s.dvel[i * 3 + k];
        for (int k = 0; k < 3; k++) s.dvel[i * 3 + k] = 0.0;
    }
}

// checkDynamicParticleCollisions (src/MpsParticleCollision.cpp:124-292)
void collisionDPC(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
    const double Wij5 = 0.5 * 0.5 * 0.5 * 0.5 * 3.0;
    const double gmod = sqrt(P.gravity[0] * P.gravity[0] + P.gravity[1] * P.gravity[1] + P.gravity[2] * P.gravity[2]);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] != P.fluid) continue;
        double mi = (s.ptype[i] == 1) ? P.dns_fl1 : P.dns_fl2;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double ux = s.vel[i * 3], uy = s.vel[i * 3 + 1], uz = s.vel[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        double dux = 0, duy = 0, duz = 0;
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.part_dist * P.part_dist && (d2 < dm2 || pw) && j != i) {
                double mj;
                if (s.type[j] == P.fluid) mj = (s.ptype[j] == 1) ? P.dns_fl1 : P.dns_fl2;
                else mj = P.dns_wall;
                double dst = sqrt(d2);
                double Wij = 0.0;
                if (dst > kEps) {
                    double w1 = 1.0 - dst * P.inv_part_dist;
                    double w2 = 4.0 * dst * P.inv_part_dist + 1.0;
                    Wij = w1 * w1 * w1 * w1 * w2;
                }
                double chi = sqrt(Wij / Wij5);
                double kappa = (dst < 0.5 * P.part_dist) ? 1.0 : chi;
                double fDT = (ux - s.vel[j * 3]) * v0 + (uy - s.vel[j * 3 + 1]) * v1 + (uz - s.vel[j * 3 + 2]) * v2;
                if (fDT > 0.0) {
                    fDT *= kappa * 2.0 * mj / (mi + mj) / d2;
                    if (s.type[j] == P.fluid) {
                        dux -= v0 * fDT; duy -= v1 * fDT; duz -= v2 * fDT;
                    } else {
                        dux -= 2.0 * v0 * fDT; duy -= 2.0 * v1 * fDT; duz -= 2.0 * v2 * fDT;
                    }
                } else {
                    double pmax = 2.0 / 3.0 * s.RHO[i] * gmod * 0.2;
                    double pmin = s.RHO[i] * gmod * P.part_dist;
                    double ptil = std::max(std::min(P.lambda_collision * fabs(s.press[i] + s.press[j]), P.lambda_collision * pmax), pmin);
                    double pb = ptil * chi;
                    double rep = P.time_step / mi * chi * pb / d2;
                    if (s.type[j] == P.fluid) {
                        dux -= v0 * rep; duy -= v1 * rep; duz -= v2 * rep;
                    } else {
                        dux -= 2.0 * v0 * rep; duy -= 2.0 * v1 * rep; duz -= 2.0 * v2 * rep;
                    }
                }
            }
        });
        s.dvel[i * 3] = dux; s.dvel[i * 3 + 1] = duy; s.dvel[i * 3 + 2] = duz;
    }
#pragma omp parallel for
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] == P.fluid) {
            for (int k = 0; k < 3; k++) s.vel[i * 3 + k] += s.dvel[i * 3 + k];
            for (int k = 0; k < 3; k++) s.pos[i * 3 + k] += s.dvel[i * 3 + k] * P.time_step;
        }
        for (int k = 0; k < 3; k++) s.dvel[i * 3 + k] = 0.0;
    }
}

// --------------------------------------------------------- PolygonMesh
// PolygonMesh::initWijnNeigh (src/PolygonMesh.cpp:377-527)
void initWijnNeigh(Sys& s) {
    Mesh& m = s.mesh;
    const int dim = (int)s.p.dim, wt = s.p.weight_type;
    const double lo = s.p.part_dist, reL = s.p.re_l, reS = s.p.re_s;
    const double reS2 = reS * reS, reL2 = reL * reL;
    const int n = 50;
    m.xPND.clear(); m.xNeigh.clear(); m.Zij.clear(); m.nNeighWall.clear();
    int limS = (int)ceil(reS / lo), limL = (int)ceil(reL / lo);
    for (int i = 1; i <= n; i++) {
        double xi = 0.0, yi = i * reS / n, zi = 0.0;
        double riw = (dim == 2) ? sqrt((0.0 - xi) * (0.0 - xi) + (0.0 - yi) * (0.0 - yi))
                                : sqrt((0.0 - xi) * (0.0 - xi) + (0.0 - yi) * (0.0 - yi) + (0.0 - zi) * (0.0 - zi));
        m.xPND.push_back(riw / reS);
        m.xNeigh.push_back(riw / reL);
        double n0 = 0.0;
        int zlo = dim == 2 ? 0 : -limS, zhi = dim == 2 ? 0 : limS;
        for (int ix = -limS; ix <= limS; ix++)
            for (int iy = -limS; iy <= 0; iy++)
                for (int iz = zlo; iz <= zhi; iz++) {
                    double xj = lo * (double)ix, yj = lo * (double)(iy - 0.5), zj = lo * (double)iz;
                    double r2 = (dim == 2) ? (xj - xi) * (xj - xi) + (yj - yi) * (yj - yi)
                                           : (xj - xi) * (xj - xi) + (yj - yi) * (yj - yi) + (zj - zi) * (zj - zi);
                    if (r2 < reS2) {
                        if (r2 == 0.0) continue;
                        double r = sqrt(r2);
                        if (wt == 0) n0 += reS / r - 1.0;
                        else if (wt == 1) n0 += reS / r + r / reS - 2.0;
                        else if (wt == 2) n0 += reS / r - r / reS;
                        else if (wt == 3) n0 += pow(1.0 - r / reS, 3.0);
                        else if (wt == 4) n0 += pow(1.0 - r / reS, 2.0);
                        else n0 += reS / r - 1.0;
                    }
                }
        m.Zij.push_back(n0);
        int nn = 0;
        zlo = dim == 2 ? 0 : -limL; zhi = dim == 2 ? 0 : limL;
        for (int ix = -limL; ix <= limL; ix++)
            for (int iy = -limL; iy <= 0; iy++)
                for (int iz = zlo; iz <= zhi; iz++) {
                    double xj = lo * (double)ix, yj = lo * (double)(iy - 0.5), zj = lo * (double)iz;
                    double r2 = (dim == 2) ? (xj - xi) * (xj - xi) + (yj - yi) * (yj - yi)
                                           : (xj - xi) * (xj - xi) + (yj - yi) * (yj - yi) + (zj - zi) * (zj - zi);
                    if (r2 < reL2) {
                        if (r2 == 0.0) continue;
                        nn += 1;
                    }
                }
        m.nNeighWall.push_back(nn);
    }
}

// src/PolygonMesh.cpp:533-554
double interpolateWij(const Mesh& m, double re, double x) {
    int sz = (int)m.xPND.size();
    double r = x / re;
    int i = 0;
    if (r >= m.xPND[sz - 2]) i = sz - 2;
    else while (r > m.xPND[i + 1]) i++;
    double xL = m.xPND[i], yL = m.Zij[i], xR = m.xPND[i + 1], yR = m.Zij[i + 1];
    double dydx = (yR - yL) / (xR - xL);
    return yL + dydx * (r - xL);
}
// src/PolygonMesh.cpp:560-581
int interpolateNumNeighWall(const Mesh& m, double re, double x) {
    int sz = (int)m.xNeigh.size();
    double r = x / re;
    int i = 0;
    if (r >= m.xNeigh[sz - 2]) i = sz - 2;
    else while (r > m.xNeigh[i + 1]) i++;
    double xL = m.xNeigh[i], yL = (double)m.nNeighWall[i], xR = m.xNeigh[i + 1], yR = (double)m.nNeighWall[i + 1];
    double dydx = (yR - yL) / (xR - xL);
    return (int)(yL + dydx * (r - xL));
}

// Closest point on triangle (a,b,c) to p: Ericson, Real-Time Collision
// Detection, section 5.1.5 (Voronoi-region case analysis).
inline void closestPtTriangle(const double* p, const double* a, const double* b, const double* c, double* q) {
    double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
    double d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
    if (d1 <= 0.0 && d2 <= 0.0) { q[0] = a[0]; q[1] = a[1]; q[2] = a[2]; return; }
    double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    double d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
    double d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
    if (d3 >= 0.0 && d4 <= d3) { q[0] = b[0]; q[1] = b[1]; q[2] = b[2]; return; }
    double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        double v = d1 / (d1 - d3);
        q[0] = a[0] + v * ab[0]; q[1] = a[1] + v * ab[1]; q[2] = a[2] + v * ab[2];
        return;
    }
    double cp[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    double d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
    double d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
    if (d6 >= 0.0 && d5 <= d6) { q[0] = c[0]; q[1] = c[1]; q[2] = c[2]; return; }
    double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        double w = d2 / (d2 - d6);
        q[0] = a[0] + w * ac[0]; q[1] = a[1] + w * ac[1]; q[2] = a[2] + w * ac[2];
        return;
    }
    double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        q[0] = b[0] + w * (c[0] - b[0]); q[1] = b[1] + w * (c[1] - b[1]); q[2] = b[2] + w * (c[2] - b[2]);
        return;
    }
    double denom = 1.0 / (va + vb + vc);
    double v = vb * denom, w = vc * denom;
    q[0] = a[0] + ab[0] * v + ac[0] * w;
    q[1] = a[1] + ab[1] * v + ac[1] * w;
    q[2] = a[2] + ab[2] * v + ac[2] * w;
}

// PolygonMesh::closestPointPNDBoundaryAABB (src/PolygonMesh.cpp:585-715) for
// one mesh, brute force over faces: a strictly closer point replaces the
// particle's wall point, mirror, normal and mesh type (:615-650).
void closestPointMesh(Sys& s, const Mesh& m, int meshType) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < s.N; i++) {
        const double* p = &s.pos[i * 3];
        double best = DBL_MAX, q[3] = {0, 0, 0}, bq[3] = {0, 0, 0};
        int bf = -1;
        for (int f = 0; f < m.nf; f++) {
            const int* t = &m.F[f * 3];
            closestPtTriangle(p, &m.V[t[0] * 3], &m.V[t[1] * 3], &m.V[t[2] * 3], q);
            double e0 = p[0] - q[0], e1 = p[1] - q[1], e2 = p[2] - q[2];
            double d2 = e0 * e0 + e1 * e1 + e2 * e2;
            if (d2 < best) { best = d2; bf = f; bq[0] = q[0]; bq[1] = q[1]; bq[2] = q[2]; }
        }
        if (bf < 0) continue;
        if (best < s.riw2[i]) {
            s.riw2[i] = best;
            for (int k = 0; k < 3; k++) {
                s.wallPos[i * 3 + k] = bq[k];
                s.mirror[i * 3 + k] = p[k] + 2.0 * (bq[k] - p[k]);
                s.polyN[i * 3 + k] = m.Nrm[bf * 3 + k];
            }
            s.nearMeshType[i] = meshType;
        }
    }
}

// main.cpp's loop over the meshes in type order (src/main.cpp:296-300)
void closestPoint(Sys& s) {
    closestPointMesh(s, s.mesh, 0);
    for (int t = 1; t <= 2; t++)
        if (s.extraOn[t - 1]) closestPointMesh(s, s.extra[t - 1], t);
}

// unit face normals (v1 - v0) x (v2 - v0), zero for a degenerate face
// (igl::per_face_normals as updateForcedPolygonMesh calls it, :362-369)
void faceNormals(Mesh& m) {
    m.Nrm.assign(3 * (size_t)m.nf, 0.0);
    for (int t = 0; t < m.nf; t++) {
        const double* a = &m.V[(size_t)m.F[t * 3] * 3];
        const double* b = &m.V[(size_t)m.F[t * 3 + 1] * 3];
        const double* c = &m.V[(size_t)m.F[t * 3 + 2] * 3];
        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double r = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; k++) m.Nrm[(size_t)t * 3 + k] = (r == 0) ? 0.0 : n[k] / r;
    }
}

// PolygonMesh::updateForcedPolygonMesh (src/PolygonMesh.cpp:173-374): the
// reference's active law is the rotation about the y axis through
// (Xcg, Zcg) with steady amplitude (its ramp is off, :289-313); each node
// moves by omega x r dt with r taken before the move (:332-356).  The
// translation branch sets uniformVelWall = (0, 0, vel) (:322-330).
int forcedMeshUpdate(Sys& s, double dt, double time) {
    if (!s.extraOn[1]) { s.err = "no forced mesh"; return -1; }
    mps_params& P = s.p;
    Mesh& m = s.extra[1];
    const bool rotation = P.forced_motion == 1;
    double omegaAmp = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
    if (rotation) {
        const double kPi = 3.14159265358979323846;   // M_PI
        const double Tmotion = P.forced_period;
        const double R_ksi = P.forced_theta_max, R_ksi_dot = 0.0;
        omegaAmp = R_ksi_dot * sin(2.0 * kPi * time / Tmotion) + R_ksi * 2.0 * kPi / Tmotion * cos(2.0 * kPi * time / Tmotion);
        omegaAmp = omegaAmp * kPi / 180.0;
    } else {
        const double vel = P.forced_velocity;
        P.uniform_vel_wall[0] = 0.0;
        P.uniform_vel_wall[1] = 0.0;
        P.uniform_vel_wall[2] = vel;
        dz = vel * dt;
    }
    const double Xcg = P.forced_center[0], Zcg = P.forced_center[1];
    for (int q = 0; q < m.nv; q++) {
        double* v = &m.V[(size_t)q * 3];
        if (rotation) {
            const double xi = v[0] - Xcg, zi = v[2] - Zcg;
            const double vx = omegaAmp * zi, vz = -omegaAmp * xi;
            dx = vx * dt;
            dz = vz * dt;
        }
        v[0] += dx;
        v[1] += dy;
        v[2] += dz;
    }
    faceNormals(m);
    return 0;
}

// PolygonMesh::updateParticlesNearPolygonMesh (src/PolygonMesh.cpp:718-795)
void updateNearMesh(Sys& s) {
    const mps_params& P = s.p;
#pragma omp parallel for
    for (int i = 0; i < s.N; i++) {
        s.pndWall[i] = 0.0;
        s.numNeighWall[i] = 0;
        double x2 = s.riw2[i];
        if (x2 < P.re_l2) {
            s.nearWall[i] = 1;
            double x = sqrt(x2);
            s.pndWall[i] = interpolateWij(s.mesh, sqrt(P.re_s2), x);
            s.numNeighWall[i] = interpolateNumNeighWall(s.mesh, sqrt(P.re_l2), x);
        }
    }
}

// ---------------------------------------------------------- MpsPndNeigh
// setInitialPndNumberOfNeighNPCD (src/MpsPndNeigh.cpp:21-101)
void initialPnd(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        double wSum = 0.0;
        s.numNeigh[i] = 0;
        s.npcd[i * 3] = s.npcd[i * 3 + 1] = s.npcd[i * 3 + 2] = 0.0;
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_l2 && (d2 < dm2 || pw) && j != i) {
                s.numNeigh[i] += 1;
                if (d2 < P.re_s2) {
                    double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                    s.pndi[i] += wS;
                    s.npcd[i * 3] += v0 * wS;
                    s.npcd[i * 3 + 1] += v1 * wS;
                    s.npcd[i * 3 + 2] += v2 * wS;
                    wSum += wS;
                }
            }
        });
        s.pndi[i] += s.pndWall[i];
        if (s.type[i] == P.wall) s.pndi[i] = P.pnd_small_zero;
        s.pndSmall[i] = s.pndi[i];
        s.pndki[i] = s.pndi[i];
        s.numNeigh[i] += s.numNeighWall[i];
        if (wSum > kEps) {
            s.npcd[i * 3] /= s.pndSmall[i];
            s.npcd[i * 3 + 1] /= s.pndSmall[i];
            s.npcd[i * 3 + 2] /= s.pndSmall[i];
        }
        s.npcd2[i] = s.npcd[i * 3] * s.npcd[i * 3] + s.npcd[i * 3 + 1] * s.npcd[i * 3 + 1] + s.npcd[i * 3 + 2] * s.npcd[i * 3 + 2];
        if (s.npcd[i * 3] * s.polyN[i * 3] + s.npcd[i * 3 + 1] * s.polyN[i * 3 + 1] + s.npcd[i * 3 + 2] * s.polyN[i * 3 + 2] < 0.0)
            s.devDotN[i] = 1;
        else
            s.devDotN[i] = -1;
    }
}

// calcWallNPCD (src/MpsPndNeigh.cpp:105-195)
void wallNPCD(Sys& s) {
    const mps_params& P = s.p;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (!(s.type[i] == P.fluid && s.nearWall[i])) continue;
        double nx = 0, ny = 0, nz = 0;
        const double* xi = &s.pos[i * 3];
        const double* xm = &s.mirror[i * 3];
        double n[3], R[9];
        reflection(xi, xm, n, R, nullptr);
        walk(s, xi[0], xi[1], xi[2], [&](int j) {
            double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2 && j != i) {
                double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
                nx += w0 * wS; ny += w1 * wS; nz += w2 * wS;
            }
        });
        double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
        double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (dm2 < P.re_s2) {
            double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
            nx += w0 * wS; ny += w1 * wS; nz += w2 * wS;
        }
        s.npcd[i * 3] += R[0] * nx + R[1] * ny + R[2] * nz;
        s.npcd[i * 3 + 1] += R[3] * nx + R[4] * ny + R[5] * nz;
        s.npcd[i * 3 + 2] += R[6] * nx + R[7] * ny + R[8] * nz;
    }
}

// calcPndnNeighNPCD (src/MpsPndNeigh.cpp:199-331)
void pndNeighNPCD(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        double ni = 0.0, wSum = 0.0;
        int nn = s.numNeighWall[i];
        double c0 = s.npcd[i * 3], c1 = s.npcd[i * 3 + 1], c2 = s.npcd[i * 3 + 2];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_l2 && (d2 < dm2 || pw) && j != i) {
                nn += 1;
                if (d2 < P.re_s2) {
                    double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                    ni += wS;
                    c0 += v0 * wS; c1 += v1 * wS; c2 += v2 * wS;
                    wSum += wS;
                }
            }
        });
        s.numNeigh[i] = nn;
        if (P.pnd_type == MPS_PND_SUM_WIJ || P.pnd_type == MPS_PND_MEAN_SUM_WIJ) {
            s.pndki[i] = s.pndi[i];
            s.pndi[i] = ni + s.pndWall[i];
        }
        s.pndSmall[i] = ni + s.pndWall[i];
        if (wSum > kEps) {
            c0 /= s.pndSmall[i]; c1 /= s.pndSmall[i]; c2 /= s.pndSmall[i];
        }
        s.npcd[i * 3] = c0; s.npcd[i * 3 + 1] = c1; s.npcd[i * 3 + 2] = c2;
        s.npcd2[i] = c0 * c0 + c1 * c1 + c2 * c2;
        if (c0 * s.polyN[i * 3] + c1 * s.polyN[i * 3 + 1] + c2 * s.polyN[i * 3 + 2] < 0.0) s.devDotN[i] = 1;
        else s.devDotN[i] = -1;
    }
}

// calcPndDiffusiveTerm (src/MpsPndNeigh.cpp:334-489)
void pndDiffusive(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        double Di = 0.0, DivV = 0.0, flagDi = 1.0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double ux = s.vel[i * 3], uy = s.vel[i * 3 + 1], uz = s.vel[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        const double* MC0 = &s.cm1[i * 3]; const double* MC1 = &s.cm2[i * 3]; const double* MC2 = &s.cm3[i * 3];
        double ni = s.pndi[i];
        if (ni < kEps) continue;
        double Pi = s.press[i];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_l2 && (d2 < dm2 || pw) && j != i) {
                double dst = sqrt(d2);
                double wL = weight(dst, P.re_l, P.weight_type);
                if (s.type[i] == P.fluid && s.type[j] == P.fluid) {
                    double pgh = -2.0 * (s.RHO[i] * s.RHO[j] / (s.RHO[i] + s.RHO[j])) *
                                 (P.gravity[0] * v0 + P.gravity[1] * v1 + P.gravity[2] * v2);
                    double Pj = s.press[j];
                    Di += P.time_step / s.RHO[i] * P.coeff_visc_multiphase * (Pj - Pi - pgh) * wL;
                }
                if (d2 < P.re_s2) {
                    double vx = s.vel[j * 3] - ux, vy = s.vel[j * 3 + 1] - uy, vz = s.vel[j * 3 + 2] - uz;
                    double wS = weight(dst, P.re_s, P.weight_type);
                    if (!P.gradient_correction) {
                        if (ni > kEps)
                            DivV += (P.dim / P.pnd_small_zero) * (s.pndi[j] / ni) * (vx * v0 + vy * v1 + vz * v2) * wS / d2;
                    } else {
                        double c0 = (v0 * MC0[0] + v1 * MC0[1] + v2 * MC0[2]);
                        double c1 = (v0 * MC1[0] + v1 * MC1[1] + v2 * MC1[2]);
                        double c2 = (v0 * MC2[0] + v1 * MC2[1] + v2 * MC2[2]);
                        DivV += (P.dim / P.pnd_small_zero) * (vx * c0 + vy * c1 + vz * c2) * wS / d2;
                    }
                }
            }
        });
        s.aux1[i] = s.pndi[i] * (1.0 + P.time_step * (-DivV + Di * flagDi));
        s.divU[i] = DivV;
        s.diffTerm[i] = Di;
    }
}

// calcWallSlipPndDiffusiveTerm / calcWallNoSlipPndDiffusiveTerm (src/MpsPndNeigh.cpp:492-744)
void wallPndDiffusive(Sys& s) {
    const mps_params& P = s.p;
    const bool slip = P.slip_condition == MPS_FREE_SLIP;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        double ni = s.pndi[i];
        if (!(s.type[i] == P.fluid && s.nearWall[i] && ni > kEps)) continue;
        double DivV = 0.0;
        const double* xi = &s.pos[i * 3];
        const double* xm = &s.mirror[i * 3];
        const double ux = s.vel[i * 3], uy = s.vel[i * 3 + 1], uz = s.vel[i * 3 + 2];
        double n[3], R[9];
        reflectionMod2(xi, xm, n, R);
        double umx, umy, umz;
        if (slip) {
            umx = R[0] * ux + R[1] * uy + R[2] * uz;
            umy = R[3] * ux + R[4] * uy + R[5] * uz;
            umz = R[6] * ux + R[7] * uy + R[8] * uz;
        } else {
            double vw[3] = {0.0, 0.0, 0.0};   // wall velocity: uniformVelWall near the forced mesh
            if (s.nearMeshType[i] == 2) for (int k = 0; k < 3; k++) vw[k] = P.uniform_vel_wall[k];
            double dotnv = n[0] * vw[0] + n[1] * vw[1] + n[2] * vw[2];
            double vt0 = ux - 2.0 * (vw[0] - dotnv * n[0]);
            double vt1 = uy - 2.0 * (vw[1] - dotnv * n[1]);
            double vt2 = uz - 2.0 * (vw[2] - dotnv * n[2]);
            umx = -1.0 * vt0 + 0.0 * vt1 + 0.0 * vt2;
            umy = 0.0 * vt0 + -1.0 * vt1 + 0.0 * vt2;
            umz = 0.0 * vt0 + 0.0 * vt1 + -1.0 * vt2;
        }
        walk(s, xi[0], xi[1], xi[2], [&](int j) {
            double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2 && j != i) {
                double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
                if (slip) {
                    double vx = s.vel[j * 3] - umx, vy = s.vel[j * 3 + 1] - umy, vz = s.vel[j * 3 + 2] - umz;
                    DivV += (P.dim / P.pnd_small_zero) * (s.pndi[j] / ni) * (vx * w0 + vy * w1 + vz * w2) * wS / dm2;
                } else {
                    double vx = -(s.vel[j * 3] - umx), vy = -(s.vel[j * 3 + 1] - umy), vz = -(s.vel[j * 3 + 2] - umz);
                    double m0 = (R[0] * w0 + R[1] * w1 + R[2] * w2);
                    double m1 = (R[3] * w0 + R[4] * w1 + R[5] * w2);
                    double m2 = (R[6] * w0 + R[7] * w1 + R[8] * w2);
                    DivV += (P.dim / P.pnd_small_zero) * (s.pndi[j] / ni) * (vx * m0 + vy * m1 + vz * m2) * wS / dm2;
                }
            }
        });
        double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
        double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (dm2 < P.re_s2) {
            double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
            if (slip) {
                double vx = ux - umx, vy = uy - umy, vz = uz - umz;
                DivV += (P.dim / P.pnd_small_zero) * (s.pndi[i] / ni) * (vx * w0 + vy * w1 + vz * w2) * wS / dm2;
            } else {
                double vx = -(ux - umx), vy = -(uy - umy), vz = -(uz - umz);
                double m0 = (R[0] * w0 + R[1] * w1 + R[2] * w2);
                double m1 = (R[3] * w0 + R[4] * w1 + R[5] * w2);
                double m2 = (R[6] * w0 + R[7] * w1 + R[8] * w2);
                DivV += (P.dim / P.pnd_small_zero) * (s.pndi[i] / ni) * (vx * m0 + vy * m1 + vz * m2) * wS / dm2;
            }
        }
        s.aux1[i] += -s.pndi[i] * P.time_step * DivV;
    }
}

// updatePnd (src/MpsPndNeigh.cpp:747-774)
void updatePnd(Sys& s) {
    for (int i = 0; i < s.N; i++) {
        s.pndi[i] = s.aux1[i];
        s.aux1[i] = 0.0;
    }
}

// meanPndParticlesWallDummySurface (src/MpsPndNeigh.cpp:777-852)
void meanPndWallDummySurface(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        s.aux1[i] = 0.0; s.aux2[i] = 0.0;
        if (!(s.type[i] == P.wall || s.bc[i] == P.surface)) continue;
        double up = 0, dn = 0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i) {
                double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                up += s.pndi[j] * wS;
                dn += wS;
            }
        });
        s.aux1[i] = up; s.aux2[i] = dn;
    }
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] == P.wall || s.bc[i] == P.surface) {
            if (s.aux2[i] < kEps) s.pndi[i] = s.aux1[i];
            else s.pndi[i] = s.aux1[i] / (s.aux2[i]);
        }
        s.aux1[i] = 0.0; s.aux2[i] = 0.0;
    }
}

// meanPnd (src/MpsPndNeigh.cpp:855-920)
void meanPnd(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        double up = 0, dn = 0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i) {
                double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                up += s.pndi[j] * wS;
                dn += wS;
            }
        });
        s.aux1[i] += up;
        s.aux2[i] += dn;
    }
    for (int i = 0; i < s.N; i++) {
        if (s.numNeigh[i] < 1) s.pndi[i] = s.aux1[i];
        else s.pndi[i] = s.aux1[i] / s.aux2[i];
        s.aux1[i] = 0.0; s.aux2[i] = 0.0;
    }
}

// meanWallPnd (src/MpsPndNeigh.cpp:923-994)
void meanWallPnd(Sys& s) {
    const mps_params& P = s.p;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        s.aux1[i] = 0.0; s.aux2[i] = 0.0;
        if (!s.nearWall[i]) continue;
        double up = 0, dn = 0;
        const double* xi = &s.pos[i * 3];
        const double* xm = &s.mirror[i * 3];
        walk(s, xi[0], xi[1], xi[2], [&](int j) {
            double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2 && j != i) {
                double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
                up += s.pndi[j] * wS;
                dn += wS;
            }
        });
        double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
        double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (dm2 < P.re_s2) {
            double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
            up += s.pndi[i] * wS;
            dn += wS;
        }
        s.aux1[i] = up; s.aux2[i] = dn;
    }
}

// meanNeighFluidPnd (src/MpsPndNeigh.cpp:997-1050)
void meanNeighFluidPnd(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        double up = 0, dn = 0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i && s.type[j] == P.fluid) {
                double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                up += s.pndki[j] * wS;
                dn += wS;
            }
        });
        if (dn > kEps) s.pndski[i] = up / dn;
        else s.pndski[i] = up;
    }
}

// ----------------------------------------------------- ARC normals
// calcNormalParticles (src/MpsShifting.cpp:106-175): n_i = -sum_j r_ij w / |r_ij| / n0S
void normalParticles(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        double dx = 0.0, dy = 0.0, dz = 0.0;
        walk(s, xi, yi, zi, [&](int j) {
            const double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i) {
                const double dst = sqrt(d2);
                const double wS = weight(dst, P.re_s, P.weight_type);
                dx += v0 * wS / dst;
                dy += v1 * wS / dst;
                dz += v2 * wS / dst;
            }
        });
        s.normal[i * 3] = -dx / P.pnd_small_zero;
        s.normal[i * 3 + 1] = -dy / P.pnd_small_zero;
        s.normal[i * 3 + 2] = -dz / P.pnd_small_zero;
    }
}

// calcWallNormalParticles (src/MpsShifting.cpp:319-430): the mirror
// neighbourhood's part, reflected back by R, for near-wall fluid particles
void wallNormalParticles(Sys& s) {
    const mps_params& P = s.p;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (!(s.type[i] == P.fluid && s.nearWall[i])) continue;
        const double* xi = &s.pos[i * 3];
        const double* xm = &s.mirror[i * 3];
        double n[3], R[9], ns;
        reflection(xi, xm, n, R, &ns);
        double dx = 0.0, dy = 0.0, dz = 0.0;
        walk(s, xi[0], xi[1], xi[2], [&](int j) {
            const double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2 && j != i) {
                const double dst = sqrt(dm2);
                const double wS = weight(dst, P.re_s, P.weight_type);
                dx += w0 * wS / dst;
                dy += w1 * wS / dst;
                dz += w2 * wS / dst;
            }
        });
        const double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (dm2 < P.re_s2) {
            const double dst = sqrt(dm2);
            const double wS = weight(dst, P.re_s, P.weight_type);
            dx += w0 * wS / dst;
            dy += w1 * wS / dst;
            dz += w2 * wS / dst;
        }
        const double rx = R[0] * dx + R[1] * dy + R[2] * dz;
        const double ry = R[3] * dx + R[4] * dy + R[5] * dz;
        const double rz = R[6] * dx + R[7] * dy + R[8] * dz;
        s.normal[i * 3] += -rx / P.pnd_small_zero;
        s.normal[i * 3 + 1] += -ry / P.pnd_small_zero;
        s.normal[i * 3 + 2] += -rz / P.pnd_small_zero;
    }
}

// ----------------------------------------------- MpsBoundaryCondition
// The ARC rule of updateParticleBC (src/MpsBoundaryCondition.cpp:205-392): a
// particle is inner if, in walk order, a neighbour j (or, with polygon walls,
// j seen from the mirror point) either lies far (r >= sqrt(2) 1.33 l0) but
// close to the point one l0 along the normal, or near with an angle to the
// normal below thetaArc; the first such j ends the walk.  Otherwise the last
// neighbour examined leaves the particle at the surface.
int arcRule(const Sys& s, int i, int bc) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
    double nx = s.normal[i * 3], ny = s.normal[i * 3 + 1], nz = s.normal[i * 3 + 2];
    const double norm2 = nx * nx + ny * ny + nz * nz;
    if (s.pndi[i] >= P.beta_pnd && s.numNeigh[i] >= P.beta_neigh && norm2 <= P.norm_threshold2) return P.inner;
    const double* xi = &s.pos[i * 3];
    const double* xm = &s.mirror[i * 3];
    if (norm2 > kEps) {
        const double nn = sqrt(norm2);
        nx /= nn; ny /= nn; nz /= nn;
    } else {
        nx = 0.0; ny = 0.0; nz = 0.0;
    }
    double n[3], R[9];
    reflectionMod2(xi, xm, n, R);
    double mx = R[0] * nx + R[1] * ny + R[2] * nz;
    double my = R[3] * nx + R[4] * ny + R[5] * nz;
    double mz = R[6] * nx + R[7] * ny + R[8] * nz;
    const double mn2 = mx * mx + my * my + mz * mz;
    if (mn2 > kEps) {
        const double mn = sqrt(mn2);
        mx /= mn; my /= mn; mz /= mn;
    } else {
        mx = 0.0; my = 0.0; mz = 0.0;
    }
    const double l0 = P.part_dist;
    const double ax = xi[0] + l0 * nx, ay = xi[1] + l0 * ny, az = xi[2] + l0 * nz;
    const double bx = xm[0] + l0 * mx, by = xm[1] + l0 * my, bz = xm[2] + l0 * mz;
    bool done = false;
    walk(s, xi[0], xi[1], xi[2], [&](int j) {
        if (done) return;
        const double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i) {
            const double u0 = PJ(s, j, 0) - ax, u1 = PJ(s, j, 1) - ay, u2 = PJ(s, j, 2) - az;
            const double dn2 = u0 * u0 + u1 * u1 + u2 * u2;
            const double rn = v0 * nx + v1 * ny + v2 * nz;
            const double ang = acos(rn / sqrt(d2));
            if (d2 >= P.dst_threshold2 && dn2 < P.h_threshold2) { bc = P.inner; done = true; return; }
            else if (d2 < P.dst_threshold2 && ang < P.theta_arc) { bc = P.inner; done = true; return; }
            else bc = P.surface;
        }
        if (!pw && d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2 && j != i) {
            const double u0 = PJ(s, j, 0) - bx, u1 = PJ(s, j, 1) - by, u2 = PJ(s, j, 2) - bz;
            const double dn2 = u0 * u0 + u1 * u1 + u2 * u2;
            const double rn = w0 * mx + w1 * my + w2 * mz;
            const double ang = acos(rn / sqrt(dm2));
            if (dm2 >= P.dst_threshold2 && dn2 < P.h_threshold2) { bc = P.inner; done = true; return; }
            else if (dm2 < P.dst_threshold2 && ang < P.theta_arc) { bc = P.inner; done = true; return; }
            else bc = P.surface;
        }
    });
    return bc;
}

// updateParticleBC (src/MpsBoundaryCondition.cpp:66-400): PND / NEIGH / NPCD / ARC rules
int updateBC(Sys& s) {
    const mps_params& P = s.p;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (s.type[i] == P.dummy_wall) { s.bc[i] = P.other; continue; }
        if (s.pndSmall[i] < P.beta_pnd) s.bc[i] = P.surface;
        else s.bc[i] = P.inner;
        if (P.free_surf_type == MPS_FS_PND_NEIGH) {
            if (s.pndSmall[i] < P.beta_pnd && s.numNeigh[i] < P.beta_neigh) s.bc[i] = P.surface;
            else s.bc[i] = P.inner;
        } else if (P.free_surf_type == MPS_FS_PND_NPCD) {
            if (s.bc[i] == P.surface)
                if (s.numNeigh[i] > 4 && s.npcd2[i] < P.delta2) s.bc[i] = P.inner;
        } else if (P.free_surf_type == MPS_FS_PND_ARC) {
            s.bc[i] = arcRule(s, i, s.bc[i]);
        }
    }
    return 0;
}

// ------------------------------------------------------ MpsVectorMatrix
// detMatrix / inverseMatrix (src/MpsVectorMatrix.cpp:96-165)
int inverse3(double* M, int dim) {
    double A[3][3], B[3][3];
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) A[r][c] = M[r * 3 + c];
    if (dim == 2) { A[0][2] = A[1][2] = A[2][0] = A[2][1] = 0.0; A[2][2] = 1.0; }
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) B[r][c] = (r == c) ? 1.0 : 0.0;
    double det = (A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1]) -
                 (A[0][2] * A[1][1] * A[2][0] + A[0][1] * A[1][0] * A[2][2] + A[0][0] * A[1][2] * A[2][1]);
    if (det <= kEps) {
        for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) M[r * 3 + c] = (r == c) ? 1.0 : 0.0;
        return 0;
    }
    for (int k = 0; k < dim; k++)
        for (int r = 0; r < dim; r++) {
            if (r == k) continue;
            double m = A[r][k] / A[k][k];
            for (int c = 0; c < dim; c++) {
                A[r][c] -= m * A[k][c];
                B[r][c] -= m * B[k][c];
            }
        }
    for (int r = 0; r < dim; r++) for (int c = 0; c < dim; c++) B[r][c] /= A[r][r];
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) M[r * 3 + c] = B[r][c];
    return 1;
}

// correctionMatrix (src/MpsVectorMatrix.cpp:20-93)
void correctionMatrix(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i) {
                double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                double inv = 1.0 / d2;
                M[0] += wS * v0 * v0 * inv; M[1] += wS * v0 * v1 * inv; M[2] += wS * v0 * v2 * inv;
                M[3] += wS * v1 * v0 * inv; M[4] += wS * v1 * v1 * inv; M[5] += wS * v1 * v2 * inv;
                M[6] += wS * v2 * v0 * inv; M[7] += wS * v2 * v1 * inv; M[8] += wS * v2 * v2 * inv;
            }
        });
        for (int k = 0; k < 9; k++) M[k] *= -P.coeff_press_grad;
        inverse3(M, (int)P.dim);
        if (s.numNeigh[i] < 4) {
            for (int k = 0; k < 9; k++) M[k] = (k % 4 == 0) ? 1.0 : 0.0;
        }
        for (int k = 0; k < 3; k++) {
            s.cm1[i * 3 + k] = M[k]; s.cm2[i * 3 + k] = M[3 + k]; s.cm3[i * 3 + k] = M[6 + k];
        }
    }
}

// ----------------------------------------------------------- MpsPressure
// calcPressEMPS / calcPressWCMPS (src/MpsPressure.cpp:62-163)
void pressExplicit(Sys& s) {
    const mps_params& P = s.p;
#pragma omp parallel for
    for (int i = 0; i < s.N; i++) {
        double mi = (s.ptype[i] == 1) ? P.dns_fl1 : P.dns_fl2;
        double pr = 0.0;
        if (s.bc[i] == P.inner) {
            if (P.mps_type == MPS_EXPLICIT) pr = (s.pndi[i] - P.pnd_small_zero) * P.coeff_press_emps * mi;
            else pr = (mi * P.coeff_press_wcmps / P.gamma) * (pow(s.pndi[i] / P.pnd_small_zero, P.gamma) - 1);
        }
        if (pr < 0.0) pr = 0.0;
        s.press[i] = pr;
    }
}

// calcVelDivergence (src/MpsPressure.cpp:479-551)
void velDivergence(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        s.divU[i] = 0.0;
        if (s.type[i] != P.fluid) continue;
        double DivV = 0.0, ni = s.pndi[i];
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double ux = s.vel[i * 3], uy = s.vel[i * 3 + 1], uz = s.vel[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        const double* MC0 = &s.cm1[i * 3]; const double* MC1 = &s.cm2[i * 3]; const double* MC2 = &s.cm3[i * 3];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i) {
                double vx = s.vel[j * 3] - ux, vy = s.vel[j * 3 + 1] - uy, vz = s.vel[j * 3 + 2] - uz;
                double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                if (!P.gradient_correction) {
                    if (ni > kEps)
                        DivV += (P.dim / P.pnd_small_zero) * (s.pndi[j] / ni) * (vx * v0 + vy * v1 + vz * v2) * wS / d2;
                } else {
                    double c0 = (v0 * MC0[0] + v1 * MC0[1] + v2 * MC0[2]);
                    double c1 = (v0 * MC1[0] + v1 * MC1[1] + v2 * MC1[2]);
                    double c2 = (v0 * MC2[0] + v1 * MC2[1] + v2 * MC2[2]);
                    DivV += (P.dim / P.pnd_small_zero) * (vx * c0 + vy * c1 + vz * c2) * wS / d2;
                }
            }
        });
        s.divU[i] = DivV;
    }
}

// calcWallSlipVelDivergence / calcWallNoSlipVelDivergence
// (src/MpsPressure.cpp:554-657, 658-788): the mirror neighbourhood of a
// near-wall fluid particle adds to its velocity divergence.  The mirror
// velocity is R v (free slip) or -(v - 2 (v_wall - (n.v_wall) n)) with a
// fixed wall (no slip, where the relative position is reflected by R too).
void wallVelDivergence(Sys& s) {
    const mps_params& P = s.p;
    const bool slip = P.slip_condition == MPS_FREE_SLIP;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        const double ni = s.pndi[i];
        if (!(s.type[i] == P.fluid && s.nearWall[i] && ni > kEps)) continue;
        double DivV = 0.0;
        const double* xi = &s.pos[i * 3];
        const double* xm = &s.mirror[i * 3];
        const double velXi = s.vel[i * 3], velYi = s.vel[i * 3 + 1], velZi = s.vel[i * 3 + 2];
        double n[3], R[9];
        reflectionMod2(xi, xm, n, R);
        double vmx, vmy, vmz;
        if (slip) {
            vmx = (R[0] * velXi + R[1] * velYi + R[2] * velZi);
            vmy = (R[3] * velXi + R[4] * velYi + R[5] * velZi);
            vmz = (R[6] * velXi + R[7] * velYi + R[8] * velZi);
        } else {
            const double Rinv[9] = {-1.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0, -1.0};
            double vw[3] = {0.0, 0.0, 0.0};   // wall velocity: uniformVelWall near the forced mesh
            if (s.nearMeshType[i] == 2) for (int k = 0; k < 3; k++) vw[k] = P.uniform_vel_wall[k];
            const double dotnv = n[0] * vw[0] + n[1] * vw[1] + n[2] * vw[2];
            const double vt[3] = {velXi - 2.0 * (vw[0] - dotnv * n[0]), velYi - 2.0 * (vw[1] - dotnv * n[1]),
                                  velZi - 2.0 * (vw[2] - dotnv * n[2])};
            vmx = (Rinv[0] * vt[0] + Rinv[1] * vt[1] + Rinv[2] * vt[2]);
            vmy = (Rinv[3] * vt[0] + Rinv[4] * vt[1] + Rinv[5] * vt[2]);
            vmz = (Rinv[6] * vt[0] + Rinv[7] * vt[1] + Rinv[8] * vt[2]);
        }
        auto term = [&](double pj, double ux, double uy, double uz, double w0, double w1, double w2, double dm2) {
            const double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
            if (slip) {
                const double vx = ux - vmx, vy = uy - vmy, vz = uz - vmz;
                DivV += (P.dim / P.pnd_small_zero) * (pj / ni) * (vx * w0 + vy * w1 + vz * w2) * wS / dm2;
            } else {
                const double vx = -(ux - vmx), vy = -(uy - vmy), vz = -(uz - vmz);
                const double m0 = (R[0] * w0 + R[1] * w1 + R[2] * w2);
                const double m1 = (R[3] * w0 + R[4] * w1 + R[5] * w2);
                const double m2 = (R[6] * w0 + R[7] * w1 + R[8] * w2);
                DivV += (P.dim / P.pnd_small_zero) * (pj / ni) * (vx * m0 + vy * m1 + vz * m2) * wS / dm2;
            }
        };
        walk(s, xi[0], xi[1], xi[2], [&](int j) {
            const double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2 && j != i)
                term(s.pndi[j], s.vel[j * 3], s.vel[j * 3 + 1], s.vel[j * 3 + 2], w0, w1, w2, dm2);
        });
        // i is a neighbour of its own mirror
        const double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (dm2 < P.re_s2) term(s.pndi[i], velXi, velYi, velZi, w0, w1, w2, dm2);
        s.divU[i] += DivV;
    }
}

// PPE assembly: solvePressurePoissonPnd / ...PndDivU (src/MpsPressure.cpp:169-421).
// Rows are stored in triplet order (neighbours in walk order, then the
// diagonal); identity rows for surface / other particles.
void assemblePPE(Sys& s) {
    const mps_params& P = s.p;
    s.asmCount++;
    const bool pw = partWall(s);
    const bool divu = P.mps_type == MPS_IMPLICIT_PND_DIVU;
    s.ppeB.assign(s.N, 0.0);
    s.rowPtr.assign(s.N + 1, 0);
    s.colIdx.clear();
    s.vals.clear();
    for (int i = 0; i < s.N; i++) {
        s.rowPtr[i] = (int)s.colIdx.size();
        if (s.bc[i] == P.other || s.bc[i] == P.surface) {
            s.colIdx.push_back(i); s.vals.push_back(1.0);
            continue;
        }
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        double ni = s.pndSmall[i];
        double sum = 0.0;
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_l2 && (d2 < dm2 || pw) && j != i) {
                double wL = weight(sqrt(d2), P.re_l, P.weight_type);
                double mij = wL * P.coeff_ppe;
                if (s.type[j] == P.dummy_wall) {
                    double pgh = s.RHO[j] * (v0 * P.gravity[0] + v1 * P.gravity[1] + v2 * P.gravity[2]) * wL;
                    s.ppeB[i] -= P.coeff_ppe * pgh;
                } else {
                    sum -= mij;
                    if (s.bc[j] == P.inner) { s.colIdx.push_back(j); s.vals.push_back(mij); }
                }
            }
        });
        double density = (s.ptype[i] == 1) ? P.dns_fl1 : P.dns_fl2;
        sum -= P.alpha_compressibility * density / (P.time_step * P.time_step);
        s.colIdx.push_back(i); s.vals.push_back(sum);
        if (!divu)
            s.ppeB[i] += -P.coeff_ppe_source * density * (ni - P.pnd_small_zero);
        else
            s.ppeB[i] += -P.coeff_ppe_source * density * (s.pndki[i] - P.pnd_small_zero) +
                         (1.0 - P.relax_pnd) * density * s.divU[i] / P.time_step;
    }
    s.rowPtr[s.N] = (int)s.colIdx.size();
}

// Restated Eigen ConjugateGradient with DiagonalPreconditioner and a guess.
// The matrix used for iteration is A^T when ppe_transpose (Eigen applies the
// row-major transpose wrapper for UpLo = Lower|Upper on a column-major
// matrix), else A.  Products accumulate in ascending column index of the
// iterated matrix, as a compressed sparse traversal does.
// The iterated matrix in row form with ascending columns: A^T (what
// Eigen's Lower|Upper CG uses on the row-filled matrix) or A.
static double wallSeconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void iterMatrix(const Sys& s, bool transpose, std::vector<int>& rp, std::vector<int>& ci, std::vector<double>& va,
                std::vector<double>& diag) {
    const int n = s.N;
    rp.assign(n + 1, 0);
    ci.clear();
    va.clear();
    diag.assign(n, 0.0);
    std::vector<std::vector<std::pair<int, double>>> rows(n);
    for (int r = 0; r < n; r++)
        for (int k = s.rowPtr[r]; k < s.rowPtr[r + 1]; k++) {
            int c = s.colIdx[k];
            double v = s.vals[k];
            if (c == r) diag[r] = v;
            if (transpose) rows[c].push_back({r, v});
            else rows[r].push_back({c, v});
        }
    for (int r = 0; r < n; r++) {
        std::stable_sort(rows[r].begin(), rows[r].end(),
                         [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.first < b.first; });
        rp[r + 1] = rp[r] + (int)rows[r].size();
        for (auto& e : rows[r]) { ci.push_back(e.first); va.push_back(e.second); }
    }
}

// Eigen BiCGSTAB with its default DiagonalPreconditioner, restated from the
// published algorithm (Eigen/src/IterativeLinearSolvers/BiCGSTAB.h, not in
// the reference tree): restart on rho below eps^2 |r0|^2 (the first restart
// also resets the iteration count), stop on |r|^2 <= tol^2 |b|^2 or 2N
// iterations; iterates with A itself (MpsPressure::
// solveBiConjugateGradientStabilized, src/MpsPressure.cpp:448-468).
void solveBiCGSTAB(Sys& s) {
    const mps_params& P = s.p;
    const int n = s.N;
    std::vector<int> rp, ci;
    std::vector<double> va, diag;
    iterMatrix(s, false, rp, ci, va, diag);
    auto spmv = [&](const std::vector<double>& x, std::vector<double>& y) {
#pragma omp parallel for
        for (int r = 0; r < n; r++) {
            double acc = 0.0;
            for (int k = rp[r]; k < rp[r + 1]; k++) acc += va[k] * x[ci[k]];
            y[r] = acc;
        }
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
        double acc = 0.0;
        for (int r = 0; r < n; r++) acc += a[r] * b[r];
        return acc;
    };
    std::vector<double> invd(n);
    for (int r = 0; r < n; r++) invd[r] = (diag[r] != 0.0) ? 1.0 / diag[r] : 1.0;
    std::vector<double>& x = s.ppeX;
    if ((int)x.size() < n) x.resize(n, 0.0);
    const std::vector<double>& b = s.ppeB;
    const double tol = P.solver_tol > 0 ? P.solver_tol : DBL_EPSILON;
    const int maxIt = P.solver_max_iter > 0 ? P.solver_max_iter : 2 * n;
    std::vector<double> r(n), r0(n), tmp(n), v(n, 0.0), p(n, 0.0), y(n), z(n), sv(n), t(n);
    spmv(x, tmp);
    for (int k = 0; k < n; k++) r[k] = b[k] - tmp[k];
    r0 = r;
    double r0n2 = dot(r0, r0);
    const double rhsN2 = dot(b, b);
    if (rhsN2 == 0) {
        for (int k = 0; k < n; k++) x[k] = 0.0;
        s.solverIter = 0; s.solverErr = 0.0;
        return;
    }
    double rho = 1.0, alpha = 1.0, w = 1.0;
    const double tol2 = tol * tol * rhsN2;
    const double eps2 = DBL_EPSILON * DBL_EPSILON;
    int i = 0, restarts = 0;
    while (dot(r, r) > tol2 && i < maxIt) {
        const double rhoOld = rho;
        rho = dot(r0, r);
        if (std::fabs(rho) < eps2 * r0n2) {
            spmv(x, tmp);
            for (int k = 0; k < n; k++) r[k] = b[k] - tmp[k];
            r0 = r;
            rho = r0n2 = dot(r, r);
            if (restarts++ == 0) i = 0;
        }
        const double beta = (rho / rhoOld) * (alpha / w);
        for (int k = 0; k < n; k++) p[k] = r[k] + beta * (p[k] - w * v[k]);
        for (int k = 0; k < n; k++) y[k] = invd[k] * p[k];
        spmv(y, v);
        alpha = rho / dot(r0, v);
        for (int k = 0; k < n; k++) sv[k] = r[k] - alpha * v[k];
        for (int k = 0; k < n; k++) z[k] = invd[k] * sv[k];
        spmv(z, t);
        const double tt = dot(t, t);
        w = tt > 0.0 ? dot(t, sv) / tt : 0.0;
        for (int k = 0; k < n; k++) x[k] += alpha * y[k] + w * z[k];
        for (int k = 0; k < n; k++) r[k] = sv[k] - w * t[k];
        ++i;
    }
    s.solverIter = i;
    s.solverErr = sqrt(dot(r, r) / rhsN2);
}

void solveCG(Sys& s) {
    const mps_params& P = s.p;
    const int n = s.N;
    // the iteration matrix of the last assembly is built once and kept
    const double tSetup = wallSeconds();
    const int want = P.ppe_transpose != 0 ? 1 : 0;
    if (s.itAsm != s.asmCount || s.itTrans != want) {
        iterMatrix(s, want != 0, s.itRp, s.itCi, s.itVa, s.itDiag);
        s.itAsm = s.asmCount;
        s.itTrans = want;
    }
    const std::vector<int>& rp = s.itRp;
    const std::vector<int>& ci = s.itCi;
    const std::vector<double>& va = s.itVa;
    const std::vector<double>& diag = s.itDiag;
    const double tIter = wallSeconds();
    s.solveSetupSeconds = tIter - tSetup;
    auto spmv = [&](const std::vector<double>& x, std::vector<double>& y) {
#pragma omp parallel for
        for (int r = 0; r < n; r++) {
            double acc = 0.0;
            for (int k = rp[r]; k < rp[r + 1]; k++) acc += va[k] * x[ci[k]];
            y[r] = acc;
        }
    };
    const bool par = s.parallelCg;
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& b) {
        double acc = 0.0;
        if (par) {
#pragma omp parallel for reduction(+ : acc)
            for (int r = 0; r < n; r++) acc += a[r] * b[r];
        } else {
            for (int r = 0; r < n; r++) acc += a[r] * b[r];
        }
        return acc;
    };
    std::vector<double> invd(n);
    for (int r = 0; r < n; r++) invd[r] = (diag[r] != 0.0) ? 1.0 / diag[r] : 1.0;
    std::vector<double>& x = s.ppeX;
    if ((int)x.size() < n) x.resize(n, 0.0);
    const std::vector<double>& b = s.ppeB;
    double tol = P.solver_tol > 0 ? P.solver_tol : DBL_EPSILON;
    int maxIt = P.solver_max_iter > 0 ? P.solver_max_iter : 2 * n;
    std::vector<double> r(n), p(n), z(n), tmp(n);
    spmv(x, tmp);
    for (int k = 0; k < n; k++) r[k] = b[k] - tmp[k];
    double rhsN2 = dot(b, b);
    if (rhsN2 == 0) {
        for (int k = 0; k < n; k++) x[k] = 0.0;
        s.solverIter = 0; s.solverErr = 0.0;
        return;
    }
    double thr = std::max(tol * tol * rhsN2, DBL_MIN);
    double rN2 = dot(r, r);
    if (rN2 < thr) {
        s.solverIter = 0; s.solverErr = sqrt(rN2 / rhsN2);
        return;
    }
    for (int k = 0; k < n; k++) p[k] = invd[k] * r[k];
    double absNew = dot(r, p);
    int it = 0;
    while (it < maxIt) {
        spmv(p, tmp);
        double alpha = absNew / dot(p, tmp);
#pragma omp parallel for if (par)
        for (int k = 0; k < n; k++) {
            x[k] += alpha * p[k];
            r[k] -= alpha * tmp[k];
        }
        rN2 = dot(r, r);
        if (rN2 < thr) break;
#pragma omp parallel for if (par)
        for (int k = 0; k < n; k++) z[k] = invd[k] * r[k];
        double absOld = absNew;
        absNew = dot(r, z);
        double beta = absNew / absOld;
#pragma omp parallel for if (par)
        for (int k = 0; k < n; k++) p[k] = z[k] + beta * p[k];
        it++;
    }
    s.solverErr = sqrt(rN2 / rhsN2);
    s.solverIter = it;
    s.solveIterSeconds = wallSeconds() - tIter;
}

int calcPress(Sys& s) {
    const mps_params& P = s.p;
    if (P.mps_type == MPS_EXPLICIT || P.mps_type == MPS_WEAKLY) {
        pressExplicit(s);
        return 0;
    }
    if (P.mps_type == MPS_IMPLICIT_PND_DIVU) {
        velDivergence(s);
        if (P.wall_type == MPS_WALL_POLYGON) wallVelDivergence(s);
    }
    assemblePPE(s);
    if (P.solver_type == MPS_SOLVER_BICGSTAB) solveBiCGSTAB(s);
    else solveCG(s);
    for (int i = 0; i < s.N; i++) s.press[i] = s.ppeX[i];
    for (int i = 0; i < s.N; i++) if (s.press[i] < 0) s.press[i] = 0.0;
    return 0;
}

// extrapolatePressParticlesWallDummy (src/MpsPressure.cpp:792-846)
void extrapolateWallDummy(Sys& s) {
    const mps_params& P = s.p;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (!(s.type[i] == P.dummy_wall || (P.mps_type == MPS_WEAKLY && s.type[i] == P.wall))) continue;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        double ni = 0.0, pr = 0.0;
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            if (d2 < P.re_s2 && j != i && s.type[j] == P.fluid) {
                double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                ni += wS;
                pr += (s.press[j] - s.RHO[j] * (P.gravity[0] * v0 + P.gravity[1] * v1 + P.gravity[2] * v2)) * wS;
            }
        });
        if (pr < 0.0) continue;
        if (ni > 0) s.press[i] = pr / ni;
        else s.press[i] = pr;
    }
}

// calcPressGradient (src/MpsPressure.cpp:1217-1338)
void pressGradient(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
    const int gt = P.gradient_type;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        double ax = 0, ay = 0, az = 0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        double pmin = s.press[i], Pi = s.press[i], ni = s.pndi[i];
        const double* MC0 = &s.cm1[i * 3]; const double* MC1 = &s.cm2[i * 3]; const double* MC2 = &s.cm3[i * 3];
        auto ok = [&](int j, double& v0, double& v1, double& v2, double& d2) {
            v0 = PJ(s, j, 0) - xi; v1 = PJ(s, j, 1) - yi; v2 = PJ(s, j, 2) - zi;
            d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            return d2 < P.re_s2 && (d2 < dm2 || pw) && j != i;
        };
        if (gt == 0 || gt == 2)
            walk(s, xi, yi, zi, [&](int j) {
                double v0, v1, v2, d2;
                if (ok(j, v0, v1, v2, d2))
                    if (pmin > s.press[j]) pmin = s.press[j];
            });
        walk(s, xi, yi, zi, [&](int j) {
            double v0, v1, v2, d2;
            if (!ok(j, v0, v1, v2, d2)) return;
            double wS = weight(sqrt(d2), P.re_s, P.weight_type);
            if (gt == 0) wS *= (s.press[j] - pmin) / d2;
            else if (gt == 1) wS *= (s.press[j] + Pi) / d2;
            else if (gt == 2) wS *= (s.press[j] + Pi - 2.0 * pmin) / d2;
            else if (gt == 3) {
                double nj = s.pndi[j];
                if (ni > kEps && nj > kEps) wS *= (ni * s.press[j] / nj + nj * Pi / ni) / d2;
            }
            if (!P.gradient_correction) {
                ax += v0 * wS; ay += v1 * wS; az += v2 * wS;
            } else {
                ax += (v0 * MC0[0] + v1 * MC0[1] + v2 * MC0[2]) * wS;
                ay += (v0 * MC1[0] + v1 * MC1[1] + v2 * MC1[2]) * wS;
                az += (v0 * MC2[0] + v1 * MC2[1] + v2 * MC2[2]) * wS;
            }
        });
        s.acc[i * 3] = P.relax_press * ax * P.coeff_press_grad / s.RHO[i];
        s.acc[i * 3 + 1] = P.relax_press * ay * P.coeff_press_grad / s.RHO[i];
        s.acc[i * 3 + 2] = P.relax_press * az * P.coeff_press_grad / s.RHO[i];
    }
}

// predictionPressGradient (src/MpsPressure.cpp:932-1028): the (1 - relax)
// part of the pressure gradient before the predictor, without the gradient
// correction, added to acc of every particle
void predictionPressGradient(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
    const int gt = P.gradient_type;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        double accX = 0.0, accY = 0.0, accZ = 0.0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        const double Pi = s.press[i], ni = s.pndi[i];
        double pressMin = Pi;
        if (gt == 0 || gt == 2)
            walk(s, xi, yi, zi, [&](int j) {
                const double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
                const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
                const double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
                const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
                if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i)
                    if (pressMin > s.press[j]) pressMin = s.press[j];
            });
        walk(s, xi, yi, zi, [&](int j) {
            const double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (!(d2 < P.re_s2 && (d2 < dm2 || pw) && j != i)) return;
            double wS = weight(sqrt(d2), P.re_s, P.weight_type);
            if (gt == 0) wS *= (s.press[j] - pressMin) / d2;
            else if (gt == 1) wS *= (s.press[j] + Pi) / d2;
            else if (gt == 2) wS *= (s.press[j] + Pi - 2.0 * pressMin) / d2;
            else if (gt == 3) {
                const double nj = s.pndi[j];
                if (ni > kEps && nj > kEps) wS *= (ni * s.press[j] / nj + nj * Pi / ni) / d2;
            }
            accX += v0 * wS; accY += v1 * wS; accZ += v2 * wS;
        });
        s.acc[i * 3] += (1.0 - P.relax_press) * accX * P.coeff_press_grad / s.RHO[i];
        s.acc[i * 3 + 1] += (1.0 - P.relax_press) * accY * P.coeff_press_grad / s.RHO[i];
        s.acc[i * 3 + 2] += (1.0 - P.relax_press) * accZ * P.coeff_press_grad / s.RHO[i];
    }
}

// repulsive forces (src/MpsPressure.cpp:1554-1611)
inline void repulsive(const Sys& s, int i, const double* n, double ns, double vmax2, double* f) {
    const mps_params& P = s.p;
    double w;
    if (P.repulsive_force_type == MPS_REP_HARADA) {
        w = s.RHO[i] / (P.time_step * P.time_step) * (P.re_repulsive_force - ns);
    } else if (P.repulsive_force_type == MPS_REP_MITSUME) {
        w = P.rep_coef_mitsume * weight(ns, P.re_repulsive_force, P.weight_type);
    } else if (P.repulsive_force_type == MPS_REP_LENNARD_JONES) {
        double R1 = (P.re_repulsive_force / ns) * (P.re_repulsive_force / ns);
        double R2 = R1 * R1;
        w = (P.rep_coef_lennard_jones * vmax2 / ns) * (R2 - R1) * s.RHO[i];
    } else {
        double W1 = (1.0 + 3.0 * 0.5 * ns / (P.re_repulsive_force));
        double W2 = (1.0 - ns / (P.re_repulsive_force)) * (1.0 - ns / (P.re_repulsive_force)) * (1.0 - ns / (P.re_repulsive_force));
        w = (P.rep_coef_monaghan_kajtar * vmax2 / (ns - 0.0 * P.part_dist)) * (1.0 / 8.0) * (W1) * (W2) * s.RHO[i];
    }
    f[0] = -w * n[0]; f[1] = -w * n[1]; f[2] = -w * n[2];
}

// calcWallPressGradient (src/MpsPressure.cpp:1341-1550)
void wallPressGradient(Sys& s) {
    const mps_params& P = s.p;
    const double vol = pow(P.part_dist, P.dim);
    double vm = std::min(s.velMax, P.expect_max_velocity);
    double vmax2 = vm * vm;
    const int gt = P.gradient_type;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (!(s.type[i] == P.fluid && s.nearWall[i])) continue;
        double ax = 0, ay = 0, az = 0;
        const double* xi = &s.pos[i * 3];
        const double* xm = &s.mirror[i * 3];
        double pmin = s.press[i], Pi = s.press[i], ni = s.pndi[i];
        double n[3], R[9], ns;
        reflection(xi, xm, n, R, &ns);
        if (gt == 0 || gt == 2)
            walk(s, xi[0], xi[1], xi[2], [&](int j) {
                double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
                double d2 = v0 * v0 + v1 * v1 + v2 * v2;
                double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
                double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
                if (d2 < P.re_s2 && d2 < dm2 && j != i)
                    if (pmin > s.press[j]) pmin = s.press[j];
            });
        auto term = [&](double Pj, double nj, double w0, double w1, double w2, double dm2) {
            double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
            if (gt == 0) wS *= (Pj - pmin) / dm2;
            else if (gt == 1) wS *= (Pj + Pi) / dm2;
            else if (gt == 2) wS *= (Pj + Pi - 2.0 * pmin) / dm2;
            else if (gt == 3) {
                if (ni > kEps && nj > kEps) wS *= (ni * Pj / nj + nj * Pi / ni) / dm2;
            }
            ax += w0 * wS; ay += w1 * wS; az += w2 * wS;
        };
        walk(s, xi[0], xi[1], xi[2], [&](int j) {
            double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2 && j != i) term(s.press[j], s.pndi[j], w0, w1, w2, dm2);
        });
        {
            double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (dm2 < P.re_s2) term(Pi, s.pndi[i], w0, w1, w2, dm2);
        }
        double f[3] = {0.0, 0.0, 0.0};
        if (ns < P.re_repulsive_force && ns > kEps) repulsive(s, i, n, ns, vmax2, f);
        double g0 = P.relax_press * (R[0] * ax + R[1] * ay + R[2] * az) * P.coeff_press_grad - f[0];
        double g1 = P.relax_press * (R[3] * ax + R[4] * ay + R[5] * az) * P.coeff_press_grad - f[1];
        double g2 = P.relax_press * (R[6] * ax + R[7] * ay + R[8] * az) * P.coeff_press_grad - f[2];
        s.acc[i * 3] += g0 / s.RHO[i];
        s.acc[i * 3 + 1] += g1 / s.RHO[i];
        s.acc[i * 3 + 2] += g2 / s.RHO[i];
        s.forceWall[i * 3] += -g0 * vol;
        s.forceWall[i * 3 + 1] += -g1 * vol;
        s.forceWall[i * 3 + 2] += -g2 * vol;
    }
}

// predictionWallPressGradient (src/MpsPressure.cpp:1031-1214): the polygon
// wall part of the (1 - relax) gradient with the repulsive force; the wall
// force is not accumulated here
void predictionWallPressGradient(Sys& s) {
    const mps_params& P = s.p;
    double vm = std::min(s.velMax, P.expect_max_velocity);
    double vmax2 = vm * vm;
    const int gt = P.gradient_type;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        if (!(s.type[i] == P.fluid && s.nearWall[i])) continue;
        double accX = 0.0, accY = 0.0, accZ = 0.0;
        const double* xi = &s.pos[i * 3];
        const double* xm = &s.mirror[i * 3];
        const double Pi = s.press[i], ni = s.pndi[i];
        double pressMin = Pi;
        double n[3], R[9], ns;
        reflection(xi, xm, n, R, &ns);
        if (gt == 0 || gt == 2)
            walk(s, xi[0], xi[1], xi[2], [&](int j) {
                const double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
                const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
                const double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
                const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
                if (d2 < P.re_s2 && d2 < dm2 && j != i)
                    if (pressMin > s.press[j]) pressMin = s.press[j];
            });
        auto term = [&](double Pj, double nj, double w0, double w1, double w2, double dm2) {
            double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
            if (gt == 0) wS *= (Pj - pressMin) / dm2;
            else if (gt == 1) wS *= (Pj + Pi) / dm2;
            else if (gt == 2) wS *= (Pj + Pi - 2.0 * pressMin) / dm2;
            else if (gt == 3) {
                if (ni > kEps && nj > kEps) wS *= (ni * Pj / nj + nj * Pi / ni) / dm2;
            }
            accX += w0 * wS; accY += w1 * wS; accZ += w2 * wS;
        };
        walk(s, xi[0], xi[1], xi[2], [&](int j) {
            const double v0 = PJ(s, j, 0) - xi[0], v1 = PJ(s, j, 1) - xi[1], v2 = PJ(s, j, 2) - xi[2];
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2 && j != i) term(s.press[j], s.pndi[j], w0, w1, w2, dm2);
        });
        {
            const double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (dm2 < P.re_s2) term(Pi, s.pndi[i], w0, w1, w2, dm2);
        }
        double f[3] = {0.0, 0.0, 0.0};
        if (ns < P.re_repulsive_force && ns > kEps) repulsive(s, i, n, ns, vmax2, f);
        s.acc[i * 3] += ((1.0 - P.relax_press) * (R[0] * accX + R[1] * accY + R[2] * accZ) * P.coeff_press_grad - f[0]) / s.RHO[i];
        s.acc[i * 3 + 1] += ((1.0 - P.relax_press) * (R[3] * accX + R[4] * accY + R[5] * accZ) * P.coeff_press_grad - f[1]) / s.RHO[i];
        s.acc[i * 3 + 2] += ((1.0 - P.relax_press) * (R[6] * accX + R[7] * accY + R[8] * accZ) * P.coeff_press_grad - f[2]) / s.RHO[i];
    }
}

// ----------------------------------------------------------- MpsShifting
// calcShifting (src/MpsShifting.cpp:23-104).  sequential = in-place update in
// ascending i (the reference at one thread); Jacobi = all reads see old values.
void shifting1(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
    const bool seq = P.shift_semantics == MPS_SHIFT_SEQUENTIAL;
    std::vector<double> out(s.vel);
    std::vector<double>& velR = s.vel;
    auto body = [&](int i, std::vector<double>& dst) {
        if (s.type[i] != P.fluid) return;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double ux = velR[i * 3], uy = velR[i * 3 + 1], uz = velR[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        double dx = 0, dy = 0, dz = 0;
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i) {
                double dw = delWeight(sqrt(d2), P.re_s, P.weight_type);
                dx += dw * (velR[j * 3] - ux);
                dy += dw * (velR[j * 3 + 1] - uy);
                dz += dw * (velR[j * 3 + 2] - uz);
            }
        });
        if (s.bc[i] == P.inner) {
            dst[i * 3] -= P.coeff_shifting1 * dx;
            dst[i * 3 + 1] -= P.coeff_shifting1 * dy;
            dst[i * 3 + 2] -= P.coeff_shifting1 * dz;
        }
    };
    if (seq) {
        for (int i = 0; i < s.N; i++) body(i, s.vel);
    } else {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < s.N; i++) body(i, out);
        s.vel.swap(out);
    }
}

// calcConcAndConcGradient + calcWallConcAndConcGradient (src/MpsShifting.cpp:432-692)
void shifting2(Sys& s) {
    const mps_params& P = s.p;
    const bool pw = partWall(s);
    const bool seq = P.shift_semantics == MPS_SHIFT_SEQUENTIAL;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < s.N; i++) {
        double c = 0.0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i) c += weight(sqrt(d2), P.re_s, P.weight_type);
        });
        c += s.pndWall[i];
        s.conc[i] = c / P.pnd_small_zero;
    }
    std::vector<double> newPos(s.pos);
    auto body = [&](int i, std::vector<double>& dst) {
        if (s.type[i] != P.fluid) return;
        double g0 = 0, g1 = 0, g2 = 0;
        const double xi = s.pos[i * 3], yi = s.pos[i * 3 + 1], zi = s.pos[i * 3 + 2];
        const double mx = s.mirror[i * 3], my = s.mirror[i * 3 + 1], mz = s.mirror[i * 3 + 2];
        double ci = s.conc[i];
        walk(s, xi, yi, zi, [&](int j) {
            double v0 = PJ(s, j, 0) - xi, v1 = PJ(s, j, 1) - yi, v2 = PJ(s, j, 2) - zi;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - mx, w1 = PJ(s, j, 1) - my, w2 = PJ(s, j, 2) - mz;
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw) && j != i) {
                double wS = weight(sqrt(d2), P.re_s, P.weight_type);
                g0 += (ci + s.conc[j]) * v0 * wS / d2;
                g1 += (ci + s.conc[j]) * v1 * wS / d2;
                g2 += (ci + s.conc[j]) * v2 * wS / d2;
            }
        });
        g0 *= -P.coeff_press_grad; g1 *= -P.coeff_press_grad; g2 *= -P.coeff_press_grad;
        s.gradC[i * 3] = g0; s.gradC[i * 3 + 1] = g1; s.gradC[i * 3 + 2] = g2;
        if (s.bc[i] == P.inner) {
            dst[i * 3] -= P.coeff_shifting2 * g0;
            dst[i * 3 + 1] -= P.coeff_shifting2 * g1;
            dst[i * 3 + 2] -= P.coeff_shifting2 * g2;
        }
    };
    if (seq) {
        for (int i = 0; i < s.N; i++) body(i, s.pos);
    } else {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < s.N; i++) body(i, newPos);
        s.pos.swap(newPos);
    }
    if (P.wall_type != MPS_WALL_POLYGON) return;
    newPos = s.pos;
    auto wbody = [&](int i, std::vector<double>& dst) {
        if (!(s.type[i] == P.fluid && s.nearWall[i])) return;
        const double* xi = &s.pos[i * 3];
        const double* xm = &s.mirror[i * 3];
        double d0 = 0, d1 = 0, dd2 = 0, ci = s.conc[i];
        double n[3], R[9];
        reflection(xi, xm, n, R, nullptr);
        double x0 = xi[0], x1 = xi[1], x2 = xi[2];
        walk(s, x0, x1, x2, [&](int j) {
            double v0 = PJ(s, j, 0) - x0, v1 = PJ(s, j, 1) - x1, v2 = PJ(s, j, 2) - x2;
            double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            double w0 = PJ(s, j, 0) - xm[0], w1 = PJ(s, j, 1) - xm[1], w2 = PJ(s, j, 2) - xm[2];
            double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2 && j != i) {
                double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
                d0 += (ci + s.conc[j]) * w0 * wS / dm2;
                d1 += (ci + s.conc[j]) * w1 * wS / dm2;
                dd2 += (ci + s.conc[j]) * w2 * wS / dm2;
            }
        });
        double w0 = x0 - xm[0], w1 = x1 - xm[1], w2 = x2 - xm[2];
        double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (dm2 < P.re_s2) {
            double wS = weight(sqrt(dm2), P.re_s, P.weight_type);
            d0 += (ci + ci) * w0 * wS / dm2;
            d1 += (ci + ci) * w1 * wS / dm2;
            dd2 += (ci + ci) * w2 * wS / dm2;
        }
        double gx = -(R[0] * d0 + R[1] * d1 + R[2] * dd2) * P.coeff_press_grad;
        double gy = -(R[3] * d0 + R[4] * d1 + R[5] * dd2) * P.coeff_press_grad;
        double gz = -(R[6] * d0 + R[7] * d1 + R[8] * dd2) * P.coeff_press_grad;
        s.gradC[i * 3] += gx; s.gradC[i * 3 + 1] += gy; s.gradC[i * 3 + 2] += gz;
        if (s.bc[i] == P.inner) {
            dst[i * 3] -= P.coeff_shifting2 * gx;
            dst[i * 3 + 1] -= P.coeff_shifting2 * gy;
            dst[i * 3 + 2] -= P.coeff_shifting2 * gz;
        }
    };
    if (seq) {
        for (int i = 0; i < s.N; i++) wbody(i, s.pos);
    } else {
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < s.N; i++) wbody(i, newPos);
        s.pos.swap(newPos);
    }
}

// ------------------------------------------------------------ driver
void wallContributions(Sys& s) {
    closestPoint(s);
    updateNearMesh(s);
    wallNPCD(s);
}

// main() initial pipeline (src/main.cpp:105-180, initMesh :471-556)
int initSystem(Sys& s) {
    const mps_params& P = s.p;
    updateParticlesID(s);
    checkOutDomain(s);
    if (P.wall_type == MPS_WALL_POLYGON) {
        if (!s.hasMesh) { s.err = "polygon walls need a mesh"; return -1; }
        if ((P.fem_on && !s.extraOn[0]) || (P.forced_on && !s.extraOn[1])) {
            s.err = "fem_MESH / forced_MESH need their meshes";
            return -1;
        }
        initWijnNeigh(s);
        wallContributions(s);
    }
    initialPnd(s);
    if (P.wall_type == MPS_WALL_POLYGON) meanWallPnd(s);
    meanPnd(s);
    meanNeighFluidPnd(s);
    if (P.free_surf_type == MPS_FS_PND_ARC) {
        normalParticles(s);
        if (P.wall_type == MPS_WALL_POLYGON) wallNormalParticles(s);
    }
    if (updateBC(s)) return -1;
    if (calcPress(s)) return -1;
    return 0;
}

// mainLoopOfSimulation body (src/main.cpp:252-467)
int stepSystem(Sys& s) {
    const mps_params& P = s.p;
    updateParticlesID(s);
    if (P.fluid_type == 1) nonNewtonian(s, false);
    calcViscosity(s);
    addGravity(s);
    if (P.relax_press < 1.0) {
        predictionPressGradient(s);
        if (P.wall_type == MPS_WALL_POLYGON) predictionWallPressGradient(s);
    }
    prediction(s);
    checkOutDomain(s);
    if (P.collision_type == MPS_COLLISION_PC) collisionPC(s);
    else collisionDPC(s);
    if (P.wall_type == MPS_WALL_POLYGON) wallContributions(s);
    if (P.gradient_correction) correctionMatrix(s);
    pndNeighNPCD(s);
    if (P.pnd_type == MPS_PND_DIFFUSIVE) {
        pndDiffusive(s);
        if (P.wall_type == MPS_WALL_POLYGON) wallPndDiffusive(s);
        updatePnd(s);
        meanPndWallDummySurface(s);
    } else if (P.pnd_type == MPS_PND_MEAN_SUM_WIJ) {
        if (P.wall_type == MPS_WALL_POLYGON) meanWallPnd(s);
        meanPnd(s);
    }
    if (P.free_surf_type == MPS_FS_PND_ARC) {
        normalParticles(s);
        if (P.wall_type == MPS_WALL_POLYGON) wallNormalParticles(s);
    }
    if (updateBC(s)) return -1;
    if (calcPress(s)) return -1;
    if (P.wall_type == MPS_WALL_PARTICLE) extrapolateWallDummy(s);
    if (P.gradient_correction) correctionMatrix(s);
    pressGradient(s);
    if (P.wall_type == MPS_WALL_POLYGON) wallPressGradient(s);
    if (P.fluid_type == 1) nonNewtonian(s, true);
    if (P.wall_type == MPS_WALL_POLYGON) calcWallViscosity(s);
    // src/main.cpp:410-414
    if (P.forced_on && forcedMeshUpdate(s, P.time_step, s.timeCurrent)) return -1;
    correction(s);
    checkOutDomain(s);
    if (P.shifting_type == 1) shifting1(s);
    else if (P.shifting_type == 2) shifting2(s);
    if (P.wall_type == MPS_WALL_PARTICLE) wallDummyVelocity(s);
    checkOutDomain(s);
    s.numIter++;
    s.timeCurrent += P.time_step;
    return 0;
}

struct Field {
    const char* name;
    int kind;   // 0 double, 1 int, 2 char
    int comps;
};

void* fieldPtr(Sys& s, const std::string& f, int& kind, int& comps) {
    comps = 1; kind = 0;
#define D1(nm, v) if (f == nm) { kind = 0; comps = 1; return v.data(); }
#define D3(nm, v) if (f == nm) { kind = 0; comps = 3; return v.data(); }
#define I1(nm, v) if (f == nm) { kind = 1; comps = 1; return v.data(); }
    D3("pos", s.pos) D3("vel", s.vel) D3("acc", s.acc) D3("acc_star", s.accStar)
    D3("npcd", s.npcd) D3("grad_conc", s.gradC) D3("cm1", s.cm1) D3("cm2", s.cm2) D3("cm3", s.cm3)
    D3("wall_pos", s.wallPos) D3("mirror", s.mirror) D3("polygon_normal", s.polyN) D3("force_wall", s.forceWall)
    D3("normal", s.normal)
    D3("dvel", s.dvel)
    D1("press", s.press) D1("press_avg", s.pressAvg) D1("pndi", s.pndi) D1("pndki", s.pndki) D1("pndski", s.pndski)
    D1("pnd_small", s.pndSmall) D1("npcd2", s.npcd2) D1("conc", s.conc) D1("div_u", s.divU)
    D1("diffusive", s.diffTerm) D1("pnd_wall", s.pndWall) D1("riw2", s.riw2) D1("rho", s.RHO) D1("meu", s.MEU)
    D1("ppe_x", s.ppeX) D1("ppe_b", s.ppeB) D1("dev_dot_n", s.devDotN)
    I1("type", s.type) I1("bc", s.bc) I1("num_neigh", s.numNeigh) I1("num_neigh_wall", s.numNeighWall)
    I1("ptype", s.ptype) I1("near_mesh_type", s.nearMeshType)
    D1("rho", s.RHO) D1("meu", s.MEU) D1("cv", s.Cv) D1("ii", s.II) D1("meu_y", s.MEUY) D1("inertia", s.inertia)
    D1("p_rheo", s.pRheo) D1("p_smooth", s.pSmooth) D1("vf", s.VF) D1("s11", s.S11) D1("s12", s.S12)
    D1("s13", s.S13) D1("s22", s.S22) D1("s23", s.S23) D1("s33", s.S33)
    if (f == "near_wall") { kind = 2; comps = 1; return s.nearWall.data(); }
#undef D1
#undef D3
#undef I1
    return nullptr;
}

}  // namespace

extern "C" {

void* orc_create(const mps_params* p, int n, const int* type, const double* pos, const double* vel,
                 const double* press, const double* press_avg) {
    Sys* s = new Sys();
    s->p = *p;
    s->N = s->N0 = n;
    auto z = [&](std::vector<double>& v, int k, double val = 0.0) { v.assign((size_t)n * k, val); };
    s->type.assign(type, type + n);
    s->pos.assign(pos, pos + 3 * n);
    s->vel.assign(vel, vel + 3 * n);
    s->press.assign(press, press + n);
    s->pressAvg.assign(press_avg, press_avg + n);
    s->bc.assign(n, 0); s->numNeigh.assign(n, 0); s->numNeighWall.assign(n, 0); s->elementID.assign(n, 0);
    s->nearWall.assign(n, 0); s->nearMeshType.assign(n, 0);
    z(s->pndi, 1); z(s->pndki, 1); z(s->pndski, 1); z(s->pndSmall, 1); z(s->npcd2, 1); z(s->conc, 1);
    z(s->divU, 1); z(s->diffTerm, 1); z(s->pndWall, 1); z(s->devDotN, 1); z(s->nNeighSurf, 1);
    z(s->aux1, 1); z(s->aux2, 1);
    z(s->riw2, 1, 10e8 * p->part_dist);
    z(s->acc, 3); z(s->accStar, 3); z(s->npcd, 3); z(s->gradC, 3); z(s->cm1, 3); z(s->cm2, 3); z(s->cm3, 3);
    z(s->normal, 3); z(s->dvel, 3); z(s->wallPos, 3); z(s->mirror, 3); z(s->wpf1, 3); z(s->wpf2, 3);
    z(s->polyN, 3); z(s->forceWall, 3);
    s->ptype.assign(n, 1);
    s->RHO.assign(n, p->dns_fl1);
    s->MEU.assign(n, p->knm_vs1 * p->dns_fl1);
    for (auto* v : {&s->Cv, &s->II, &s->MEUY, &s->inertia, &s->pRheo, &s->pSmooth, &s->VF, &s->S11, &s->S12, &s->S13,
                    &s->S22, &s->S23, &s->S33})
        v->assign(n, 0.0);
    s->ppeX.assign(n, 0.0);
    s->ppeB.assign(n, 0.0);
    if (deriveParams(*s) != 0) {
        fprintf(stderr, "oracle: %s\n", s->err.c_str());
        delete s;
        return nullptr;
    }
    // RHO / MEU use DNS_FL1 of the derived parameter block (readMpsParticleFile :833-837);
    // with the granular phase, particles of type 1 in the file are its fluid
    // particles (PTYPE 2, density and viscosity of phase 2, :839-855)
    s->RHO.assign(n, s->p.dns_fl1);
    s->MEU.assign(n, s->p.knm_vs1 * s->p.dns_fl1);
    if (s->p.fluid_type == 1)
        for (int i = 0; i < n; i++)
            if (s->type[i] == 1) {
                s->type[i] = 0;
                s->ptype[i] = 2;
                s->RHO[i] = s->p.dns_fl2;
                s->MEU[i] = s->p.knm_vs2 * s->p.dns_fl2;
            }
    s->first.assign(s->p.num_buckets_xyz, -1);
    s->last.assign(s->p.num_buckets_xyz, -1);
    s->next.assign(n, -1);
    return s;
}

void orc_destroy(void* h) { delete (Sys*)h; }

void orc_get_params(void* h, mps_params* out) { *out = ((Sys*)h)->p; }

int orc_num_particles(void* h) { return ((Sys*)h)->N; }

const char* orc_last_error(void* h) { return ((Sys*)h)->err.c_str(); }

void orc_set_threads(void* h, int t) {
    ((Sys*)h)->nThreads = t;
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#endif
}

int orc_set_mesh(void* h, int nv, const double* V, int nf, const int* F, const double* Nrm) {
    Sys* s = (Sys*)h;
    s->mesh.nv = nv; s->mesh.nf = nf;
    s->mesh.V.assign(V, V + 3 * nv);
    s->mesh.F.assign(F, F + 3 * nf);
    s->mesh.Nrm.assign(Nrm, Nrm + 3 * nf);
    s->hasMesh = true;
    initWijnNeigh(*s);
    return 0;
}

int orc_set_mesh_type(void* h, int type, int nv, const double* V, int nf, const int* F, const double* Nrm) {
    if (type == 0) return orc_set_mesh(h, nv, V, nf, F, Nrm);
    if (type != 1 && type != 2) return -1;
    Sys* s = (Sys*)h;
    Mesh& m = s->extra[type - 1];
    m.nv = nv; m.nf = nf;
    m.V.assign(V, V + 3 * nv);
    m.F.assign(F, F + 3 * nf);
    m.Nrm.assign(Nrm, Nrm + 3 * nf);
    s->extraOn[type - 1] = true;
    return 0;
}

int orc_get_mesh(void* h, int type, int* nv, double* V, int* nf, int* F, double* Nrm) {
    Sys* s = (Sys*)h;
    const Mesh* m = type == 0 ? &s->mesh : ((type == 1 || type == 2) && s->extraOn[type - 1] ? &s->extra[type - 1] : nullptr);
    if (!m) return -1;
    if (nv) *nv = m->nv;
    if (nf) *nf = m->nf;
    if (V) std::copy(m->V.begin(), m->V.end(), V);
    if (F) std::copy(m->F.begin(), m->F.end(), F);
    if (Nrm) std::copy(m->Nrm.begin(), m->Nrm.end(), Nrm);
    return 0;
}

int orc_forced_mesh_update(void* h, double dt, double time) { return forcedMeshUpdate(*(Sys*)h, dt, time); }

int orc_get_wall_tables(void* h, double* xpnd, double* xneigh, double* zij, int* nneigh) {
    Sys* s = (Sys*)h;
    if (s->mesh.xPND.empty()) initWijnNeigh(*s);
    for (size_t k = 0; k < s->mesh.xPND.size(); k++) {
        xpnd[k] = s->mesh.xPND[k]; xneigh[k] = s->mesh.xNeigh[k];
        zij[k] = s->mesh.Zij[k]; nneigh[k] = s->mesh.nNeighWall[k];
    }
    return (int)s->mesh.xPND.size();
}

int orc_field(void* h, const char* name, void* buf, int write) {
    Sys* s = (Sys*)h;
    int kind, comps;
    void* p = fieldPtr(*s, name, kind, comps);
    if (!p) { s->err = std::string("unknown field ") + name; return -1; }
    size_t n = (size_t)s->N0 * comps;
    size_t bytes = n * (kind == 0 ? 8 : kind == 1 ? 4 : 1);
    if (write) memcpy(p, buf, bytes); else memcpy(buf, p, bytes);
    return 0;
}

int orc_scalar(void* h, const char* name, double* out) {
    Sys* s = (Sys*)h;
    std::string f(name);
    if (f == "vel_max") *out = s->velMax;
    else if (f == "cfl") *out = s->cflCurrent;
    else if (f == "time") *out = s->timeCurrent;
    else if (f == "solver_iter") *out = s->solverIter;
    else if (f == "solver_err") *out = s->solverErr;
    else if (f == "num_particles") *out = s->N;
    else if (f == "cfl_visc") *out = s->cflVisc;
    else if (f == "solve_setup_ms") *out = s->solveSetupSeconds * 1e3;
    else if (f == "solve_iter_ms") *out = s->solveIterSeconds * 1e3;
    else return -1;
    return 0;
}

int orc_set_scalar(void* h, const char* name, double v) {
    Sys* s = (Sys*)h;
    std::string f(name);
    if (f == "vel_max") s->velMax = v;
    else if (f == "solver_tol") s->p.solver_tol = v;        // 0: machine epsilon, as Eigen's default
    else if (f == "parallel_cg") s->parallelCg = v != 0.0;
    else if (f == "threads") orc_set_threads(h, (int)v);
    else return -1;
    return 0;
}

// PPE rows in triplet order: returns nnz; arrays may be null to query size.
int orc_ppe_matrix(void* h, int* row_ptr, int* cols, double* vals) {
    Sys* s = (Sys*)h;
    int nnz = (int)s->colIdx.size();
    if (row_ptr) for (int k = 0; k <= s->N; k++) row_ptr[k] = s->rowPtr[k];
    if (cols) for (int k = 0; k < nnz; k++) cols[k] = s->colIdx[k];
    if (vals) for (int k = 0; k < nnz; k++) vals[k] = s->vals[k];
    return nnz;
}

// Bucket list as (first, next) arrays.
int orc_buckets(void* h, int* first, int* next) {
    Sys* s = (Sys*)h;
    if (first) memcpy(first, s->first.data(), sizeof(int) * s->first.size());
    if (next) memcpy(next, s->next.data(), sizeof(int) * s->next.size());
    return (int)s->first.size();
}

// Neighbour list of particle i in walk order for radius^2 r2 with the
// mirror test (mode 0: d2<r2 && (d2<dm2 || particle walls); mode 1: none).
int orc_neighbors(void* h, int i, double r2, int mode, int* out, int cap) {
    Sys* s = (Sys*)h;
    const bool pw = partWall(*s);
    int cnt = 0;
    const double xi = s->pos[i * 3], yi = s->pos[i * 3 + 1], zi = s->pos[i * 3 + 2];
    const double mx = s->mirror[i * 3], my = s->mirror[i * 3 + 1], mz = s->mirror[i * 3 + 2];
    walk(*s, xi, yi, zi, [&](int j) {
        if (j == i) return;
        double v0 = PJ(*s, j, 0) - xi, v1 = PJ(*s, j, 1) - yi, v2 = PJ(*s, j, 2) - zi;
        double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        double w0 = PJ(*s, j, 0) - mx, w1 = PJ(*s, j, 1) - my, w2 = PJ(*s, j, 2) - mz;
        double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        bool ok = d2 < r2 && (mode == 1 || d2 < dm2 || pw);
        if (ok) { if (cnt < cap) out[cnt] = j; cnt++; }
    });
    return cnt;
}

int orc_call(void* h, const char* op) {
    Sys* s = (Sys*)h;
    std::string o(op);
    if (o == "update_buckets") updateParticlesID(*s);
    else if (o == "viscosity") calcViscosity(*s);
    else if (o == "volume_fraction") volumeFraction(*s);
    else if (o == "viscosity_interaction") viscosityInteractionVal(*s);
    else if (o == "wall_viscosity_interaction") wallViscosityInteractionVal(*s);
    else if (o == "add_gravity") addGravity(*s);
    else if (o == "prediction") prediction(*s);
    else if (o == "check_out_domain") checkOutDomain(*s);
    else if (o == "collision") { if (s->p.collision_type == MPS_COLLISION_PC) collisionPC(*s); else collisionDPC(*s); }
    else if (o == "closest_point") closestPoint(*s);
    else if (o == "update_near_mesh") updateNearMesh(*s);
    else if (o == "wall_npcd") wallNPCD(*s);
    else if (o == "wall_contributions") wallContributions(*s);
    else if (o == "correction_matrix") correctionMatrix(*s);
    else if (o == "initial_pnd") initialPnd(*s);
    else if (o == "pnd_neigh_npcd") pndNeighNPCD(*s);
    else if (o == "pnd_diffusive") pndDiffusive(*s);
    else if (o == "wall_pnd_diffusive") wallPndDiffusive(*s);
    else if (o == "update_pnd") updatePnd(*s);
    else if (o == "mean_pnd_wall_dummy_surface") meanPndWallDummySurface(*s);
    else if (o == "mean_pnd") meanPnd(*s);
    else if (o == "mean_wall_pnd") meanWallPnd(*s);
    else if (o == "mean_neigh_fluid_pnd") meanNeighFluidPnd(*s);
    else if (o == "update_bc") return updateBC(*s);
    else if (o == "vel_divergence") velDivergence(*s);
    else if (o == "wall_vel_divergence") wallVelDivergence(*s);
    else if (o == "prediction_press_gradient") predictionPressGradient(*s);
    else if (o == "normal_particles") normalParticles(*s);
    else if (o == "wall_normal_particles") wallNormalParticles(*s);
    else if (o == "prediction_wall_press_gradient") predictionWallPressGradient(*s);
    else if (o == "assemble_ppe") assemblePPE(*s);
    else if (o == "solve_ppe") {
        if (s->p.solver_type == MPS_SOLVER_BICGSTAB) solveBiCGSTAB(*s); else solveCG(*s);
        for (int i = 0; i < s->N; i++) s->press[i] = s->ppeX[i] < 0 ? 0.0 : s->ppeX[i]; }
    else if (o == "calc_press") return calcPress(*s);
    else if (o == "extrapolate_wall_dummy") extrapolateWallDummy(*s);
    else if (o == "press_gradient") pressGradient(*s);
    else if (o == "wall_press_gradient") wallPressGradient(*s);
    else if (o == "wall_viscosity") calcWallViscosity(*s);
    else if (o == "correction") correction(*s);
    else if (o == "shifting") { if (s->p.shifting_type == 1) shifting1(*s); else if (s->p.shifting_type == 2) shifting2(*s); }
    else if (o == "wall_dummy_velocity") wallDummyVelocity(*s);
    else if (o == "init") return initSystem(*s);
    else if (o == "step") return stepSystem(*s);
    else { s->err = "unknown op " + o; return -1; }
    return 0;
}

int orc_steps(void* h, int k) {
    Sys* s = (Sys*)h;
    for (int t = 0; t < k; t++)
        if (stepSystem(*s)) return -1;
    return 0;
}

}  // extern "C"
