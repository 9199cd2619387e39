// Granular (non-Newtonian) phase, fluid_type 1: the viscosity with
// per-particle MEU and RHO (MpsViscosity::calcViscosity,
// src/MpsViscosity.cpp:20-102), the volume fraction (calcVolumeFraction
// :104-191), the strain rate and mu(I) rheology (calcViscosityInteractionVal
// :194-700) and their polygon-wall forms (calcWall{Slip,NoSlip}
// ViscosityInteractionVal :703-1526).  Every FP64 expression keeps the
// reference's operation order; sin / cos / tan / exp / pow come from the
// device library and may differ from glibc in the last place.
#include <climits>
#include <cstring>
#include <string>

#include "kern_common.h"

#define PX(k) a.D[DX][k]
#define PY(k) a.D[DY][k]
#define PZ(k) a.D[DZ][k]

namespace mpsgpu {

namespace {

constexpr int kGB = 256;

// calcViscosity (src/MpsViscosity.cpp:20-102) with MEU and RHO per particle
template <bool LIST>
__global__ void kViscosityNN(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double meuI = a.D[DMEU][i], rhoI = a.D[DRHO][i];
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double ax = 0, ay = 0, az = 0;
    forNbrsPosType<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj, int tj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_l2 && (d2 < dm2 || pw)) {
            const double wL = wgt(sqrt(d2), P.re_l, P.weight_type);
            const double meuJ = a.D[DMEU][j];
            double nu;
            if ((meuI + meuJ) > KEPS) nu = 2.0 * meuI * meuJ / (meuI + meuJ);
            else nu = 0.0;
            if (tj == P.wall) nu = 2.0 * meuI;
            nu = nu / rhoI;
            ax += (a.D[DU][j] - ux) * wL * nu;
            ay += (a.D[DV][j] - uy) * wL * nu;
            az += (a.D[DW][j] - uz) * wL * nu;
        }
    });
    a.D[DAX][i] = P.coeff_visc_multiphase * ax;
    a.D[DAY][i] = P.coeff_visc_multiphase * ay;
    a.D[DAZ][i] = P.coeff_visc_multiphase * az;
}

// calcVolumeFraction (src/MpsViscosity.cpp:104-191)
template <bool LIST>
__global__ void kVolumeFraction(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (a.I[ITYPE][i] != P.fluid) return;
    const bool weighted = P.fraction_method == 2;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    double sum1 = 0.0, sum2 = 0.0;
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        if (d2 < P.re_s2 && a.I[ITYPE][j] == P.fluid) {
            if (weighted) {
                const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
                sum1 += wS;
                if (a.I[IPTYPE][j] >= 2) sum2 += wS;
            } else {
                sum1 += 1.0;
                if (a.I[IPTYPE][j] >= 2) sum2 += 1.0;
            }
        }
    });
    a.D[DCV][i] = (sum1 < KEPS) ? 0.0 : sum2 / sum1;
}

// ------------------------------------------------------------ bed levels
// Grid of 2 l0 columns (src/MpsViscosity.cpp:236-305) holding the bed level
// BL (highest fluid particle with Cv > 0.5), the water level WL (highest
// phase-1 particle) and PS, the pressure of the bed-level particle.  The
// reference fills it from an unsynchronised parallel loop; the result
// defined here is that of its loop in index order: BL and WL are maxima
// over the particles above the initial level, PS belongs to the lowest
// reference index at the bed level.  Levels are kept as order-preserving
// 64-bit keys (0 = never raised) so that atomicMax gives the maximum.
struct BedArgs {
    unsigned long long* keys;   // [0, cells): BL, [cells, 2 cells): WL, [2 cells]: MEU maximum
    int* idx;                   // [cells]: lowest reference index at the bed level
    int kxMax, kyMax, cells;
    double xmin, ymin, aa, level0;
    bool wallForm;
};

__device__ __forceinline__ unsigned long long orderedKey(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double keyValue(unsigned long long k) {
    const unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

// column of a particle (1-based kx, ky as the reference's arrays), -1 outside
// the grid (an out-of-bounds access in the reference)
__device__ __forceinline__ int bedCell(const BedArgs& g, bool d3, double x, double y) {
    const int kx = int((x - g.xmin) / g.aa) + 1;
    const int ky = d3 ? int((y - g.ymin) / g.aa) + 1 : 1;
    if (kx < 1 || kx > g.kxMax || ky < 1 || ky > g.kyMax) return -1;
    return (kx - 1) * g.kyMax + (ky - 1);
}

__global__ void kBedInit(BedArgs g) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < g.cells) {
        g.keys[c] = 0ull;
        g.keys[g.cells + c] = 0ull;
        g.idx[c] = INT_MAX;
    }
    if (c == 0 && !g.wallForm) g.keys[2 * g.cells] = 0ull;   // the wall form leaves CFLvisc of the fluid pass
}

__global__ void kBedLevels(KArgs a, BedArgs g) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (a.I[ITYPE][i] != P.fluid) return;
    const bool d3 = (int)P.dim == 3;
    const int c = bedCell(g, d3, a.D[DX][i], a.D[DY][i]);
    if (c < 0) return;
    const double h = d3 ? a.D[DZ][i] : a.D[DY][i];
    if (!(h > g.level0)) return;
    if (a.D[DCV][i] > 0.5) atomicMax(&g.keys[c], orderedKey(h + 0.0));   // + 0.0: -0 and +0 share a key
    if (a.I[IPTYPE][i] == 1) atomicMax(&g.keys[g.cells + c], orderedKey(h + 0.0));
}

__global__ void kBedIndex(KArgs a, BedArgs g) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (a.I[ITYPE][i] != P.fluid || !(a.D[DCV][i] > 0.5)) return;
    const bool d3 = (int)P.dim == 3;
    const int c = bedCell(g, d3, a.D[DX][i], a.D[DY][i]);
    if (c < 0) return;
    const double h = d3 ? a.D[DZ][i] : a.D[DY][i];
    const unsigned long long k = g.keys[c];
    if (k != 0ull && h > g.level0 && h == keyValue(k)) atomicMin(&g.idx[c], a.I[ICANON][i]);
}

struct Levels {
    double BL, WL, PS;
};
__device__ __forceinline__ Levels bedLevels(const KArgs& a, const BedArgs& g, const int* slotOf, int c) {
    Levels L{g.level0, g.level0, 0.0};
    if (c < 0) return L;
    const unsigned long long kb = g.keys[c], kw = g.keys[g.cells + c];
    if (kb != 0ull) L.BL = keyValue(kb);
    if (kw != 0ull) L.WL = keyValue(kw);
    const int q = g.idx[c];
    if (q != INT_MAX) L.PS = a.D[DPRESS][slotOf[q]];
    return L;
}

// ----------------------------------------------------------- strain rate
// src/MpsViscosity.cpp:307-440: strain-rate tensor, its second invariant
// and the smoothed pressure of the fluid particles
template <bool LIST>
__global__ void kStrainRate(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (a.I[ITYPE][i] != P.fluid) return;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double MC[9];
    for (int q = 0; q < 9; q++) MC[q] = a.D[DCM0 + q][i];
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0, s6 = 0.0, s7 = 0.0, s8 = 0.0, s9 = 0.0, s10 = 0.0,
           sw = 0.0;
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
            const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
            const double ex = a.D[DU][j] - ux, ey = a.D[DV][j] - uy, ez = a.D[DW][j] - uz;
            const double inv = 1.0 / d2;
            if (!P.gradient_correction) {
                s1 += ex * v0 * wS * inv; s2 += ex * v1 * wS * inv; s3 += ex * v2 * wS * inv;
                s4 += ey * v0 * wS * inv; s5 += ey * v1 * wS * inv; s6 += ey * v2 * wS * inv;
                s7 += ez * v0 * wS * inv; s8 += ez * v1 * wS * inv; s9 += ez * v2 * wS * inv;
            } else {
                const double c0 = (v0 * MC[0] + v1 * MC[1] + v2 * MC[2]);
                const double c1 = (v0 * MC[3] + v1 * MC[4] + v2 * MC[5]);
                const double c2 = (v0 * MC[6] + v1 * MC[7] + v2 * MC[8]);
                s1 += ex * c0 * wS * inv; s2 += ex * c1 * wS * inv; s3 += ex * c2 * wS * inv;
                s4 += ey * c0 * wS * inv; s5 += ey * c1 * wS * inv; s6 += ey * c2 * wS * inv;
                s7 += ez * c0 * wS * inv; s8 += ez * c1 * wS * inv; s9 += ez * c2 * wS * inv;
            }
            s10 += a.D[DPRESS][j] * wS;
            sw += wS;
        }
    });
    const double c = P.coeff_press_grad;
    const double Uxx = -c * s1, Uxy = -c * s2, Uxz = -c * s3;
    const double Uyx = -c * s4, Uyy = -c * s5, Uyz = -c * s6;
    const double Uzx = -c * s7, Uzy = -c * s8, Uzz = -c * s9;
    double ps = (sw > KEPS) ? s10 / sw : a.D[DPRESS][i];
    if (ps < KEPS) ps = 0.0;
    a.D[DPSMOOTH][i] = ps;
    const double S11 = 0.5 * (Uxx + Uxx), S12 = 0.5 * (Uxy + Uyx), S13 = 0.5 * (Uxz + Uzx);
    const double S22 = 0.5 * (Uyy + Uyy), S23 = 0.5 * (Uyz + Uzy), S33 = 0.5 * (Uzz + Uzz);
    a.D[DS11][i] = S11; a.D[DS12][i] = S12; a.D[DS13][i] = S13;
    a.D[DS22][i] = S22; a.D[DS23][i] = S23; a.D[DS33][i] = S33;
    double ii = 0.5 * (S11 * S11 + 2.0 * S12 * S12 + 2.0 * S13 * S13 + S22 * S22 + 2.0 * S23 * S23 + S33 * S33);
    if (ii < KEPS || ii * 0 != 0) ii = 0.0;
    a.D[DII][i] = ii;
}

// the MEU maximum of a block, one atomic per block (CFLvisc, :691)
__device__ __forceinline__ void blockMaxKey(unsigned long long* dst, double v, bool has) {
    __shared__ unsigned long long sh[kGB / 64];
    unsigned long long k = has ? orderedKey(v + 0.0) : 0ull;
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_down(k, o, 64);
        k = t > k ? t : k;
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = k;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long m = 0ull;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) m = sh[q] > m ? sh[q] : m;
        if (m != 0ull) atomicMax(dst, m);
    }
}

// mu(I) rheology of the fluid particles (src/MpsViscosity.cpp:461-691)
__global__ __launch_bounds__(kGB) void kRheology(KArgs a, BedArgs g, const int* slotOf, double timeCurrent) {
    const int i = slotId();
    const mps_params& P = a.P;
    bool hasMax = false;
    double meuMax = 0.0;
    if (live(a, i) && a.I[ITYPE][i] == P.fluid) {
        const bool d3 = (int)P.dim == 3;
        const double gmod =
            sqrt(P.gravity[0] * P.gravity[0] + P.gravity[1] * P.gravity[1] + P.gravity[2] * P.gravity[2]);
        const double f21 = P.dns_fl2 - P.dns_fl1;
        const double xi = a.D[DX][i], yi = a.D[DY][i], zi = a.D[DZ][i];
        const double vx = a.D[DU][i], vy = a.D[DV][i], vz = a.D[DW][i];
        const double vel2i = vx * vx + vy * vy + vz * vz;
        const int pt = a.I[IPTYPE][i];
        const double cv = a.D[DCV][i];
        if (pt == 1) {
            a.D[DMEU][i] = P.knm_vs1 * P.dns_fl1;
        } else if (pt == 2) {
            double phi = (cv - 0.25) * P.phi_1 / (1.0 - 0.25);
            double phi2 = (cv - 0.25) * P.phi_2 / (1.0 - 0.25);
            if (cv <= 0.25) { phi = KEPS; phi2 = KEPS; }
            double ps = a.D[DPSMOOTH][i];
            const Levels L = bedLevels(a, g, slotOf, bedCell(g, d3, xi, yi));
            double ns;
            if (!d3) {
                // the reference forms (p_smooth - PS) f21 VF / RHO first and overwrites it
                ns = (L.WL < L.BL) ? ps : ps - L.PS;
            } else {
                ns = (L.BL - zi + P.part_dist * 0.5) * f21 * gmod - vel2i * f21 * 0.5;
                if (ps - (L.WL - zi) * P.dns_fl1 * gmod < KEPS) ps = (L.WL - zi) * P.dns_fl1 * gmod;
                if (timeCurrent <= 1.0)
                    ns = (1.0 - timeCurrent) * (ps - (L.WL - zi) * P.dns_fl1 * gmod) + timeCurrent * ns;
                a.D[DPSMOOTH][i] = ps;
            }
            if (ns < 1.0 || cv < 0.5) ns = 1.0;
            a.D[DPRHEO][i] = ns;
            const double ii = a.D[DII][i];
            const double modE = sqrt(ii);
            double inertia;
            if (L.WL < L.BL) inertia = modE * P.dg / sqrt(ns / P.dns_sdt);
            else inertia = modE * P.dg / sqrt(ns / (P.dns_fl1 * P.cd));
            a.D[DINERTIA][i] = inertia;
            double vf = P.vf_max - (P.vf_max - P.vf_min) * inertia;
            if (vf < P.vf_min) vf = P.vf_min;
            a.D[DVF][i] = vf;
            a.D[DRHO][i] = P.dns_sdt * vf + (1.0 - vf) * P.dns_fl1;
            phi *= (vf / P.vf_max);
            const double alpha1 = 2.0 * sqrt(3.0) * sin(phi) / (3.0 - sin(phi));
            const double beta1 = 2.0 * sqrt(3.0) * cos(phi) / (3.0 - sin(phi));
            double yield = P.cohes * beta1 + ns * alpha1;
            if (yield < KEPS) yield = 0.0;
            double viscMax = (P.meu0 + yield * P.mm * 0.5);
            const double meuY = yield / (2.0 * sqrt(modE * modE + KEPS));
            a.D[DMEUY][i] = meuY;
            double meu0;
            if (L.WL < L.BL)
                meu0 = 0.5 * (tan(phi2) - tan(phi)) * ns * P.dg / (P.i0 * sqrt(ns / P.dns_fl2) + modE * P.dg) * modE /
                       (sqrt(modE * modE + KEPS));
            else
                meu0 = 0.5 * (tan(phi2) - tan(phi)) * ns * P.dg / (P.i0 * sqrt(ns / (P.dns_fl1 * P.cd)) + modE * P.dg) *
                       modE / (sqrt(modE * modE + KEPS));
            if (modE <= KEPS || (meu0 * 0) != 0) meu0 = P.meu0;
            viscMax = (meu0 + yield * P.mm * 0.5);
            double meu = meuY + meu0;
            if (ii <= KEPS) meu = viscMax;
            a.D[DMEU][i] = meu;
            hasMax = true;
            meuMax = meu;
        }
        if (pt >= 2) a.D[DRHO][i] = (cv > 0.5) ? P.dns_fl2 : cv * P.dns_fl2 + (1.0 - cv) * P.dns_fl1;
    }
    blockMaxKey(&g.keys[2 * g.cells], meuMax, hasMax);
}

// ------------------------------------------------------ polygon walls
// calcWall{Slip,NoSlip}ViscosityInteractionVal (src/MpsViscosity.cpp:785-923,
// 1194-1336): the mirror-neighbourhood strain rate of the fluid particles
// near the wall replaces S; II and the smoothed pressure add to the values
// of the fluid pass
template <bool LIST>
__global__ void kWallStrainRate(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (!(a.I[ITYPE][i] == P.fluid && a.I[INEARWALL][i])) return;
    const bool noSlip = P.slip_condition == MPS_NO_SLIP;
    const double xi[3] = {PX(i), PY(i), PZ(i)};
    const double xm[3] = {a.D[DMX][i], a.D[DMY][i], a.D[DMZ][i]};
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    double n[3], R[9];
    reflectMod2(xi, xm, n, R);
    double vmx, vmy, vmz;
    if (!noSlip) {
        vmx = (R[0] * ux + R[1] * uy + R[2] * uz);
        vmy = (R[3] * ux + R[4] * uy + R[5] * uz);
        vmz = (R[6] * ux + R[7] * uy + R[8] * uz);
    } else {
        const bool forced = a.I[INEARMESH][i] == 2;
        const double vw0 = forced ? P.uniform_vel_wall[0] : 0.0;
        const double vw1 = forced ? P.uniform_vel_wall[1] : 0.0;
        const double vw2 = forced ? P.uniform_vel_wall[2] : 0.0;
        const double dotnv = n[0] * vw0 + n[1] * vw1 + n[2] * vw2;
        const double t0 = ux - 2.0 * (vw0 - dotnv * n[0]);
        const double t1 = uy - 2.0 * (vw1 - dotnv * n[1]);
        const double t2 = uz - 2.0 * (vw2 - dotnv * n[2]);
        vmx = (-1.0 * t0 + 0.0 * t1 + 0.0 * t2);
        vmy = (0.0 * t0 + -1.0 * t1 + 0.0 * t2);
        vmz = (0.0 * t0 + 0.0 * t1 + -1.0 * t2);
    }
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0, s6 = 0.0, s7 = 0.0, s8 = 0.0, s9 = 0.0, s10 = 0.0;
    forNbrsPosS<LIST>(a, i, xi[0], xi[1], xi[2], [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi[0], v1 = yj - xi[1], v2 = zj - xi[2];
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - xm[0], w1 = yj - xm[1], w2 = zj - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2) {
            const double wS = wgt(sqrt(dm2), P.re_s, P.weight_type);
            const double ex = a.D[DU][j] - vmx, ey = a.D[DV][j] - vmy, ez = a.D[DW][j] - vmz;
            s1 += ex * w0 * wS / dm2; s2 += ex * w1 * wS / dm2; s3 += ex * w2 * wS / dm2;
            s4 += ey * w0 * wS / dm2; s5 += ey * w1 * wS / dm2; s6 += ey * w2 * wS / dm2;
            s7 += ez * w0 * wS / dm2; s8 += ez * w1 * wS / dm2; s9 += ez * w2 * wS / dm2;
            s10 += a.D[DPRESS][j] * wS;
        }
    });
    const double c = P.coeff_press_grad;
    double Uxx, Uxy, Uxz, Uyx, Uyy, Uyz, Uzx, Uzy, Uzz;
    if (!noSlip) {   // R grad(u) R
        const double aUxx = -c * (R[0] * s1 + R[1] * s4 + R[2] * s7);
        const double aUxy = -c * (R[0] * s2 + R[1] * s5 + R[2] * s8);
        const double aUxz = -c * (R[0] * s3 + R[1] * s6 + R[2] * s9);
        const double aUyx = -c * (R[3] * s1 + R[4] * s4 + R[5] * s7);
        const double aUyy = -c * (R[3] * s2 + R[4] * s5 + R[5] * s8);
        const double aUyz = -c * (R[3] * s3 + R[4] * s6 + R[5] * s9);
        const double aUzx = -c * (R[6] * s1 + R[7] * s4 + R[8] * s7);
        const double aUzy = -c * (R[6] * s2 + R[7] * s5 + R[8] * s8);
        const double aUzz = -c * (R[6] * s3 + R[7] * s6 + R[8] * s9);
        Uxx = aUxx * R[0] + aUxy * R[3] + aUxz * R[6];
        Uxy = aUxx * R[1] + aUxy * R[4] + aUxz * R[7];
        Uxz = aUxx * R[2] + aUxy * R[5] + aUxz * R[8];
        Uyx = aUyx * R[0] + aUyy * R[3] + aUyz * R[6];
        Uyy = aUyx * R[1] + aUyy * R[4] + aUyz * R[7];
        Uyz = aUyx * R[2] + aUyy * R[5] + aUyz * R[8];
        Uzx = aUzx * R[0] + aUzy * R[3] + aUzz * R[6];
        Uzy = aUzx * R[1] + aUzy * R[4] + aUzz * R[7];
        Uzz = aUzx * R[2] + aUzy * R[5] + aUzz * R[8];
    } else {         // -grad(u) R
        Uxx = c * (s1 * R[0] + s2 * R[3] + s3 * R[6]);
        Uxy = c * (s1 * R[1] + s2 * R[4] + s3 * R[7]);
        Uxz = c * (s1 * R[2] + s2 * R[5] + s3 * R[8]);
        Uyx = c * (s4 * R[0] + s5 * R[3] + s6 * R[6]);
        Uyy = c * (s4 * R[1] + s5 * R[4] + s6 * R[7]);
        Uyz = c * (s4 * R[2] + s5 * R[5] + s6 * R[8]);
        Uzx = c * (s7 * R[0] + s8 * R[3] + s9 * R[6]);
        Uzy = c * (s7 * R[1] + s8 * R[4] + s9 * R[7]);
        Uzz = c * (s7 * R[2] + s8 * R[5] + s9 * R[8]);
    }
    double ps = a.D[DPSMOOTH][i] + s10 / P.pnd_gradient_zero;
    if (ps < KEPS) ps = 0.0;
    a.D[DPSMOOTH][i] = ps;
    const double S11 = 0.5 * (Uxx + Uxx), S12 = 0.5 * (Uxy + Uyx), S13 = 0.5 * (Uxz + Uzx);
    const double S22 = 0.5 * (Uyy + Uyy), S23 = 0.5 * (Uyz + Uzy), S33 = 0.5 * (Uzz + Uzz);
    a.D[DS11][i] = S11; a.D[DS12][i] = S12; a.D[DS13][i] = S13;
    a.D[DS22][i] = S22; a.D[DS23][i] = S23; a.D[DS33][i] = S33;
    double ii = a.D[DII][i] +
                0.5 * (S11 * S11 + 2.0 * S12 * S12 + 2.0 * S13 * S13 + S22 * S22 + 2.0 * S23 * S23 + S33 * S33);
    if (ii < KEPS || ii * 0 != 0) ii = 0.0;
    a.D[DII][i] = ii;
}

// the near-wall rheology (src/MpsViscosity.cpp:952-1062, 1362-1476):
// Papanastasiou-regularised yield viscosity plus the power law
__global__ void kWallRheology(KArgs a, BedArgs g, double timeCurrent) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (!(a.I[ITYPE][i] == P.fluid && a.I[INEARWALL][i])) return;
    const bool noSlip = P.slip_condition == MPS_NO_SLIP;
    const double gmod = sqrt(P.gravity[0] * P.gravity[0] + P.gravity[1] * P.gravity[1] + P.gravity[2] * P.gravity[2]);
    const double f21 = P.dns_fl2 - P.dns_fl1;
    const double xi = a.D[DX][i], yi = a.D[DY][i], zi = a.D[DZ][i];
    const double vx = a.D[DU][i], vy = a.D[DV][i], vz = a.D[DW][i];
    const double vel2i = vx * vx + vy * vy + vz * vz;
    const int pt = a.I[IPTYPE][i];
    const double cv = a.D[DCV][i];
    if (pt == 1) {
        a.D[DMEU][i] = P.knm_vs1 * P.dns_fl1;
    } else if (pt == 2) {
        double phi = (cv - 0.25) * P.phi_1 / (1.0 - 0.25);
        double phi2 = (cv - 0.25) * P.phi_2 / (1.0 - 0.25);
        if (cv <= 0.25) { phi = KEPS; phi2 = KEPS; }
        const int c = bedCell(g, true, xi, yi);
        double BL = g.level0, WL = g.level0;
        if (c >= 0) {
            if (g.keys[c] != 0ull) BL = keyValue(g.keys[c]);
            if (g.keys[g.cells + c] != 0ull) WL = keyValue(g.keys[g.cells + c]);
        }
        double ps = a.D[DPSMOOTH][i];
        double ns = (BL - zi + P.part_dist * 0.5) * f21 * gmod - vel2i * f21 * 0.5;
        if (ps < (WL - zi) * P.dns_fl1 * gmod) ps = (WL - zi) * P.dns_fl1 * gmod;
        a.D[DPSMOOTH][i] = ps;
        if (timeCurrent <= 1.0) ns = (1.0 - timeCurrent) * (ps - (WL - zi) * P.dns_fl1 * gmod) + timeCurrent * ns;
        if (ns < 1.0 || cv < 0.5) ns = 1.0;
        a.D[DPRHEO][i] = ns;
        const double ii = a.D[DII][i];
        const double inertia = sqrt(ii) * P.dg / sqrt(ns / (P.dns_fl1 * P.cd));
        a.D[DINERTIA][i] = inertia;
        double vf = P.vf_max - (P.vf_max - P.vf_min) * inertia;
        if (vf < P.vf_min) vf = P.vf_min;
        a.D[DVF][i] = vf;
        a.D[DRHO][i] = P.dns_sdt * vf + (1.0 - vf) * P.dns_fl1;
        phi = phi * vf / P.vf_max;
        double yield = P.cohes * cos(phi) + ns * sin(phi);
        if (noSlip ? yield < 0.0 : yield < KEPS) yield = 0.0;
        double viscMax = (yield * P.mm * 0.5 + P.meu0);
        double meuY;
        if (ii > KEPS) meuY = yield * (1.0 - exp(-P.mm * sqrt(ii))) * 0.5 / sqrt(ii);
        else meuY = viscMax;
        a.D[DMEUY][i] = meuY;
        double meu0 = 0.5 * (tan(phi2) - tan(phi)) * ns * P.dg / (P.i0 * sqrt(ns / (P.dns_fl1 * P.cd)) + sqrt(ii) * P.dg);
        if (ii <= KEPS || (meu0 * 0) != 0) meu0 = P.meu0;
        viscMax = (yield * P.mm * 0.5 + meu0);
        double meu = meuY + P.meu0 * pow(4.0 * ii, (P.power_n - 1.0) * 0.5);
        if (ii == 0 || meu > viscMax) meu = viscMax;
        a.D[DMEU][i] = meu;
    }
    if (pt >= 2) a.D[DRHO][i] = (cv > 0.5) ? P.dns_fl2 : cv * P.dns_fl2 + (1.0 - cv) * P.dns_fl1;
}

// phase of every particle from the file's types (readMpsParticleFile,
// src/MpsInputOutput.cpp:832-855): with the granular phase, type 1 marks its
// particles (fluid, PTYPE 2, phase-2 density and viscosity)
__global__ void kInitPhase(KArgs a) {
    const int i = slotId();
    if (i >= a.N) return;
    const mps_params& P = a.P;
    int pt = 1;
    double rho = P.dns_fl1, meu = P.knm_vs1 * P.dns_fl1;
    if (P.fluid_type == 1 && a.I[ITYPE][i] == 1) {
        a.I[ITYPE][i] = 0;
        pt = 2;
        rho = P.dns_fl2;
        meu = P.knm_vs2 * P.dns_fl2;
    }
    a.I[IPTYPE][i] = pt;
    a.D[DRHO][i] = rho;
    a.D[DMEU][i] = meu;
    for (int f = DCV; f <= DS33; f++) a.D[f][i] = 0.0;
}

BedArgs bedArgs(Ctx* ctx, bool wallForm) {
    const mps_params& P = ctx->P;
    const bool d3 = (int)P.dim == 3;
    BedArgs g;
    g.keys = ctx->bedKeys;
    g.idx = ctx->bedIdx;
    g.kxMax = ctx->bedKx;
    g.kyMax = ctx->bedKy;
    g.cells = ctx->bedKx * ctx->bedKy;
    g.xmin = P.domain_min[0];
    g.ymin = P.domain_min[1];
    g.aa = 2.0 * P.part_dist;
    g.level0 = (d3 || wallForm) ? (d3 ? P.domain_min[2] : 0.0) : P.domain_min[1];
    g.wallForm = wallForm;
    return g;
}

}  // namespace

// Grid size of the bed levels (src/MpsViscosity.cpp:230-235)
void granularGridSize(const mps_params& P, int* kx, int* ky) {
    const bool d3 = (int)P.dim == 3;
    const double aa = 2.0 * P.part_dist;
    const double Ymin = P.domain_min[1], Ymax = d3 ? P.domain_max[1] : Ymin;
    *kx = int((P.domain_max[0] - P.domain_min[0]) / aa) + 1;
    *ky = int((Ymax - Ymin) / aa) + 1;
}

int launchInitPhase(Ctx* ctx) {
    KArgs a = ctx->args();
    hipLaunchKernelGGL(kInitPhase, dim3(gridFor(ctx->N, kGB)), dim3(kGB), 0, ctx->stream, a);
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

int launchViscosityNN(Ctx* ctx) {
    KArgs a = ctx->args();
    const int g = gridFor(ctx->N, kGB);
    if (ctx->listValid) hipLaunchKernelGGL(kViscosityNN<true>, dim3(g), dim3(kGB), 0, ctx->stream, a);
    else hipLaunchKernelGGL(kViscosityNN<false>, dim3(g), dim3(kGB), 0, ctx->stream, a);
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

// op: "volume_fraction", "viscosity_interaction" (bed levels, strain rate,
// rheology) or "wall_viscosity_interaction" (polygon walls, 3-D)
int launchGranular(Ctx* ctx, const char* op) {
    const std::string o(op);
    const mps_params& P = ctx->P;
    if (P.fluid_type != 1) {
        ctx->err = o + " needs the granular phase (fluid_type 1)";
        return MPS_E_ARG;
    }
    KArgs a = ctx->args();
    const int g = gridFor(ctx->N, kGB);
    const bool L = ctx->listValid;
#define GL(K)                                                                                      \
    do {                                                                                           \
        if (L) hipLaunchKernelGGL(K<true>, dim3(g), dim3(kGB), 0, ctx->stream, a);                 \
        else hipLaunchKernelGGL(K<false>, dim3(g), dim3(kGB), 0, ctx->stream, a);                  \
    } while (0)
    if (o == "volume_fraction") {
        if (P.fraction_method == 1 || P.fraction_method == 2) GL(kVolumeFraction);
    } else if (o == "viscosity_interaction" || o == "wall_viscosity_interaction") {
        const bool wall = o == "wall_viscosity_interaction";
        if (wall && (P.wall_type != MPS_WALL_POLYGON || (int)P.dim != 3)) {
            ctx->err = "the wall rheology applies to polygon walls in 3-D";
            return MPS_E_UNSUPPORTED;
        }
        const BedArgs b = bedArgs(ctx, wall);
        hipLaunchKernelGGL(kBedInit, dim3(gridFor(b.cells, kGB)), dim3(kGB), 0, ctx->stream, b);
        hipLaunchKernelGGL(kBedLevels, dim3(g), dim3(kGB), 0, ctx->stream, a, b);
        if (!wall) {
            hipLaunchKernelGGL(kBedIndex, dim3(g), dim3(kGB), 0, ctx->stream, a, b);
            GL(kStrainRate);
            hipLaunchKernelGGL(kRheology, dim3(g), dim3(kGB), 0, ctx->stream, a, b, (const int*)ctx->slotOf,
                               ctx->timeCurrent);
        } else {
            GL(kWallStrainRate);
            hipLaunchKernelGGL(kWallRheology, dim3(g), dim3(kGB), 0, ctx->stream, a, b, ctx->timeCurrent);
        }
    } else {
        ctx->err = "unknown granular operator " + o;
        return MPS_E_ARG;
    }
#undef GL
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

// CFLvisc of the last fluid rheology (src/MpsViscosity.cpp:691)
int granularCflVisc(Ctx* ctx, double* out) {
    *out = 0.0;
    if (ctx->P.fluid_type != 1 || !ctx->bedKeys) return MPS_OK;
    const int cells = ctx->bedKx * ctx->bedKy;
    unsigned long long k = 0;
    MPS_HIP(hipMemcpyAsync(&k, ctx->bedKeys + 2 * cells, sizeof(k), hipMemcpyDeviceToHost, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    double miMax = 0.0;
    if (k != 0ull) {
        const unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
        memcpy(&miMax, &b, sizeof(miMax));
        if (miMax < 0.0) miMax = 0.0;
    }
    const mps_params& P = ctx->P;
    *out = P.time_step * miMax / ((P.dns_sdt * P.vf_max + (1.0 - P.vf_max) * P.dns_fl1) * P.part_dist * P.part_dist);
    return MPS_OK;
}

}  // namespace mpsgpu
