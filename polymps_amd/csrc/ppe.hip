// Pressure Poisson equation on the device.
//
// Assembly follows MpsPressure::solvePressurePoissonPnd / ...PndDivU
// (src/MpsPressure.cpp:169-421): inner rows get coeffPPE w(r_ij) for inner,
// non-dummy neighbours, the diagonal collects -coeffPPE w over all non-dummy
// neighbours minus alpha rho / dt^2, dummy neighbours move a hydrostatic term
// into the right-hand side; surface / other rows are identity rows with zero
// right-hand side.  The matrix is stored ELL (column-major [K][N]) without
// its diagonal, which is kept apart for the Jacobi preconditioner.
//
// The solver is the preconditioned CG of Eigen's ConjugateGradient with a
// DiagonalPreconditioner (the solver the reference calls at
// src/MpsPressure.cpp:424-445): same start residual, same stopping rule on the
// recursively updated residual (|r|^2 < max(tol^2 |b|^2, DBL_MIN)), same
// iteration count convention.  Dot products are FP64 reductions in a fixed
// order, so a solve is deterministic from run to run.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <vector>

#include "kern_common.h"

namespace mpsgpu {

constexpr int kB = 256;

// LDS window of a 256-row block for the SpMVs of the CG solves.  The
// columns of a block's rows lie in the union, over its rows, of their
// stencil rows' slot ranges (three in 2-D, nine in 3-D); merged into
// disjoint ascending segments these hold a few thousand slots, which the
// SpMV copies from the vector into LDS with coalesced loads before its rows
// gather from LDS instead of from the caches (the gathers of a wave scatter
// over a stencil row's slots, so from global memory each touches many cache
// lines).  The assembly writes, next to the column code, each entry's index
// in its block's window, and a table per block:
//   [0] 1 = no window (wider than kWinMax, a periodic domain, or a column
//       outside it): the block's rows gather from global memory;
//   [1] segment count m, [2 + k] first slot of segment k,
//   [11 + k] its LDS offset (k = 0 .. m; offset m = the window's size).
constexpr int kWinMax = 5000;    // slots: 20 KB of FP32 / 40 KB of FP64 values per block
constexpr int kWinStride = 24;   // ints per block in the table

// The block's window from its rows' current stencil rows (rows with
// entries only), merged; every thread returns the segments in sS/sE/sB.
__device__ __forceinline__ void ppeWindowBegin(const KArgs& a, int i, int* sS, int* sE, int* sB, int& sM,
                                               int& sOut) {
    __shared__ int ps[kB / 64][9], pe[kB / 64][9];
    const mps_params& P = a.P;
    const int rows = ((int)P.dim == 3) ? 9 : 3;
    bool has = i < a.N && live(a, i);
    if (has) {
        const int bci = a.I[IBC][i];
        has = !(bci == P.other || bci == P.surface);
    }
    const double xi = has ? a.D[DX][i] : 0.0, yi = has ? a.D[DY][i] : 0.0, zi = has ? a.D[DZ][i] : 0.0;
    const int ix = (int)((xi - P.domain_min[0]) * P.inv_bucket_side + KEPS);
    const int iy = (int)((yi - P.domain_min[1]) * P.inv_bucket_side + KEPS);
    const int iz = (int)((zi - P.domain_min[2]) * P.inv_bucket_side + KEPS);
    const int w = threadIdx.x >> 6;
    for (int r = 0; r < rows; r++) {
        int s = 0, e = 0;
        if (has) {
            const int jz = ((int)P.dim == 3) ? iz - 1 + r / 3 : 0;
            const int jy = iy - 1 + r % 3;
            if (!(jz < 0 || jz >= P.num_buckets_z || jy < 0 || jy >= P.num_buckets_y)) {
                const int row = jz * P.num_buckets_xy + jy * P.num_buckets_x;
                int jx0 = ix - 1, jx1 = ix + 1;
                if (jx0 < 0) jx0 = 0;
                if (jx1 >= P.num_buckets_x) jx1 = P.num_buckets_x - 1;
                if (jx0 <= jx1) {
                    s = a.cellStart[row + jx0];
                    e = a.cellStart[row + jx1 + 1];
                }
            }
        }
        const int mn = waveMinInt(e > s ? s : INT_MAX);
        const int mx = waveMaxInt(e > s ? e : INT_MIN);
        if ((threadIdx.x & 63) == 0) {
            ps[w][r] = mn;
            pe[w][r] = mx;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // per stencil row the union over the block; starts and ends ascend
        // with the row, so overlapping neighbours merge in one pass
        int m = 0, W = 0;
        for (int r = 0; r < rows; r++) {
            int s = INT_MAX, e = INT_MIN;
            for (int q = 0; q < (int)(blockDim.x >> 6); q++) {
                s = min(s, ps[q][r]);
                e = max(e, pe[q][r]);
            }
            if (s >= e) continue;
            if (m > 0 && s <= sE[m - 1]) {
                if (e > sE[m - 1]) {
                    W += e - sE[m - 1];
                    sE[m - 1] = e;
                }
            } else {
                sS[m] = s;
                sE[m] = e;
                sB[m] = W;
                W += e - s;
                m++;
            }
        }
        sB[m] = W;
        sM = m;
        sOut = (W > kWinMax || periodicDomain(P)) ? 1 : 0;
    }
    __syncthreads();
}

// RES: the row also gives the FP64 true residual r_i = b_i - (A x)_i of the
// initial guess x (the refinement loop's first residual, which kTrueResidual
// would otherwise compute from the stored row), and |r_i|^2, |b_i|^2 for the
// block partials.  kTrueResidual sums diag x_i first; here the off-diagonal
// products are summed first: the same FP64 quantity, rounded differently.
// lc8 / lcEsc (with a window): each entry's index in the block's LDS window
// (segments wS/wE/wB, wM of them) as a one-byte step from the previous
// entry's, or an escape to the next of the row's absolute indices; outside =
// some column is not in the window or the row's indices do not fit the code.
template <bool LIST, bool TRANS, bool RES>
__device__ __forceinline__ void assembleRow(const KArgs& a, int i, double* val, unsigned short* col, int* esc,
                                            int* cnt, double* diag, double* b, float* val32, float* diag32,
                                            float* invd32, double* invd64, double* x, double* r, double& rn2,
                                            double& bn2, unsigned char* lc8, unsigned short* lcEsc, const int* wS,
                                            const int* wE, const int* wB, int wM, bool& outside, int& nEsc,
                                            int& nLc) {
    if (!live(a, i)) {
        cnt[i] = 0; diag[i] = 1.0; b[i] = 0.0;
        if (invd32) { invd32[i] = 1.0f; diag32[i] = 1.0f; }
        if (RES) { x[i] = 0.0; r[i] = 0.0; }
        if (invd64) invd64[i] = 1.0;
        return;
    }
    const mps_params& P = a.P;
    const int bci = a.I[IBC][i];
    if (bci == P.other || bci == P.surface) {
        cnt[i] = 0; diag[i] = 1.0; b[i] = 0.0;
        if (invd32) { invd32[i] = 1.0f; diag32[i] = 1.0f; }
        if (invd64) invd64[i] = 1.0;
        if (RES) {
            const double ri = 0.0 - 1.0 * x[i];
            r[i] = ri;
            rn2 = ri * ri;
        }
        return;
    }
    double ax = 0.0;   // sum of the row's off-diagonal products with x (RES)
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = a.D[DX][i], yi = a.D[DY][i], zi = a.D[DZ][i];
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    const double rho = phaseDensity(a, i);   // DNS_FL1 or DNS_FL2 by phase (src/MpsPressure.cpp:237-239)
    // A neighbour j whose wall distance is at least 2 reL has its mirror point
    // at least 3 reL away from any particle of its reL ball, so the mirror test
    // of row j passes for i without reading j's mirror point.
    const double farWall2 = 4.0 * P.re_l2;
    const int* __restrict__ BC = a.I[IBC];
    const double* __restrict__ RIW2 = a.D[DRIW2];
    double sum = 0.0, rhs = 0.0;
    int c = 0, prev = -1, ne = 0;
    const int KE = ellWidth(a.K);
    bool full = false;
    typedef double D2 __attribute__((ext_vector_type(2)));
    double pend = 0.0;   // FP64 value of the last even entry, stored with the next one
    typedef float F4 __attribute__((ext_vector_type(4)));
    typedef unsigned short U4 __attribute__((ext_vector_type(4)));
    F4 g32 = {0.0f, 0.0f, 0.0f, 0.0f};   // the current group of four FP32 values / codes
    U4 gcd = {0, 0, 0, 0};
    // window index codes: eight one-byte codes per group (lc8), the escapes'
    // absolute indices in lcEsc
    unsigned g8lo = 0, g8hi = 0;
    int prevLc = -1, nle = 0;
    // the window segment of the current column (columns ascend)
    int wk = 0, curS = 0, curE = 0, curB = 0;
    if (lc8 && wM > 0) {
        curS = wS[0];
        curE = wE[0];
        curB = wB[0];
    }
    const bool periodic = periodicDomain(P);
    // with the list, the inner and far-from-wall tests of j come with its
    // position (P4.w flags of update_bc) instead of two more gathers
    const bool flags = LIST && a.p4Flags;
    auto entry = [&](int j, double xj, double yj, double zj, int tj, int fj, auto xOf) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        // Entry (j, i) of A is formed in row j of the reference, where i is
        // seen through the opposite ghost layer: (x_i - pl) - x_j, whose
        // rounding differs from -(x_j + pl - x_i) (periodic domains only)
        double d2T = d2;
        int pbc = 0;   // ghost shift of j as seen from i (periodic domains)
        if (TRANS && periodic) {
            const int pd = P.periodic_dir[0] ? 0 : (P.periodic_dir[1] ? 1 : 2);
            const double raw = a.D[DX + pd][j], sh = (pd == 0) ? xj : ((pd == 1) ? yj : zj);
            pbc = sh > raw ? 1 : (sh < raw ? -1 : 0);
            if (pbc != 0) {
                const double u0 = (xi + (-pbc) * P.periodic_length[0]) - a.D[DX][j];
                const double u1 = (yi + (-pbc) * P.periodic_length[1]) - a.D[DY][j];
                const double u2 = (zi + (-pbc) * P.periodic_length[2]) - a.D[DZ][j];
                d2T = u0 * u0 + u1 * u1 + u2 * u2;
            }
        }
        if (!(d2 < P.re_l2)) return;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        const double wL = wgt(sqrt(d2), P.re_l, P.weight_type);
        const double mij = wL * P.coeff_ppe;
        const bool okRow = d2 < dm2 || pw;   // j is a neighbour of row i
        const bool dummy = tj == P.dummy_wall;
        // both accumulators are updated every time (x - 0.0 == x) so that no
        // accumulator is selected through memory
        double subSum = 0.0, subRhs = 0.0;
        if (okRow && dummy) {
            const double pgh = rhoOf(a, j) * (v0 * P.gravity[0] + v1 * P.gravity[1] + v2 * P.gravity[2]) * wL;
            subRhs = P.coeff_ppe * pgh;
        }
        if (okRow && !dummy) subSum = mij;
        sum -= subSum;
        rhs -= subRhs;
        bool put = false;
        const bool innerJ = flags ? (fj & 1) != 0 : BC[j] == P.inner;
        if (!TRANS) {
            put = okRow && !dummy && innerJ;
        } else if (innerJ) {
            // entry (j, i) of A: i seen from row j (through the opposite
            // ghost layer when j is across a periodic seam), mirror test
            // with j's mirror
            put = pw || (flags ? (fj & 2) != 0 : RIW2[j] >= farWall2);
            if (!put) {
                const double ix = xi + (-pbc) * P.periodic_length[0], iy = yi + (-pbc) * P.periodic_length[1],
                             iz = zi + (-pbc) * P.periodic_length[2];
                const double u0 = ix - a.D[DMX][j], u1 = iy - a.D[DMY][j], u2 = iz - a.D[DMZ][j];
                const double dmj = u0 * u0 + u1 * u1 + u2 * u2;
                put = d2T < dmj;
            }
        }
        if (put) {
            // column code: delta from the previous entry (walk order is
            // ascending slot order), 0 = the slot is in the escape table.  A
            // row that does not fit keeps its first entries (consistent for
            // the solver) and is reported as a capacity error.
            const int dj = j - prev;
            const bool delta = prev >= 0 && dj > 0 && dj < 65536;
            if (c < a.K && (delta || ne < kPpeEsc)) {
                const double mv = (TRANS && d2T != d2) ? wgt(sqrt(d2T), P.re_l, P.weight_type) * P.coeff_ppe : mij;
                // FP64 values leave in pairs (one 16-byte store per two entries)
                if (c & 1) {
                    D2 w;
                    w.x = pend;
                    w.y = mv;
                    reinterpret_cast<D2*>(val)[ellElem(c >> 1, i, KE >> 1)] = w;
                } else {
                    pend = mv;
                }
                // FP32 values and codes leave in groups of four (ellIdx4)
                const float f = (float)mv;
                const unsigned short cd = delta ? (unsigned short)dj : (unsigned short)0;
                if (lc8) {
                    while (wk < wM && j >= curE) {
                        wk++;
                        if (wk < wM) {
                            curS = wS[wk];
                            curE = wE[wk];
                            curB = wB[wk];
                        }
                    }
                    int lc = -1;
                    if (wk < wM && j >= curS) lc = curB + (j - curS);
                    else outside = true;
                    // code: the step from the previous entry's index (1-255)
                    // or 0, an escape to the next absolute index of the row
                    // (at most kLcEsc per row, or the block gathers from
                    // global memory)
                    if ((c & 7) == 0) {
                        g8lo = 0;
                        g8hi = 0;
                    }
                    const int dl = lc - prevLc;
                    unsigned code = 0;
                    if (prevLc >= 0 && lc >= 0 && dl > 0 && dl < 256) {
                        code = (unsigned)dl;
                    } else {
                        if (lc >= 0 && nle < kLcEsc) lcEsc[ellElem(nle, i, kLcEsc)] = (unsigned short)lc;
                        else outside = true;
                        nle++;
                    }
                    prevLc = lc;
                    if (c & 4) g8hi |= code << (8 * (c & 3));
                    else g8lo |= code << (8 * (c & 3));
                    if ((c & 7) == 7) reinterpret_cast<uint2*>(lc8)[ellElem(c >> 3, i, KE >> 3)] = make_uint2(g8lo, g8hi);
                }
                switch (c & 3) {
                    case 0: g32.x = f; gcd.x = cd; break;
                    case 1: g32.y = f; gcd.y = cd; break;
                    case 2: g32.z = f; gcd.z = cd; break;
                    default: g32.w = f; gcd.w = cd; break;
                }
                if ((c & 3) == 3) {
                    const size_t gi = ellElem(c >> 2, i, KE >> 2);
                    if (val32) reinterpret_cast<F4*>(val32)[gi] = g32;
                    reinterpret_cast<U4*>(col)[gi] = gcd;
                }
                if (RES) ax += mv * xOf();
                if (!delta) {
                    esc[(size_t)ne * a.N + i] = j;
                    ne++;
                }
                prev = j;
                c++;
            } else {
                full = true;
            }
        }
    };
    if (LIST && RES) {
        // the guess x_j of the row's residual is gathered with the batch's positions
        listNbrsX(a, a.nbr, a.nbrCnt, i, [&](int j) { return x[j]; }, [&](int j, const double4& q, double xv) {
            entry(j, q.x, q.y, q.z, p4TypeOf(a, q.w), a.p4Flags ? p4FlagBits(q.w) : 0, [xv]() { return xv; });
        });
    } else {
        forNbrsPosTypeFlags<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj, int tj, int fj) {
            entry(j, xj, yj, zj, tj, fj, [x, j]() { return x[j]; });
        });
    }
    if (c & 1) val[ellIdx64(c - 1, i, KE)] = pend;
    if (c & 3) {   // the last, partial group (entries past the count are never read)
        const size_t gi = ellElem(c >> 2, i, KE >> 2);
        if (val32) reinterpret_cast<F4*>(val32)[gi] = g32;
        reinterpret_cast<U4*>(col)[gi] = gcd;
    }
    if (lc8 && (c & 7)) reinterpret_cast<uint2*>(lc8)[ellElem(c >> 3, i, KE >> 3)] = make_uint2(g8lo, g8hi);
    if (full) atomicMax(&a.is[I_OVERFLOW], c + 1);   // reported as a capacity error after the step
    nEsc = ne;
    nLc = lc8 ? nle : 0;
    sum -= P.alpha_compressibility * rho / (P.time_step * P.time_step);
    if (P.mps_type == MPS_IMPLICIT_PND_DIVU)
        rhs += -P.coeff_ppe_source * rho * (a.D[DPNDKI][i] - P.pnd_small_zero) +
               (1.0 - P.relax_pnd) * rho * a.D[DDIVU][i] / P.time_step;
    else
        rhs += -P.coeff_ppe_source * rho * (a.D[DPNDSMALL][i] - P.pnd_small_zero);
    cnt[i] = c;
    diag[i] = sum;
    b[i] = rhs;
    if (invd32) {
        invd32[i] = (sum != 0.0) ? (float)(1.0 / sum) : 1.0f;
        diag32[i] = (float)sum;
    }
    if (invd64) invd64[i] = (sum != 0.0) ? 1.0 / sum : 1.0;
    if (RES) {
        const double ri = rhs - (sum * x[i] + ax);
        r[i] = ri;
        rn2 = ri * ri;
        bn2 = rhs * rhs;
    }
}

template <bool LIST, bool TRANS, bool RES>
__global__ __launch_bounds__(256) void kPpeAssemble(KArgs a, double* val, unsigned short* col, int* esc, int* cnt,
                                                    double* diag, double* b, float* val32, float* diag32,
                                                    float* invd32, double* invd64, double* x, double* r,
                                                    int bFirst, unsigned char* lc8, unsigned short* lcEsc, int* win);

// ------------------------------------------------------------ reductions
// Every vector kernel runs one thread per row; each block leaves a pair of
// FP64 partial sums in red[2 * block], and a one-block kernel adds the
// partials in block order, so every reduction is deterministic.
__device__ __forceinline__ void blockReduce2(double& s0, double& s1) {
    __shared__ double sh[2][16];
    s0 = waveSum(s0);
    s1 = waveSum(s1);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sh[0][w] = s0; sh[1][w] = s1; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0, t1 = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) { t0 += sh[0][q]; t1 += sh[1][q]; }
        s0 = t0; s1 = t1;
    }
}
__device__ __forceinline__ void writePartials(const KArgs& a, double s0, double s1) {
    blockReduce2(s0, s1);
    if (threadIdx.x == 0) { a.red[blockIdx.x * 2] = s0; a.red[blockIdx.x * 2 + 1] = s1; }
}

// lc8 / lcEsc / win: window index codes and escapes, the block's window table
template <bool LIST, bool TRANS, bool RES>
__global__ __launch_bounds__(256) void kPpeAssemble(KArgs a, double* val, unsigned short* col, int* esc, int* cnt,
                                                    double* diag, double* b, float* val32, float* diag32,
                                                    float* invd32, double* invd64, double* x, double* r,
                                                    int bFirst, unsigned char* lc8, unsigned short* lcEsc, int* win) {
    __shared__ int wS[9], wE[9], wB[10], wM, wOut;
    const int i = slotId();
    if (lc8) ppeWindowBegin(a, i, wS, wE, wB, wM, wOut);
    double rn2 = 0.0, bn2 = 0.0;
    bool outside = false;
    int nEsc = 0, nLc = 0;
    if (i < a.N)
        assembleRow<LIST, TRANS, RES>(a, i, val, col, esc, cnt, diag, b, val32, diag32, invd32, invd64, x, r, rn2,
                                      bn2, lc8, lcEsc, wS, wE, wB, lc8 ? wM : 0, outside, nEsc, nLc);
    // escaped columns and window index escapes: one atomic per wavefront
    nEsc = waveSumInt(nEsc);
    nLc = waveSumInt(nLc);
    if ((threadIdx.x & 63) == 0) {
        if (nEsc) atomicAdd(&a.is[I_NESC], nEsc);
        if (nLc) atomicAdd(&a.is[I_NLCESC], nLc);
    }
    if (lc8) {
        if (__builtin_amdgcn_ballot_w64(outside) != 0 && (threadIdx.x & 63) == 0) atomicOr(&wOut, 1);
        __syncthreads();
        if (threadIdx.x == 0) {
            int* t = win + (size_t)blockIdx.x * kWinStride;
            t[0] = wOut;
            t[1] = wM;
            for (int k = 0; k < wM; k++) t[2 + k] = wS[k];
            for (int k = 0; k <= wM; k++) t[11 + k] = wB[k];
        }
    }
    // the partials kTrueResidualFinish (|r|^2, |b|^2) or kCgStartFinish (|b|^2, |r|^2) sums
    if (RES) {
        if (bFirst) writePartials(a, bn2, rn2);
        else writePartials(a, rn2, bn2);
    }
}
// three sums, partials at red[3 * block]
__device__ __forceinline__ void writePartials3(const KArgs& a, double s0, double s1, double s2) {
    __shared__ double sh[3][16];
    s0 = waveSum(s0);
    s1 = waveSum(s1);
    s2 = waveSum(s2);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sh[0][w] = s0; sh[1][w] = s1; sh[2][w] = s2; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0, t1 = 0, t2 = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); q++) { t0 += sh[0][q]; t1 += sh[1][q]; t2 += sh[2][q]; }
        a.red[blockIdx.x * 3] = t0; a.red[blockIdx.x * 3 + 1] = t1; a.red[blockIdx.x * 3 + 2] = t2;
    }
}
__device__ __forceinline__ void finishPartials3(const double* red, int nb, double& s0, double& s1, double& s2) {
    __shared__ double sh[3][16];
    double t0 = 0, t1 = 0, t2 = 0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) { t0 += red[q * 3]; t1 += red[q * 3 + 1]; t2 += red[q * 3 + 2]; }
    t0 = waveSum(t0);
    t1 = waveSum(t1);
    t2 = waveSum(t2);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sh[0][w] = t0; sh[1][w] = t1; sh[2][w] = t2; }
    __syncthreads();
    s0 = 0; s1 = 0; s2 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) { s0 += sh[0][q]; s1 += sh[1][q]; s2 += sh[2][q]; }
}
__device__ __forceinline__ void finishPartials(const double* red, int nb, double& s0, double& s1) {
    __shared__ double sh[2][16];
    double t0 = 0, t1 = 0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) { t0 += red[q * 2]; t1 += red[q * 2 + 1]; }
    t0 = waveSum(t0);
    t1 = waveSum(t1);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sh[0][w] = t0; sh[1][w] = t1; }
    __syncthreads();
    s0 = 0; s1 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) { s0 += sh[0][q]; s1 += sh[1][q]; }
}

// y_i = diag_i x_i + sum_c val(c, i) x[j_c]; the column of entry c is the
// previous one plus code(c, i), or the next slot of the escape table when the
// code is 0 (the first entry of every row is escaped: escape 0 is the row's
// first column).
//
// A row is read in batches of B entries (16 FP32, 8 FP64), software-
// pipelined so that a lane's streamed loads and its gathers overlap: the
// codes of batch b + 1 are loaded right behind the gathers of batch b and
// decoded after batch b's products, and the values of batch b + 1 (FP64 as
// pair loads, ellIdx64; FP32 values and the codes as loads of four, ellIdx4)
// are loaded together with its gathers.  The memory counter retires loads
// in issue order, so every wait is for the oldest loads in flight and none
// waits for a load issued after the one it needs.  The row's first escape is
// loaded with the row, so a batch decodes its columns without a dependent
// load unless some lane of the wave meets a second escaped entry in it.  The
// last batch is padded: the arrays hold ellWidth(K) entries per row (K, a
// multiple of 16, and one batch more), so the loads past a row stay inside
// them, and the entries past the count get value 0 and the previous column.
// All lanes of a wave are at the same batch, so a batch's stream addresses
// are a wave-uniform base plus the lane's 32-bit offset, and a gather is the
// vector's base plus a 32-bit offset.  Value and code streams are read once
// per launch and are loaded non-temporally so that they do not displace the
// gathered x from the caches.
typedef unsigned short EllU4 __attribute__((ext_vector_type(4)));
typedef unsigned EllW2 __attribute__((ext_vector_type(2)));

// element g of row i of an ELL array of 16- or 8-byte elements, G per row
// (ellElem): a wave-uniform tile base plus the lane's 32-bit offset
template <class E>
__device__ __forceinline__ E ellStream(const void* base, int g, int G, int i) {
    const char* b = reinterpret_cast<const char*>(base) + (((size_t)(i >> 6) * (size_t)G + (size_t)g) << 6) * sizeof(E);
    return __builtin_nontemporal_load(reinterpret_cast<const E*>(b + (unsigned)(i & 63) * (unsigned)sizeof(E)));
}
template <class T>
__device__ __forceinline__ T ellGather(const T* __restrict__ x, int j) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(x) + (unsigned)j * (unsigned)sizeof(T));
}

// values of entries c0 .. c0 + B - 1 (c0 a multiple of B)
template <class T, int B>
__device__ __forceinline__ void ellLoadValues(const T* __restrict__ val, int KE, int i, int c0, T* v) {
    typedef double D2 __attribute__((ext_vector_type(2)));
    typedef float F4 __attribute__((ext_vector_type(4)));
    if (sizeof(T) == 8) {
        // pair m of row i is element m * N + i of the pair view
#pragma unroll
        for (int k = 0; k < B / 2; k++) {
            const D2 w = ellStream<D2>(val, (c0 >> 1) + k, KE >> 1, i);
            v[2 * k] = (T)w.x;
            v[2 * k + 1] = (T)w.y;
        }
    } else {
        // group g of row i is element g * N + i of the group view
#pragma unroll
        for (int k = 0; k < B / 4; k++) {
            const F4 w = ellStream<F4>(val, (c0 >> 2) + k, KE >> 2, i);
            v[4 * k] = (T)w.x;
            v[4 * k + 1] = (T)w.y;
            v[4 * k + 2] = (T)w.z;
            v[4 * k + 3] = (T)w.w;
        }
    }
}

template <int B>
__device__ __forceinline__ void ellLoadCodes(const unsigned short* __restrict__ code, int KE, int i, int c0,
                                             EllU4* d) {
#pragma unroll
    for (int k = 0; k < B / 4; k++) d[k] = ellStream<EllU4>(code, (c0 >> 2) + k, KE >> 2, i);
}

// columns jj of a batch with rem valid entries (past them: the previous
// column); j = the last column, e = escapes used, es0 = the first escape
template <int B>
__device__ __forceinline__ void ellDecode(const EllU4* dv, int rem, const int* __restrict__ es, size_t N, int es0,
                                          int& j, int& e, int* jj) {
    unsigned d[B];
#pragma unroll
    for (int k = 0; k < B / 4; k++) {
        d[4 * k] = dv[k].x;
        d[4 * k + 1] = dv[k].y;
        d[4 * k + 2] = dv[k].z;
        d[4 * k + 3] = dv[k].w;
    }
    int zc = 0;
#pragma unroll
    for (int k = 0; k < B; k++) zc += (k < rem && d[k] == 0u) ? 1 : 0;
    int jl = j;
    if (__builtin_amdgcn_ballot_w64(e + zc > 1) != 0) {
        // a second escape in some lane's row: read from the table
        int el = e;
#pragma unroll
        for (int k = 0; k < B; k++) {
            if (k < rem) {
                int t = jl + (int)d[k];
                if (d[k] == 0u) {
                    t = (el == 0) ? es0 : es[(size_t)el * N];
                    el++;
                }
                jl = t;
            }
            jj[k] = jl;
        }
    } else {
#pragma unroll
        for (int k = 0; k < B; k++) {
            if (k < rem) jl = d[k] ? jl + (int)d[k] : es0;
            jj[k] = jl;
        }
    }
    j = jl;
    e += zc;
}

// acc += the batch's products in entry order (entries k >= rem add nothing);
// FP32 (inner solve): each run of eight is summed in FP32, then added in FP64
template <class T, int B>
__device__ __forceinline__ void ellAccum(const T* v, const T* xv, int rem, double& acc) {
    if (sizeof(T) == 4) {
#pragma unroll
        for (int h = 0; h < B; h += 8) {
            T b8 = (h < rem) ? v[h] * xv[h] : (T)0;
#pragma unroll
            for (int k = h + 1; k < h + 8; k++) b8 += (k < rem) ? v[k] * xv[k] : (T)0;
            if (h < rem) acc += (double)b8;
        }
    } else {
#pragma unroll
        for (int k = 0; k < B; k++)
            if (k < rem) acc += (double)(v[k] * xv[k]);
    }
}

template <class T>
__device__ __forceinline__ double ellRow(const T* __restrict__ val, const unsigned short* __restrict__ code,
                                         const int* __restrict__ esc, int n, size_t N, int KE,
                                         const T* __restrict__ x, int i, double acc) {
    constexpr int B = 8;
    if (n <= 0) return acc;
    const int* es = esc + i;
    const int es0 = es[0];
    int j = 0, e = 0;
    T v[B];
    EllU4 d[B / 4];
    int jj[B];
    ellLoadValues<T, B>(val, KE, i, 0, v);
    ellLoadCodes<B>(code, KE, i, 0, d);
    ellDecode<B>(d, n, es, N, es0, j, e, jj);
    for (int c = 0;;) {
        T xv[B];
#pragma unroll
        for (int k = 0; k < B; k++) xv[k] = ellGather(x, jj[k]);
        ellLoadCodes<B>(code, KE, i, c + B, d);
        ellAccum<T, B>(v, xv, n - c, acc);
        ellDecode<B>(d, n - c - B, es, N, es0, j, e, jj);
        c += B;
        if (c >= n) break;
        ellLoadValues<T, B>(val, KE, i, c, v);
    }
    return acc;
}

struct CgPtrs {
    const double* val;
    const unsigned short* col;
    const int* esc;
    const int* cnt;
    const double* diag;
    const double* b;
    double* x;
    double* r;
    double* p;
    double* z;
    double* q;
    double* invd;
    // mixed precision
    const float* val32;
    const float* diag32;
    const float* invd32;
    float* fr;
    float* fp;
    float* fz;
    float* fq;
    // the correction d = sum of alpha p, summed in FP64: stored in FP32 its
    // rounding alone leaves a residual of order 3e-8 |A| |d|, above 1e-8 |b|
    // for the larger corrections of a developed flow
    double* fd;
    // window indices and the per-block window table (null: no windows)
    const unsigned char* lc8;
    const unsigned short* lcEsc;
    const int* win;
};

// Row from the block's LDS window, in the order of ellRow (FP32: batches of
// eight summed in FP32, then added in FP64; FP64: entry by entry), so both
// give the same bits.
template <class T>
__device__ __forceinline__ double ellRowWin(const T* __restrict__ val, const unsigned char* __restrict__ lc8,
                                            const unsigned short* __restrict__ lcEsc, int n, int KE, const T* win,
                                            int i, double acc) {
    typedef float F4 __attribute__((ext_vector_type(4)));
    typedef double D2 __attribute__((ext_vector_type(2)));
    int lc = 0, k = 0;   // running window index, escapes taken
    for (int c = 0; c < n; c += 8) {
        T v[8];
        if (sizeof(T) == 4) {
            const F4 v0 = ellStream<F4>(val, c >> 2, KE >> 2, i), v1 = ellStream<F4>(val, (c >> 2) + 1, KE >> 2, i);
            v[0] = (T)v0.x; v[1] = (T)v0.y; v[2] = (T)v0.z; v[3] = (T)v0.w;
            v[4] = (T)v1.x; v[5] = (T)v1.y; v[6] = (T)v1.z; v[7] = (T)v1.w;
        } else {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const D2 w = ellStream<D2>(val, (c >> 1) + q, KE >> 1, i);
                v[2 * q] = (T)w.x;
                v[2 * q + 1] = (T)w.y;
            }
        }
        // one byte per entry; the lane's next two escape indices are loaded
        // with the group (a third escape in one group loads its own)
        const EllW2 cw = ellStream<EllW2>(lc8, c >> 3, KE >> 3, i);
        const unsigned e0 = lcEsc[ellElem(min(k, kLcEsc - 1), i, kLcEsc)];
        const unsigned e1 = lcEsc[ellElem(min(k + 1, kLcEsc - 1), i, kLcEsc)];
        const int rem = n - c;
        const int k0 = k;
        unsigned kk[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const unsigned b = ((q < 4 ? cw.x : cw.y) >> (8 * (q & 3))) & 0xffu;
            if (q < rem) {
                if (b != 0u) {
                    lc += (int)b;
                } else {
                    lc = (int)(k == k0 ? e0 : (k == k0 + 1 ? e1 : lcEsc[ellElem(min(k, kLcEsc - 1), i, kLcEsc)]));
                    k++;
                }
            }
            kk[q] = (unsigned)lc;
        }
        if (sizeof(T) == 4) {
            T b8 = v[0] * win[kk[0]];
#pragma unroll
            for (int q = 1; q < 8; q++) b8 += (q < rem) ? v[q] * win[kk[q]] : (T)0;
            acc += (double)b8;
        } else {
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (q < rem) acc += (double)(v[q] * win[kk[q]]);
        }
    }
    return acc;
}

// Copies the block's window of vector x into LDS when the block has one
// (block-wide; false: the block's rows gather from global memory).
template <class T>
__device__ __forceinline__ bool stageWindow(const int* table, const T* __restrict__ x, T* win) {
    const int* t = table ? table + (size_t)blockIdx.x * kWinStride : nullptr;
    if (t == nullptr || t[0] != 0) return false;
    const int m = t[1];
    for (int k = 0; k < m; k++) {
        const int S = t[2 + k], B0 = t[11 + k], len = t[12 + k] - B0;
        for (int q = threadIdx.x; q < len; q += kB) win[B0 + q] = x[S + q];
    }
    __syncthreads();
    return true;
}

__device__ __forceinline__ double spmv64(const KArgs& a, const CgPtrs& c, const double* x, int i) {
    return ellRow<double>(c.val, c.col, c.esc, c.cnt[i], a.N, ellWidth(a.K), x, i, c.diag[i] * x[i]);
}

// ---------------------------------------------------------- FP64 PCG
// r = b - M x; partials |b|^2, |r|^2; invd = 1/diag (1 where diag == 0)
__global__ void kCgStart(KArgs a, CgPtrs c) {
    const int i = slotId();
    double s0 = 0, s1 = 0;
    if (i < a.N) {
        if (!live(a, i)) c.x[i] = 0.0;   // slots beyond the live count do not take part
        const double d = c.diag[i];
        c.invd[i] = (d != 0.0) ? 1.0 / d : 1.0;
        const double r = c.b[i] - spmv64(a, c, c.x, i);
        c.r[i] = r;
        s0 = c.b[i] * c.b[i];
        s1 = r * r;
    }
    writePartials(a, s0, s1);
}

__global__ void kCgStartFinish(KArgs a, int nb, double tol) {
    double bn2, rn2;
    finishPartials(a.red, nb, bn2, rn2);
    if (threadIdx.x != 0) return;
    a.ds[S_RHSN2] = bn2;
    a.ds[S_RN2] = rn2;
    a.is[I_CG_ITER] = 0;
    a.is[I_CG_DONE] = 0;
    if (!isfinite(bn2) || !isfinite(rn2)) {
        a.is[I_CG_DONE] = 5;   // non-finite system or guess: nothing to iterate
        return;
    }
    if (bn2 == 0.0) {
        a.is[I_CG_DONE] = 2;   // x = 0
        return;
    }
    const double thr = fmax(tol * tol * bn2, DBL_MIN);
    a.ds[S_THR] = thr;
    if (rn2 < thr) a.is[I_CG_DONE] = 1;
}

__global__ void kCgZeroX(KArgs a, double* x) {
    const int i = slotId();
    if (a.is[I_CG_DONE] != 2 || i >= a.N) return;
    x[i] = 0.0;
}

// p = invd r; partial r.p
__global__ void kCgInitP(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    double s0 = 0;
    if (i < a.N) {
        const double p = c.invd[i] * c.r[i];
        c.p[i] = p;
        s0 = c.r[i] * p;
    }
    writePartials(a, s0, 0.0);
}
__global__ void kCgInitPFinish(KArgs a, int nb) {
    if (a.is[I_CG_DONE]) return;
    double s0, s1;
    finishPartials(a.red, nb, s0, s1);
    if (threadIdx.x == 0) a.ds[S_ABSNEW] = s0;
}

// q = M p; partial p.q
__global__ __launch_bounds__(256) void kCgSpmv(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    __shared__ double win[kWinMax];
    const bool lds = stageWindow<double>(c.win, c.p, win);
    const int i = slotId();
    double s0 = 0;
    if (i < a.N) {
        const double q = lds ? ellRowWin<double>(c.val, c.lc8, c.lcEsc, c.cnt[i], ellWidth(a.K), win, i, c.diag[i] * c.p[i])
                             : spmv64(a, c, c.p, i);
        c.q[i] = q;
        s0 = c.p[i] * q;
    }
    writePartials(a, s0, 0.0);
}
__global__ void kCgAlpha(KArgs a, int nb) {
    if (a.is[I_CG_DONE]) return;
    double s0, s1;
    finishPartials(a.red, nb, s0, s1);
    if (threadIdx.x == 0) {
        a.ds[S_PAP] = s0;
        a.ds[S_ALPHA] = a.ds[S_ABSNEW] / s0;
    }
}
// x += alpha p; r -= alpha q; z = invd r; partials |r|^2, r.z
__global__ void kCgUpdate(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    const double alpha = a.ds[S_ALPHA];
    double s0 = 0, s1 = 0;
    if (i < a.N) {
        c.x[i] += alpha * c.p[i];
        const double r = c.r[i] - alpha * c.q[i];
        c.r[i] = r;
        const double z = c.invd[i] * r;
        c.z[i] = z;
        s0 = r * r;
        s1 = r * z;
    }
    writePartials(a, s0, s1);
}
__global__ void kCgBeta(KArgs a, int nb, int maxIt) {
    if (a.is[I_CG_DONE]) return;
    double rn2, rz;
    finishPartials(a.red, nb, rn2, rz);
    if (threadIdx.x != 0) return;
    a.ds[S_RN2] = rn2;
    if (rn2 < a.ds[S_THR]) {
        a.is[I_CG_DONE] = 1;
        return;
    }
    if (!isfinite(rn2)) {
        a.is[I_CG_DONE] = 5;
        return;
    }
    const double absOld = a.ds[S_ABSNEW];
    a.ds[S_ABSOLD] = absOld;
    a.ds[S_ABSNEW] = rz;
    a.ds[S_BETA] = rz / absOld;
    const int it = a.is[I_CG_ITER] + 1;
    a.is[I_CG_ITER] = it;
    if (it >= maxIt) a.is[I_CG_DONE] = 3;
}
// p = z + beta p
__global__ void kCgP(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    if (i < a.N) c.p[i] = c.z[i] + a.ds[S_BETA] * c.p[i];
}

// true residual r = b - M x in FP64 (kept in c.r); partials |r|^2, |b|^2
__global__ __launch_bounds__(256) void kTrueResidual(KArgs a, CgPtrs c) {
    __shared__ double win[kWinMax];
    // (slots past the live count, zeroed below, are in no window)
    const bool lds = stageWindow<double>(c.win, c.x, win);
    const int i = slotId();
    double s0 = 0, s1 = 0;
    if (i < a.N) {
        if (!live(a, i)) c.x[i] = 0.0;
        const double r = c.b[i] - (lds ? ellRowWin<double>(c.val, c.lc8, c.lcEsc, c.cnt[i], ellWidth(a.K), win, i, c.diag[i] * c.x[i])
                                       : spmv64(a, c, c.x, i));
        c.r[i] = r;
        s0 = r * r;
        s1 = c.b[i] * c.b[i];
    }
    writePartials(a, s0, s1);
}
__global__ void kTrueResidualFinish(KArgs a, int nb) {
    double s0, s1;
    finishPartials(a.red, nb, s0, s1);
    if (threadIdx.x == 0) {
        a.ds[S_TRUE_RN2] = s0;
        a.ds[S_BN2] = s1;
    }
}

// --------------------------------------- mixed precision (inner FP32 PCG)
// Solves M d = r with FP32 values, SpMV and Jacobi preconditioner; every dot
// product accumulates in FP64.  Convergence: |r_inner|^2 < tol_in^2 |r_0|^2.
__global__ void kMxInit(KArgs a, CgPtrs c) {
    const int i = slotId();
    double s0 = 0, s1 = 0;
    if (i < a.N) {
        const float r = (float)c.r[i];
        const float z = c.invd32[i] * r;
        c.fr[i] = r;
        c.fd[i] = 0.0;
        c.fz[i] = z;
        c.fp[i] = z;
        s0 = (double)r * (double)z;
        s1 = (double)r * (double)r;
    }
    writePartials(a, s0, s1);
}
// The inner solve stops at tolIn relative to its start or as soon as its
// residual is below half the outer target (tol |b|), whichever comes first.
__global__ void kMxInitFinish(KArgs a, int nb, double tolIn, double tol) {
    double rz, rn2;
    finishPartials(a.red, nb, rz, rn2);
    if (threadIdx.x != 0) return;
    a.ds[S_ABSNEW] = rz;
    a.ds[S_RN2] = rn2;
    const double target = 0.25 * tol * tol * a.ds[S_BN2];
    a.ds[S_THR] = fmax(fmax(tolIn * tolIn * rn2, target), DBL_MIN);
    a.is[I_CG_ITER] = 0;
    a.is[I_CG_DONE] = (rn2 == 0.0) ? 2 : (isfinite(rn2) && isfinite(rz)) ? 0 : 5;
}
// q = M p: a block with a window first copies its segments of p into LDS
__global__ __launch_bounds__(256) void kMxSpmv(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    __shared__ float win[kWinMax];
    const int i = slotId();
    const bool lds = stageWindow<float>(c.win, c.fp, win);
    double s0 = 0;
    if (i < a.N) {
        const float p = c.fp[i];
        const double d = (double)(c.diag32[i] * p);
        const int n = c.cnt[i];
        const float q = lds ? (float)ellRowWin<float>(c.val32, c.lc8, c.lcEsc, n, ellWidth(a.K), win, i, d)
                            : (float)ellRow<float>(c.val32, c.col, c.esc, n, a.N, ellWidth(a.K), c.fp, i, d);
        c.fq[i] = q;
        s0 = (double)p * (double)q;
    }
    writePartials(a, s0, 0.0);
}
__global__ void kMxUpdate(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    const float alpha = (float)a.ds[S_ALPHA];
    double s0 = 0, s1 = 0;
    if (i < a.N) {
        c.fd[i] += (double)alpha * (double)c.fp[i];
        const float r = c.fr[i] - alpha * c.fq[i];
        c.fr[i] = r;
        const float z = c.invd32[i] * r;
        c.fz[i] = z;
        s0 = (double)r * (double)r;
        s1 = (double)r * (double)z;
    }
    writePartials(a, s0, s1);
}
__global__ void kMxP(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    if (i < a.N) c.fp[i] = c.fz[i] + (float)a.ds[S_BETA] * c.fp[i];
}
__global__ void kZero(KArgs a, double* x) {
    const int i = slotId();
    if (i < a.N) x[i] = 0.0;
}
// x += d
__global__ void kMxApply(KArgs a, CgPtrs c) {
    const int i = slotId();
    if (i < a.N && live(a, i)) c.x[i] += c.fd[i];
}

// ------------------------------------- flexible CG with the learned operator
// Inner solve of the mixed path with z = M r (M from the learned operator):
// p_{k+1} = z_{k+1} - (z_{k+1} . q_k) / (p_k . q_k) p_k, q_k = A p_k.
__global__ void kFcgInit(KArgs a, CgPtrs c) {
    const int i = slotId();
    if (i < a.N) {
        c.fr[i] = (float)c.r[i];
        c.fd[i] = 0.0;
    }
}
// after z = M r: p = z; partials r.z, |r|^2 (same finish as kMxInit)
__global__ void kFcgInitDots(KArgs a, CgPtrs c) {
    const int i = slotId();
    double s0 = 0, s1 = 0;
    if (i < a.N) {
        const float r = c.fr[i], z = c.fz[i];
        c.fp[i] = z;
        s0 = (double)r * (double)z;
        s1 = (double)r * (double)r;
    }
    writePartials(a, s0, s1);
}
__global__ void kFcgUpdate(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    const float alpha = (float)a.ds[S_ALPHA];
    if (i < a.N) {
        c.fd[i] += (double)alpha * (double)c.fp[i];
        c.fr[i] = c.fr[i] - alpha * c.fq[i];
    }
}
// partials |r|^2, r.z, z.q
__global__ void kFcgDots(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    double s0 = 0, s1 = 0, s2 = 0;
    if (i < a.N) {
        const double r = c.fr[i], z = c.fz[i], q = c.fq[i];
        s0 = r * r;
        s1 = r * z;
        s2 = z * q;
    }
    writePartials3(a, s0, s1, s2);
}
// stop test, breakdown test, beta.  The assembled matrix is negative
// definite on its coupled rows (diagonal -sum w_ij - alpha rho / dt^2,
// off-diagonals w_ij > 0), and so is a usable preconditioner (Jacobi: 1 /
// a_ii < 0): the next step length r.z / p.Ap stays positive while r.z has
// the sign of the last p.Ap, and a preconditioner that gives r.z of the
// other sign (or 0) has broken down.
__global__ void kFcgBeta(KArgs a, int nb, int maxIt) {
    if (a.is[I_CG_DONE]) return;
    double rn2, rz, zq;
    finishPartials3(a.red, nb, rn2, rz, zq);
    if (threadIdx.x != 0) return;
    a.ds[S_RN2] = rn2;
    if (rn2 < a.ds[S_THR]) {
        a.is[I_CG_DONE] = 1;
        return;
    }
    if (!isfinite(rn2)) {
        a.is[I_CG_DONE] = 5;
        return;
    }
    if (!(rz * a.ds[S_PAP] > 0.0)) {
        a.is[I_CG_DONE] = 4;
        return;
    }
    a.ds[S_ABSNEW] = rz;
    a.ds[S_BETA] = -zq / a.ds[S_PAP];
    const int it = a.is[I_CG_ITER] + 1;
    a.is[I_CG_ITER] = it;
    if (it >= maxIt) a.is[I_CG_DONE] = 3;
}

// press = max(x, 0) for live slots (setZeroOnNegativePressure :471-476)
__global__ void kPressFromX(KArgs a, const double* x) {
    const int i = slotId();
    if (!live(a, i)) return;
    const double v = x[i];
    a.D[DPRESS][i] = (v < 0) ? 0.0 : v;
}

static CgPtrs cgPtrs(Ctx* ctx) {
    CgPtrs c;
    c.val = ctx->aVal; c.col = ctx->aCol; c.esc = ctx->aEsc; c.cnt = ctx->aCnt; c.diag = ctx->aDiag; c.b = ctx->D[DPPEB];
    c.x = ctx->D[DPPEX]; c.r = ctx->cgR; c.p = ctx->cgP; c.z = ctx->cgZ; c.q = ctx->cgQ; c.invd = ctx->cgInvD;
    c.val32 = ctx->aVal32; c.diag32 = ctx->aDiag32; c.invd32 = ctx->aDiagInv32;
    c.fr = ctx->fR; c.fp = ctx->fP; c.fz = ctx->fZ; c.fq = ctx->fQ; c.fd = ctx->fX;
    c.lc8 = ctx->aLc8;
    c.lcEsc = ctx->aLcEsc;
    c.win = ctx->winValid ? ctx->aWin : nullptr;
    return c;
}

static int ensureMixed(Ctx* ctx) {
    if (ctx->aVal32) return MPS_OK;
    const size_t n = ctx->N;
    MPS_HIP(hipMalloc(&ctx->aVal32, sizeof(float) * (size_t)ellWidth(ctx->K) * ellRows(n)));
    MPS_HIP(hipMalloc(&ctx->aDiagInv32, sizeof(float) * n));
    MPS_HIP(hipMalloc(&ctx->aDiag32, sizeof(float) * n));
    MPS_HIP(hipMalloc(&ctx->fR, sizeof(float) * n));
    MPS_HIP(hipMalloc(&ctx->fP, sizeof(float) * n));
    MPS_HIP(hipMalloc(&ctx->fZ, sizeof(float) * n));
    MPS_HIP(hipMalloc(&ctx->fQ, sizeof(float) * n));
    MPS_HIP(hipMalloc(&ctx->fX, sizeof(double) * n));
    return MPS_OK;
}

static int ensureWindows(Ctx* ctx) {
    if (ctx->aLc8) return MPS_OK;
    MPS_HIP(hipMalloc(&ctx->aLc8, (size_t)ellWidth(ctx->K) * ellRows(ctx->N)));
    MPS_HIP(hipMalloc(&ctx->aLcEsc, sizeof(unsigned short) * (size_t)kLcEsc * ellRows(ctx->N)));
    MPS_HIP(hipMalloc(&ctx->aWin, sizeof(int) * (size_t)kWinStride * gridFor(ctx->N, kB)));
    return MPS_OK;
}

// The solver kind an assembly prepares: FP64 CG, mixed CG (FP32 values as
// well) or BiCGSTAB (A instead of A^T).
static int asmModeOf(const mps_params& P) {
    if (P.solver_type == MPS_SOLVER_BICGSTAB) return 2;
    return P.precision_mode == MPS_PREC_MIXED ? 1 : 0;
}

int ppeAssemble(Ctx* ctx) {
    // BiCGSTAB runs in FP64 on A; CG on A^T (ppe_transpose) and in the chosen precision
    const bool bicg = ctx->P.solver_type == MPS_SOLVER_BICGSTAB;
    const bool mixed = !bicg && ctx->P.precision_mode == MPS_PREC_MIXED;
    if (mixed) {
        int rc = ensureMixed(ctx);
        if (rc) return rc;
    }
    KArgs a = ctx->args();
    a.red = ctx->asmRed;   // the first residual's partials stay valid until the solve
    const int g = gridFor(ctx->N, kB);
    const bool T = ctx->P.ppe_transpose != 0 && !bicg;
    float* v32 = mixed ? ctx->aVal32 : nullptr;
    float* d32 = mixed ? ctx->aDiag32 : nullptr;
    float* i32 = mixed ? ctx->aDiagInv32 : nullptr;
    // I_NESC counts the escaped columns; the entry count is summed from the
    // row counts when asked for (reduceCounts), not by one same-address
    // atomic per wavefront here
    MPS_HIP(hipMemsetAsync(ctx->is + I_NESC, 0, sizeof(int), ctx->stream));
    MPS_HIP(hipMemsetAsync(ctx->is + I_NLCESC, 0, sizeof(int), ctx->stream));
    // the CG solve's first true residual comes out of the assembly (with the
    // learned warm start it is the residual the warm start corrects)
    const bool res = !bicg;
    double* x = ctx->D[DPPEX];
    double* invd64 = (res && !mixed) ? ctx->cgInvD : nullptr;
    const int bFirst = mixed ? 0 : 1;
    // the CG solves' LDS windows (ppe_window scalar)
    const bool wins = !bicg && ctx->ppeWindow;
    if (wins) {
        int rc = ensureWindows(ctx);
        if (rc) return rc;
    }
    unsigned char* lc8 = wins ? ctx->aLc8 : nullptr;
    unsigned short* lcEsc = wins ? ctx->aLcEsc : nullptr;
    int* win = wins ? ctx->aWin : nullptr;
#define ASM(L, TT, R)                                                                                          \
    hipLaunchKernelGGL((kPpeAssemble<L, TT, R>), dim3(g), dim3(kB), 0, ctx->stream, a, ctx->aVal, ctx->aCol,   \
                       ctx->aEsc, ctx->aCnt, ctx->aDiag, ctx->D[DPPEB], v32, d32, i32, invd64, x, ctx->cgR, bFirst, \
                       lc8, lcEsc, win)
    if (res) {
        if (ctx->listValid) {
            if (T) ASM(true, true, true); else ASM(true, false, true);
        } else {
            if (T) ASM(false, true, true); else ASM(false, false, true);
        }
    } else if (ctx->listValid) {
        if (T) ASM(true, true, false); else ASM(true, false, false);
    } else {
        if (T) ASM(false, true, false); else ASM(false, false, false);
    }
#undef ASM
    MPS_HIP(hipGetLastError());
    ctx->r0Fresh = res;
    ctx->r0FreshMixed = mixed;
    ctx->winValid = wins;
    ctx->asmMode = asmModeOf(ctx->P);
    return MPS_OK;
}

// Blocks of the last assembly whose rows the mixed SpMV serves from an LDS
// window (0 without windows).
int ppeWindowBlocks(Ctx* ctx, double* out) {
    *out = 0.0;
    if (!ctx->winValid) return MPS_OK;
    const int nb = gridFor(ctx->N, kB);
    std::vector<int> t((size_t)kWinStride * nb);
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    MPS_HIP(hipMemcpy(t.data(), ctx->aWin, sizeof(int) * t.size(), hipMemcpyDeviceToHost));
    int c = 0;
    for (int b = 0; b < nb; b++) c += (t[(size_t)b * kWinStride] == 0) ? 1 : 0;
    *out = c;
    return MPS_OK;
}

// Launch one SpMV kernel; when profiling, every kEventEvery-th launch of a
// batch is bracketed by HIP events (a sample: each marker pair adds a few
// microseconds to the stream, and the mean launch time needs no more).
constexpr int kEventEvery = 8;
template <typename K>
static int timedSpmv(Ctx* ctx, K kern, const KArgs& a, const CgPtrs& c, int nb, int slot) {
    hipStream_t s = ctx->stream;
    const bool ev = ctx->profile && slot % kEventEvery == 0;
    if (ev) MPS_HIP(hipEventRecord(ctx->evPool[2 * slot], s));
    hipLaunchKernelGGL(kern, dim3(nb), dim3(kB), 0, s, a, c);
    if (ev) MPS_HIP(hipEventRecord(ctx->evPool[2 * slot + 1], s));
    return MPS_OK;
}

// the timed launches among the first `real` of a solve pass (t < 0: not timed)
static void addSpmvStats(Ctx* ctx, const std::vector<float>& t, int real, double bytes) {
    if (!ctx->profile) return;
    real = std::min(real, (int)t.size());
    for (int q = 0; q < real; q++) {
        if (t[q] < 0.f) continue;
        ctx->spmvMsTotal += t[q];
        ctx->spmvBytesTotal += bytes;
        ctx->spmvLaunches++;
    }
}

// Iterations queued before the first convergence read: the last solve's
// count plus one (a typical solve then needs one host read), at least the
// batch; kernels queued past convergence exit at once on the device flag.
constexpr int kMaxBatch = 64;
static int firstBatch(int last, int batch) { return std::min(std::max(last + 1, batch), kMaxBatch); }

// Algorithmic bytes of one SpMV launch: the value (vb bytes) and column code
// of each entry, the per-row vectors (rb bytes per live row) and the column
// escapes.  With the LDS windows a code is one byte and an escape two (the
// window index); without, a code is two bytes and an escape four (the slot).
static double spmvBytes(const Ctx* ctx, double vb, double rb) {
    const double nnz = ctx->hostScalars[S_NNZ];
    if (ctx->winValid) return (vb + 1.0) * nnz + 2.0 * ctx->hostInts[I_NLCESC] + rb * ctx->nLiveHost;
    return (vb + 2.0) * nnz + 4.0 * ctx->hostInts[I_NESC] + rb * ctx->nLiveHost;
}

static int readStatus(Ctx* ctx, std::vector<float>* times, int batch) {
    hipStream_t s = ctx->stream;
    MPS_HIP(hipMemcpyAsync(ctx->hostInts, ctx->is, sizeof(int) * NIS, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipMemcpyAsync(ctx->hostScalars, ctx->ds, sizeof(double) * NDS, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipStreamSynchronize(s));
    if (ctx->profile && times)
        for (int q = 0; q < batch; q++) {
            float ms = -1.f;
            if (q % kEventEvery == 0) hipEventElapsedTime(&ms, ctx->evPool[2 * q], ctx->evPool[2 * q + 1]);
            times->push_back(ms);
        }
    return MPS_OK;
}

// ---------------------------------------------------------- BiCGSTAB
// Eigen's BiCGSTAB with its DiagonalPreconditioner
// (solveBiConjugateGradientStabilized, src/MpsPressure.cpp:448-468), in FP64
// on A itself: restart when rho falls below eps^2 |r0|^2 (the first restart
// also resets the iteration count), stop on |r|^2 <= tol^2 |b|^2 or the
// iteration limit.  Device scalars carry rho, alpha, w; each reduction is a
// block-partials kernel and a one-block finish.
struct BcgPtrs {
    CgPtrs c;   // x, r, p, z, q (= v), invd and the FP64 matrix
    double* r0;
    double* y;
    double* s;
    double* t;
};

// r = b - A x; r0 = r; p = v = 0; invd; partials |r|^2, |b|^2
__global__ void kBcgStart(KArgs a, BcgPtrs q) {
    const int i = slotId();
    double s0 = 0, s1 = 0;
    if (i < a.N) {
        const double ax = spmv64(a, q.c, q.c.x, i);
        const double r = q.c.b[i] - ax;
        q.c.r[i] = r;
        q.r0[i] = r;
        q.c.p[i] = 0.0;
        q.c.q[i] = 0.0;
        const double d = q.c.diag[i];
        q.c.invd[i] = (d != 0.0) ? 1.0 / d : 1.0;
        s0 = r * r;
        s1 = q.c.b[i] * q.c.b[i];
    }
    writePartials(a, s0, s1);
}
__global__ void kBcgStartFinish(KArgs a, int nb, double tol, int maxIt) {
    double rn2, bn2;
    finishPartials(a.red, nb, rn2, bn2);
    if (threadIdx.x != 0) return;
    a.ds[S_R0N2] = rn2;
    a.ds[S_RN2] = rn2;
    a.ds[S_RHSN2] = bn2;
    a.ds[S_THR] = tol * tol * bn2;
    a.ds[S_RHO] = 1.0;
    a.ds[S_ALPHA] = 1.0;
    a.ds[S_BW] = 1.0;
    a.is[I_CG_ITER] = 0;
    a.is[I_BCG_RESTART] = 0;
    a.is[I_BCG_RESTARTS] = 0;
    if (!isfinite(bn2) || !isfinite(rn2)) a.is[I_CG_DONE] = 5;
    else if (bn2 == 0.0) a.is[I_CG_DONE] = 2;   // x = 0
    else a.is[I_CG_DONE] = (rn2 > a.ds[S_THR] && maxIt > 0) ? 0 : 1;
}
// rho = r0.r
__global__ void kBcgRho(KArgs a, BcgPtrs q) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    writePartials(a, i < a.N ? q.r0[i] * q.c.r[i] : 0.0, 0.0);
}
__global__ void kBcgRhoFinish(KArgs a, int nb) {
    if (a.is[I_CG_DONE]) return;
    double rho, unused;
    finishPartials(a.red, nb, rho, unused);
    if (threadIdx.x != 0) return;
    a.ds[S_RHOOLD] = a.ds[S_RHO];
    a.ds[S_RHO] = rho;
    a.is[I_BCG_RESTART] = fabs(rho) < DBL_EPSILON * DBL_EPSILON * a.ds[S_R0N2] ? 1 : 0;
}
// restart: r = b - A x; r0 = r; rho = |r0|^2 = |r|^2
__global__ void kBcgRestart(KArgs a, BcgPtrs q) {
    if (a.is[I_CG_DONE] || !a.is[I_BCG_RESTART]) return;
    const int i = slotId();
    double s0 = 0;
    if (i < a.N) {
        const double r = q.c.b[i] - spmv64(a, q.c, q.c.x, i);
        q.c.r[i] = r;
        q.r0[i] = r;
        s0 = r * r;
    }
    writePartials(a, s0, 0.0);
}
__global__ void kBcgRestartFinish(KArgs a, int nb) {
    if (a.is[I_CG_DONE] || !a.is[I_BCG_RESTART]) return;
    double rn2, unused;
    finishPartials(a.red, nb, rn2, unused);
    if (threadIdx.x != 0) return;
    a.ds[S_RHO] = rn2;
    a.ds[S_R0N2] = rn2;
    if (a.is[I_BCG_RESTARTS]++ == 0) a.is[I_CG_ITER] = 0;
    a.is[I_BCG_RESTART] = 0;
}
// p = r + beta (p - w v); y = invd p
__global__ void kBcgP(KArgs a, BcgPtrs q) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    if (i >= a.N) return;
    const double w = a.ds[S_BW];
    const double beta = (a.ds[S_RHO] / a.ds[S_RHOOLD]) * (a.ds[S_ALPHA] / w);
    const double p = q.c.r[i] + beta * (q.c.p[i] - w * q.c.q[i]);
    q.c.p[i] = p;
    q.y[i] = q.c.invd[i] * p;
}
// v = A y; partial r0.v
__global__ __launch_bounds__(256) void kBcgV(KArgs a, BcgPtrs q) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    double s0 = 0;
    if (i < a.N) {
        const double v = spmv64(a, q.c, q.y, i);
        q.c.q[i] = v;
        s0 = q.r0[i] * v;
    }
    writePartials(a, s0, 0.0);
}
__global__ void kBcgAlpha(KArgs a, int nb) {
    if (a.is[I_CG_DONE]) return;
    double r0v, unused;
    finishPartials(a.red, nb, r0v, unused);
    if (threadIdx.x == 0) a.ds[S_ALPHA] = a.ds[S_RHO] / r0v;
}
// s = r - alpha v; z = invd s
__global__ void kBcgS(KArgs a, BcgPtrs q) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    if (i >= a.N) return;
    const double sv = q.c.r[i] - a.ds[S_ALPHA] * q.c.q[i];
    q.s[i] = sv;
    q.c.z[i] = q.c.invd[i] * sv;
}
// t = A z; partials t.t, t.s
__global__ __launch_bounds__(256) void kBcgT(KArgs a, BcgPtrs q) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    double s0 = 0, s1 = 0;
    if (i < a.N) {
        const double t = spmv64(a, q.c, q.c.z, i);
        q.t[i] = t;
        s0 = t * t;
        s1 = t * q.s[i];
    }
    writePartials(a, s0, s1);
}
__global__ void kBcgW(KArgs a, int nb) {
    if (a.is[I_CG_DONE]) return;
    double tt, ts;
    finishPartials(a.red, nb, tt, ts);
    if (threadIdx.x == 0) a.ds[S_BW] = (tt > 0.0) ? ts / tt : 0.0;
}
// x += alpha y + w z; r = s - w t; partial |r|^2
__global__ void kBcgUpdate(KArgs a, BcgPtrs q) {
    if (a.is[I_CG_DONE]) return;
    const int i = slotId();
    double s0 = 0;
    if (i < a.N) {
        const double alpha = a.ds[S_ALPHA], w = a.ds[S_BW];
        q.c.x[i] = q.c.x[i] + (alpha * q.y[i] + w * q.c.z[i]);
        const double r = q.s[i] - w * q.t[i];
        q.c.r[i] = r;
        s0 = r * r;
    }
    writePartials(a, s0, 0.0);
}
__global__ void kBcgCheck(KArgs a, int nb, int maxIt) {
    if (a.is[I_CG_DONE]) return;
    double rn2, unused;
    finishPartials(a.red, nb, rn2, unused);
    if (threadIdx.x != 0) return;
    a.ds[S_RN2] = rn2;
    const int it = a.is[I_CG_ITER] + 1;
    a.is[I_CG_ITER] = it;
    if (!isfinite(rn2)) a.is[I_CG_DONE] = 5;
    else if (!(rn2 > a.ds[S_THR])) a.is[I_CG_DONE] = 1;
    else if (it >= maxIt) a.is[I_CG_DONE] = 3;
}

static int ensureBicgstab(Ctx* ctx) {
    if (ctx->bcgR0) return MPS_OK;
    const size_t bytes = sizeof(double) * (size_t)ctx->N;
    MPS_HIP(hipMalloc(&ctx->bcgR0, bytes));
    MPS_HIP(hipMalloc(&ctx->bcgY, bytes));
    MPS_HIP(hipMalloc(&ctx->bcgS, bytes));
    MPS_HIP(hipMalloc(&ctx->bcgT, bytes));
    return MPS_OK;
}

static int solveBicgstab(Ctx* ctx, const KArgs& a, const CgPtrs& c, int nb, int batch) {
    int rc = ensureBicgstab(ctx);
    if (rc) return rc;
    hipStream_t s = ctx->stream;
    const BcgPtrs q{c, ctx->bcgR0, ctx->bcgY, ctx->bcgS, ctx->bcgT};
    const double tol = ctx->P.solver_tol > 0 ? ctx->P.solver_tol : DBL_EPSILON;
    const int maxIt = ctx->P.solver_max_iter > 0 ? ctx->P.solver_max_iter : 2 * ctx->nLiveHost;
    hipLaunchKernelGGL(kBcgStart, dim3(nb), dim3(kB), 0, s, a, q);
    hipLaunchKernelGGL(kBcgStartFinish, dim3(1), dim3(1024), 0, s, a, nb, tol, maxIt);
    hipLaunchKernelGGL(kCgZeroX, dim3(nb), dim3(kB), 0, s, a, c.x);
    int done = 0, iters = 0;
    // a restart resets the count once, so the host bound allows two passes
    for (int launched = 0; !done; launched += batch) {
        for (int k = 0; k < batch; k++) {
            hipLaunchKernelGGL(kBcgRho, dim3(nb), dim3(kB), 0, s, a, q);
            hipLaunchKernelGGL(kBcgRhoFinish, dim3(1), dim3(1024), 0, s, a, nb);
            hipLaunchKernelGGL(kBcgRestart, dim3(nb), dim3(kB), 0, s, a, q);
            hipLaunchKernelGGL(kBcgRestartFinish, dim3(1), dim3(1024), 0, s, a, nb);
            hipLaunchKernelGGL(kBcgP, dim3(nb), dim3(kB), 0, s, a, q);
            hipLaunchKernelGGL(kBcgV, dim3(nb), dim3(kB), 0, s, a, q);
            hipLaunchKernelGGL(kBcgAlpha, dim3(1), dim3(1024), 0, s, a, nb);
            hipLaunchKernelGGL(kBcgS, dim3(nb), dim3(kB), 0, s, a, q);
            hipLaunchKernelGGL(kBcgT, dim3(nb), dim3(kB), 0, s, a, q);
            hipLaunchKernelGGL(kBcgW, dim3(1), dim3(1024), 0, s, a, nb);
            hipLaunchKernelGGL(kBcgUpdate, dim3(nb), dim3(kB), 0, s, a, q);
            hipLaunchKernelGGL(kBcgCheck, dim3(1), dim3(1024), 0, s, a, nb, maxIt);
        }
        MPS_HIP(hipGetLastError());
        if ((rc = readStatus(ctx, nullptr, 0))) return rc;
        done = ctx->hostInts[I_CG_DONE];
        iters = ctx->hostInts[I_CG_ITER];
        if (launched > 8 * maxIt + 64) break;   // the device flag ends every batch; this only bounds the host loop
    }
    ctx->solverIter = iters;
    ctx->outerIter = 1;
    ctx->lastDone = ctx->hostInts[I_CG_DONE];
    ctx->solverErr = ctx->hostScalars[S_RHSN2] > 0 ? sqrt(ctx->hostScalars[S_RN2] / ctx->hostScalars[S_RHSN2]) : 0.0;
    return MPS_OK;
}

// Eigen-semantics Jacobi-PCG in FP64 (solveConjugateGradient, src/MpsPressure.cpp:424-445)
static int solveFp64(Ctx* ctx, const KArgs& a, const CgPtrs& c, int nb, int batch) {
    hipStream_t s = ctx->stream;
    const double tol = ctx->P.solver_tol > 0 ? ctx->P.solver_tol : DBL_EPSILON;
    const int maxIt = ctx->P.solver_max_iter > 0 ? ctx->P.solver_max_iter : 2 * ctx->nLiveHost;
    // r = b - A x, 1 / diag and the partials may come from the assembly
    const bool fresh = ctx->r0Fresh && !ctx->r0FreshMixed;
    if (!fresh) hipLaunchKernelGGL(kCgStart, dim3(nb), dim3(kB), 0, s, a, c);
    ctx->r0Fresh = false;
    KArgs a0 = a;
    if (fresh) a0.red = ctx->asmRed;
    hipLaunchKernelGGL(kCgStartFinish, dim3(1), dim3(1024), 0, s, a0, nb, tol);
    hipLaunchKernelGGL(kCgZeroX, dim3(nb), dim3(kB), 0, s, a, c.x);
    hipLaunchKernelGGL(kCgInitP, dim3(nb), dim3(kB), 0, s, a, c);
    hipLaunchKernelGGL(kCgInitPFinish, dim3(1), dim3(1024), 0, s, a, nb);
    // when profiling, the entry count (for the SpMV bytes) is summed on the
    // device and comes with the loop's status reads
    if (ctx->profile) {
        const int rcN = launchCountToScalar(ctx, ctx->aCnt, ctx->N, S_NNZ);
        if (rcN) return rcN;
    }
    std::vector<float> times;
    int done = 0, iters = 0;
    for (int launched = 0, b = firstBatch(ctx->lastCgIters, batch); !done; launched += b, b = batch) {
        for (int q = 0; q < b; q++) {
            timedSpmv(ctx, kCgSpmv, a, c, nb, q);
            hipLaunchKernelGGL(kCgAlpha, dim3(1), dim3(1024), 0, s, a, nb);
            hipLaunchKernelGGL(kCgUpdate, dim3(nb), dim3(kB), 0, s, a, c);
            hipLaunchKernelGGL(kCgBeta, dim3(1), dim3(1024), 0, s, a, nb, maxIt);
            hipLaunchKernelGGL(kCgP, dim3(nb), dim3(kB), 0, s, a, c);
        }
        int rc = readStatus(ctx, &times, b);
        if (rc) return rc;
        done = ctx->hostInts[I_CG_DONE];
        iters = ctx->hostInts[I_CG_ITER];
        if (launched > 4 * maxIt + 64) break;   // the device flag ends every batch; this only bounds the host loop
    }
    ctx->lastCgIters = iters;
    const int D = ctx->hostInts[I_CG_DONE];
    ctx->lastDone = D;
    if (ctx->profile)   // the entry count came with the loop's status reads
        addSpmvStats(ctx, times, (D == 1 || D == 3) ? iters + 1 : 0,
                     spmvBytes(ctx, 8.0, 28.0));
    ctx->solverIter = iters;
    ctx->outerIter = 1;
    ctx->solverErr = ctx->hostScalars[S_RHSN2] > 0 ? sqrt(ctx->hostScalars[S_RN2] / ctx->hostScalars[S_RHSN2]) : 0.0;
    return MPS_OK;
}

// Mixed precision: FP64 iterative refinement on the true residual around an
// FP32 Jacobi-PCG (FP64 dot products).  Stops when |b - A x| < tol |b| in FP64.
static int solveMixed(Ctx* ctx, const KArgs& a, const CgPtrs& c, int nb, int batch) {
    constexpr int kMaxOuter = 30;
    // A refinement pass with the learned preconditioner ends after this many
    // iterations: an operator that neither converges nor breaks down (a
    // poorly trained or random one can drift for thousands of iterations)
    // then faces the stall test below instead of using up the whole
    // iteration budget.  A working operator needs tens of iterations per pass.
    constexpr int kLearnedPassMax = 500;
    hipStream_t s = ctx->stream;
    bool useM = ctx->P.learned_mode == 2 && ctx->gnn != nullptr;
    // The flexible CG's iterate is not the A-norm minimiser the Jacobi PCG's
    // is: at the same residual its pressure error is larger (up to 1e-5 of
    // the largest pressure at 1e-8, against 1e-7), so with the learned
    // preconditioner the true residual is driven to learned_tol_scale times
    // the tolerance.
    const double tol = (ctx->P.solver_tol > 0 ? ctx->P.solver_tol : 1e-8) * (useM ? ctx->learnedTolScale : 1.0);
    const double tolIn = ctx->P.mixed_inner_tol > 0 ? ctx->P.mixed_inner_tol : 1e-4;
    const int maxIt = ctx->P.solver_max_iter > 0 ? ctx->P.solver_max_iter : 2 * ctx->nLiveHost;
    std::vector<float> times;
    int total = 0, outer = 0, rc;
    double prevRn2 = 0.0;
    ctx->lastDone = 0;
    // when profiling, the entry count (for the SpMV bytes) is summed on the
    // device and comes with the next status read
    if (ctx->profile && (rc = launchCountToScalar(ctx, ctx->aCnt, ctx->N, S_NNZ))) return rc;
    auto bytes32 = [ctx]() { return spmvBytes(ctx, 4.0, 16.0); };
    for (;;) {
        // the first residual may come from the assembly (r and its partials)
        const bool fresh = outer == 0 && ctx->r0Fresh && ctx->r0FreshMixed;
        if (!fresh) hipLaunchKernelGGL(kTrueResidual, dim3(nb), dim3(kB), 0, s, a, c);
        ctx->r0Fresh = false;
        KArgs a0 = a;
        if (fresh) a0.red = ctx->asmRed;
        hipLaunchKernelGGL(kTrueResidualFinish, dim3(1), dim3(1024), 0, s, a0, nb);
        if ((rc = readStatus(ctx, nullptr, 0))) return rc;
        const double rn2 = ctx->hostScalars[S_TRUE_RN2], bn2 = ctx->hostScalars[S_BN2];
        if (!std::isfinite(rn2) || !std::isfinite(bn2)) break;   // reported by ppeSolve
        if (bn2 == 0.0) {   // zero right-hand side: x = 0, as the FP64 path
            hipLaunchKernelGGL(kZero, dim3(nb), dim3(kB), 0, s, a, c.x);
            break;
        }
        if (rn2 < tol * tol * bn2 || outer >= kMaxOuter || total >= maxIt) break;
        // the learned operator stays in use while each refinement pass cuts
        // the true residual to learned_stall_ratio (default: halves it);
        // otherwise Jacobi takes over
        const double stall = ctx->learnedStallRatio;
        if (useM && prevRn2 > 0.0 && rn2 > stall * stall * prevRn2) {
            useM = false;
            ctx->fallbackUsed = 1;
        }
        prevRn2 = rn2;
        outer++;
        if (useM) {
            hipLaunchKernelGGL(kFcgInit, dim3(nb), dim3(kB), 0, s, a, c);
            gnnApply32(ctx, c.fr, c.fz);
            hipLaunchKernelGGL(kFcgInitDots, dim3(nb), dim3(kB), 0, s, a, c);
        } else {
            hipLaunchKernelGGL(kMxInit, dim3(nb), dim3(kB), 0, s, a, c);
        }
        hipLaunchKernelGGL(kMxInitFinish, dim3(1), dim3(1024), 0, s, a, nb, tolIn, tol);
        int done = 0, it = 0;
        times.clear();
        // first batch: the last solve's count for this pass plus one (a
        // second pass, which starts just above the target, is short)
        const int bq0 = outer == 1 ? firstBatch(ctx->lastInnerIters, batch)
                                   : (outer == 2 ? std::min(std::max(ctx->lastPass2Iters + 1, 2), batch) : batch);
        for (int launched = 0, bq = bq0; !done; launched += bq, bq = batch) {
            for (int q = 0; q < bq; q++) {
                timedSpmv(ctx, kMxSpmv, a, c, nb, q);
                hipLaunchKernelGGL(kCgAlpha, dim3(1), dim3(1024), 0, s, a, nb);
                if (useM) {
                    hipLaunchKernelGGL(kFcgUpdate, dim3(nb), dim3(kB), 0, s, a, c);
                    gnnApply32(ctx, c.fr, c.fz);
                    hipLaunchKernelGGL(kFcgDots, dim3(nb), dim3(kB), 0, s, a, c);
                    hipLaunchKernelGGL(kFcgBeta, dim3(1), dim3(1024), 0, s, a, nb, std::min(maxIt, kLearnedPassMax));
                } else {
                    hipLaunchKernelGGL(kMxUpdate, dim3(nb), dim3(kB), 0, s, a, c);
                    hipLaunchKernelGGL(kCgBeta, dim3(1), dim3(1024), 0, s, a, nb, maxIt);
                }
                hipLaunchKernelGGL(kMxP, dim3(nb), dim3(kB), 0, s, a, c);
            }
            if ((rc = readStatus(ctx, &times, bq))) return rc;
            done = ctx->hostInts[I_CG_DONE];
            it = ctx->hostInts[I_CG_ITER];
            if (launched > 4 * maxIt + 64) break;
        }
        if (outer == 1) ctx->lastInnerIters = it;
        if (outer == 2) ctx->lastPass2Iters = it;
        const int D = ctx->hostInts[I_CG_DONE];
        addSpmvStats(ctx, times, (D == 1 || D == 3 || D == 4) ? it + 1 : 0, bytes32());
        total += it + 1;
        if (D == 4) {   // breakdown of the flexible CG: keep its progress, continue with Jacobi
            useM = false;
            ctx->fallbackUsed = 1;
        }
        hipLaunchKernelGGL(kMxApply, dim3(nb), dim3(kB), 0, s, a, c);
        if (D == 5) {   // non-finite inner residual or product
            ctx->lastDone = 5;
            break;
        }
    }
    ctx->solverIter = total;
    ctx->outerIter = outer;
    if (ctx->lastDone != 5) ctx->lastDone = ctx->hostInts[I_CG_DONE];
    ctx->solverErr = ctx->hostScalars[S_BN2] > 0 ? sqrt(ctx->hostScalars[S_TRUE_RN2] / ctx->hostScalars[S_BN2]) : 0.0;
    return MPS_OK;
}

// Initial guess of the next solve from the history of solutions (x_n in
// ppe_x, x_{n-1}, x_{n-2}): x_n (order 0), 2 x_n - x_{n-1} (1) or
// 3 x_n - 3 x_{n-1} + x_{n-2} (2); the history shifts by one.  The
// reference starts from x_n (solveWithGuess, src/MpsPressure.cpp:433); the
// solve still runs to the same residual target.
__global__ void kPpeExtrapolate(KArgs a, int order) {
    const int i = slotId();
    if (i >= a.N) return;
    const double x = a.D[DPPEX][i], x1 = a.D[DPPEX1][i], x2 = a.D[DPPEX2][i], x3 = a.D[DPPEX3][i];
    double g = x;
    if (order == 1) g = 2.0 * x - x1;
    else if (order == 2) g = 3.0 * x - 3.0 * x1 + x2;
    else if (order == 3) g = 4.0 * x - 6.0 * x1 + 4.0 * x2 - x3;
    a.D[DPPEX3][i] = x2;
    a.D[DPPEX2][i] = x1;
    a.D[DPPEX1][i] = x;
    a.D[DPPEX][i] = g;
}

int ppeWarmStart(Ctx* ctx) {
    const int want = ctx->P.ppe_warm_start;
    if (want <= 0) return MPS_OK;
    ctx->r0Fresh = false;   // the guess changes after any assembly
    // solutions available before this solve: ppeSolves (the first guess is not one)
    const int order = std::min(want, std::max(ctx->ppeSolves - 1, 0));
    KArgs a = ctx->args();
    hipLaunchKernelGGL(kPpeExtrapolate, dim3(gridFor(ctx->N, kB)), dim3(kB), 0, ctx->stream, a, std::min(order, 3));
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

// Time of the last pressure solve from its events (waits for them).
int ppeSolveMs(Ctx* ctx, double* out) {
    if (ctx->ppeMsPending) {
        MPS_HIP(hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        MPS_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->ppeMs = ms;
        ctx->ppeMsPending = false;
    }
    *out = ctx->ppeMs;
    return MPS_OK;
}

int ppeSolve(Ctx* ctx) {
    const bool bicg = ctx->P.solver_type == MPS_SOLVER_BICGSTAB;
    const bool mixed = !bicg && ctx->P.precision_mode == MPS_PREC_MIXED;
    if (ctx->asmMode < 0) {
        ctx->err = "the pressure solve needs a PPE assembly after the last bucket update, compaction or state upload";
        return MPS_E_ARG;
    }
    if (ctx->asmMode != asmModeOf(ctx->P)) {
        static const char* kinds[] = {"FP64 CG", "mixed-precision CG", "BiCGSTAB"};
        ctx->err = std::string("the PPE was assembled for ") + kinds[ctx->asmMode] + " but the solve is " +
                   kinds[asmModeOf(ctx->P)] + " (solver_type / precision_mode changed in between)";
        return MPS_E_ARG;
    }
    KArgs a = ctx->args();
    CgPtrs c = cgPtrs(ctx);
    const int N = ctx->N;
    const int nb = gridFor(N, kB);
    const int batch = 8;
    hipStream_t s = ctx->stream;
    int rc0;
    if (ctx->profile && ctx->evPool.size() < 2 * kMaxBatch) {
        ctx->evPool.resize(2 * kMaxBatch);
        for (auto& e : ctx->evPool) MPS_HIP(hipEventCreate(&e));
    }
    if (ctx->ppeMsPending) {   // keep the previous solve's time before its events are reused
        double prev;
        if ((rc0 = ppeSolveMs(ctx, &prev))) return rc0;
    }
    MPS_HIP(hipEventRecord(ctx->ev0, s));
    ctx->fallbackUsed = 0;
    int rc;
    if (ctx->P.learned_mode >= 1) {
        // learned operator for this matrix, then x += M (b - A x)
        if ((rc = gnnForward(ctx))) return rc;
        if ((rc = gnnWarmStart(ctx))) return rc;
    }
    if (bicg) rc = solveBicgstab(ctx, a, c, nb, batch);
    else rc = mixed ? solveMixed(ctx, a, c, nb, batch) : solveFp64(ctx, a, c, nb, batch);
    if (rc) return rc;
    if (!mixed) {   // the mixed path ends on a fresh true residual already
        hipLaunchKernelGGL(kTrueResidual, dim3(nb), dim3(kB), 0, s, a, c);
        hipLaunchKernelGGL(kTrueResidualFinish, dim3(1), dim3(1024), 0, s, a, nb);
    }
    hipLaunchKernelGGL(kPressFromX, dim3(nb), dim3(kB), 0, s, a, c.x);
    MPS_HIP(hipEventRecord(ctx->ev1, s));
    // the mixed path's last status read already holds the final true
    // residual; the other paths read the one computed above.  The solve's
    // time is read from its events when asked for (ppeSolveMs).
    if (!mixed && (rc = readStatus(ctx, nullptr, 0))) return rc;
    MPS_HIP(hipGetLastError());
    ctx->ppeMsPending = true;
    const double* hs = ctx->hostScalars;
    ctx->trueResidual = hs[S_BN2] > 0 ? sqrt(hs[S_TRUE_RN2] / hs[S_BN2]) : 0.0;
    if (!std::isfinite(hs[S_BN2]) || !std::isfinite(hs[S_TRUE_RN2])) ctx->trueResidual = NAN;
    ctx->ppeSolves++;
    ctx->r0Fresh = false;
    // Outcome.  FP64 CG and BiCGSTAB follow Eigen's rule (success when the
    // recursive relative residual ends at or below the tolerance; the
    // reference prints "Failed solving" otherwise, src/MpsPressure.cpp:440-442);
    // the mixed solve stops on its FP64 true residual.  The pressure is
    // written either way, as the reference does.
    const double tol = ctx->P.solver_tol > 0 ? ctx->P.solver_tol : (mixed ? 1e-8 : DBL_EPSILON);
    const bool nonFinite = ctx->lastDone == 5 || !std::isfinite(ctx->trueResidual) || !std::isfinite(ctx->solverErr);
    const bool converged = mixed ? ctx->trueResidual < tol : ctx->solverErr <= tol;
    ctx->solverFailed = nonFinite ? 2 : converged ? 0 : 1;
    if (ctx->solverFailed) {
        ctx->solverFailures++;
        char msg[256];
        snprintf(msg, sizeof msg,
                 "%s: the pressure solve %s after %d iterations (relative residual %.3e, true %.3e, tolerance %.1e)",
                 "Failed solving", nonFinite ? "produced a non-finite residual" : "stopped above the tolerance",
                 ctx->solverIter, ctx->solverErr, ctx->trueResidual, tol);
        ctx->err = msg;
        return MPS_E_SOLVER;
    }
    return MPS_OK;
}

}  // namespace mpsgpu
