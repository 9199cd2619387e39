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

#include "kern_common.h"

namespace mpsgpu {

constexpr int kB = 256;

template <bool LIST, bool TRANS>
__global__ void kPpeAssemble(KArgs a, double* val, int* col, int* cnt, double* diag, double* b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    if (!live(a, i)) { cnt[i] = 0; diag[i] = 1.0; b[i] = 0.0; return; }
    const mps_params& P = a.P;
    const int bci = a.I[IBC][i];
    if (bci == P.other || bci == P.surface) {
        cnt[i] = 0; diag[i] = 1.0; b[i] = 0.0;
        return;
    }
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = a.D[DX][i], yi = a.D[DY][i], zi = a.D[DZ][i];
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    const double rho = P.dns_fl1;
    double sum = 0.0, rhs = 0.0;
    int c = 0;
    auto body = [&](int j) {
        const double xj = a.D[DX][j], yj = a.D[DY][j], zj = a.D[DZ][j];
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        if (!(d2 < P.re_l2)) return;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        const double wL = wgt(sqrt(d2), P.re_l, P.weight_type);
        const double mij = wL * P.coeff_ppe;
        const bool okRow = d2 < dm2 || pw;   // j is a neighbour of row i
        if (okRow) {
            if (a.I[ITYPE][j] == P.dummy_wall) {
                const double pgh = rho * (v0 * P.gravity[0] + v1 * P.gravity[1] + v2 * P.gravity[2]) * wL;
                rhs -= P.coeff_ppe * pgh;
            } else {
                sum -= mij;
                if (!TRANS && a.I[IBC][j] == P.inner) {
                    val[(size_t)c * a.N + i] = mij;
                    col[(size_t)c * a.N + i] = j;
                    c++;
                }
            }
        }
        if (TRANS && a.I[IBC][j] == P.inner) {
            // entry (j, i) of A: i seen from row j, mirror test with j's mirror
            const double u0 = xi - a.D[DMX][j], u1 = yi - a.D[DMY][j], u2 = zi - a.D[DMZ][j];
            const double dmj = u0 * u0 + u1 * u1 + u2 * u2;
            if (d2 < dmj || pw) {
                val[(size_t)c * a.N + i] = mij;
                col[(size_t)c * a.N + i] = j;
                c++;
            }
        }
    };
    if (LIST) listNbrs(a, i, body);
    else walkCells(a, xi, yi, zi, [&](int j) { if (j != i) body(j); });
    sum -= P.alpha_compressibility * rho / (P.time_step * P.time_step);
    if (P.mps_type == MPS_IMPLICIT_PND_DIVU)
        rhs += -P.coeff_ppe_source * rho * (a.D[DPNDKI][i] - P.pnd_small_zero) +
               (1.0 - P.relax_pnd) * rho * a.D[DDIVU][i] / P.time_step;
    else
        rhs += -P.coeff_ppe_source * rho * (a.D[DPNDSMALL][i] - P.pnd_small_zero);
    cnt[i] = c;
    diag[i] = sum;
    b[i] = rhs;
    atomicAdd(&a.is[I_NNZ], c);
}

// ------------------------------------------------------------ reductions
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

// Sum the per-block partial pairs in a fixed order (one block of 1024 threads).
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

__device__ __forceinline__ double spmvRow(const KArgs& a, const double* val, const int* col, const int* cnt,
                                          const double* diag, const double* x, int i) {
    const int n = cnt[i];
    double acc = diag[i] * x[i];
    for (int c = 0; c < n; c++) {
        const size_t o = (size_t)c * a.N + i;
        acc += val[o] * x[col[o]];
    }
    return acc;
}

struct CgPtrs {
    const double* val;
    const int* col;
    const int* cnt;
    const double* diag;
    const double* b;
    double* x;
    double* r;
    double* p;
    double* z;
    double* q;
    double* invd;
};

// r = b - M x; partials |b|^2, |r|^2; invd = 1/diag (1 where diag == 0)
__global__ void kCgStart(KArgs a, CgPtrs c) {
    double s0 = 0, s1 = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.N; i += gridDim.x * blockDim.x) {
        if (!live(a, i)) c.x[i] = 0.0;   // slots beyond the live count do not take part
        const double d = c.diag[i];
        c.invd[i] = (d != 0.0) ? 1.0 / d : 1.0;
        const double r = c.b[i] - spmvRow(a, c.val, c.col, c.cnt, c.diag, c.x, i);
        c.r[i] = r;
        s0 += c.b[i] * c.b[i];
        s1 += r * r;
    }
    blockReduce2(s0, s1);
    if (threadIdx.x == 0) { a.red[blockIdx.x * 2] = s0; a.red[blockIdx.x * 2 + 1] = s1; }
}

__global__ void kCgStartFinish(KArgs a, int nb, double tol) {
    double bn2, rn2;
    finishPartials(a.red, nb, bn2, rn2);
    if (threadIdx.x != 0) return;
    a.ds[S_RHSN2] = bn2;
    a.ds[S_RN2] = rn2;
    a.is[I_CG_ITER] = 0;
    a.is[I_CG_DONE] = 0;
    if (bn2 == 0.0) {
        a.is[I_CG_DONE] = 2;   // x = 0
        return;
    }
    const double thr = fmax(tol * tol * bn2, DBL_MIN);
    a.ds[S_THR] = thr;
    if (rn2 < thr) a.is[I_CG_DONE] = 1;
}

__global__ void kCgZeroX(KArgs a, double* x) {
    if (a.is[I_CG_DONE] != 2) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.N; i += gridDim.x * blockDim.x) x[i] = 0.0;
}

// p = invd r; partial r.p
__global__ void kCgInitP(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    double s0 = 0, s1 = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.N; i += gridDim.x * blockDim.x) {
        const double p = c.invd[i] * c.r[i];
        c.p[i] = p;
        s0 += c.r[i] * p;
    }
    blockReduce2(s0, s1);
    if (threadIdx.x == 0) { a.red[blockIdx.x * 2] = s0; a.red[blockIdx.x * 2 + 1] = s1; }
}
__global__ void kCgInitPFinish(KArgs a, int nb) {
    if (a.is[I_CG_DONE]) return;
    double s0, s1;
    finishPartials(a.red, nb, s0, s1);
    if (threadIdx.x == 0) a.ds[S_ABSNEW] = s0;
}

// q = M p; partial p.q
__global__ void kCgSpmv(KArgs a, CgPtrs c) {
    if (a.is[I_CG_DONE]) return;
    double s0 = 0, s1 = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.N; i += gridDim.x * blockDim.x) {
        const double q = spmvRow(a, c.val, c.col, c.cnt, c.diag, c.p, i);
        c.q[i] = q;
        s0 += c.p[i] * q;
    }
    blockReduce2(s0, s1);
    if (threadIdx.x == 0) { a.red[blockIdx.x * 2] = s0; a.red[blockIdx.x * 2 + 1] = s1; }
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
    const double alpha = a.ds[S_ALPHA];
    double s0 = 0, s1 = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.N; i += gridDim.x * blockDim.x) {
        c.x[i] += alpha * c.p[i];
        const double r = c.r[i] - alpha * c.q[i];
        c.r[i] = r;
        const double z = c.invd[i] * r;
        c.z[i] = z;
        s0 += r * r;
        s1 += r * z;
    }
    blockReduce2(s0, s1);
    if (threadIdx.x == 0) { a.red[blockIdx.x * 2] = s0; a.red[blockIdx.x * 2 + 1] = s1; }
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
    const double beta = a.ds[S_BETA];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.N; i += gridDim.x * blockDim.x)
        c.p[i] = c.z[i] + beta * c.p[i];
}

// true residual |b - M x|^2 and |b|^2 partials
__global__ void kTrueResidual(KArgs a, CgPtrs c) {
    double s0 = 0, s1 = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.N; i += gridDim.x * blockDim.x) {
        const double r = c.b[i] - spmvRow(a, c.val, c.col, c.cnt, c.diag, c.x, i);
        s0 += r * r;
        s1 += c.b[i] * c.b[i];
    }
    blockReduce2(s0, s1);
    if (threadIdx.x == 0) { a.red[blockIdx.x * 2] = s0; a.red[blockIdx.x * 2 + 1] = s1; }
}
__global__ void kTrueResidualFinish(KArgs a, int nb) {
    double s0, s1;
    finishPartials(a.red, nb, s0, s1);
    if (threadIdx.x == 0) {
        a.ds[S_TRUE_RN2] = s0;
        a.ds[S_BN2] = s1;
    }
}

// press = max(x, 0) for live slots (setZeroOnNegativePressure :471-476)
__global__ void kPressFromX(KArgs a, const double* x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (!live(a, i)) return;
    const double v = x[i];
    a.D[DPRESS][i] = (v < 0) ? 0.0 : v;
}

static CgPtrs cgPtrs(Ctx* ctx) {
    CgPtrs c;
    c.val = ctx->aVal; c.col = ctx->aCol; c.cnt = ctx->aCnt; c.diag = ctx->aDiag; c.b = ctx->D[DPPEB];
    c.x = ctx->D[DPPEX]; c.r = ctx->cgR; c.p = ctx->cgP; c.z = ctx->cgZ; c.q = ctx->cgQ; c.invd = ctx->cgInvD;
    return c;
}

int ppeAssemble(Ctx* ctx) {
    KArgs a = ctx->args();
    const int g = gridFor(ctx->N, kB);
    const bool T = ctx->P.ppe_transpose != 0;
    MPS_HIP(hipMemsetAsync(ctx->is + I_NNZ, 0, sizeof(int), ctx->stream));
#define ASM(L, TT) hipLaunchKernelGGL((kPpeAssemble<L, TT>), dim3(g), dim3(kB), 0, ctx->stream, a, ctx->aVal, ctx->aCol, ctx->aCnt, ctx->aDiag, ctx->D[DPPEB])
    if (ctx->listValid) {
        if (T) ASM(true, true); else ASM(true, false);
    } else {
        if (T) ASM(false, true); else ASM(false, false);
    }
#undef ASM
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

int ppeSolve(Ctx* ctx) {
    KArgs a = ctx->args();
    CgPtrs c = cgPtrs(ctx);
    const int N = ctx->N;
    const int nb = std::min(kRedBlocks, gridFor(N, kB));
    const double tol = ctx->P.solver_tol > 0 ? ctx->P.solver_tol : DBL_EPSILON;
    const int maxIt = ctx->P.solver_max_iter > 0 ? ctx->P.solver_max_iter : 2 * ctx->nLiveHost;
    hipStream_t s = ctx->stream;
    MPS_HIP(hipEventRecord(ctx->ev0, s));
    hipLaunchKernelGGL(kCgStart, dim3(nb), dim3(kB), 0, s, a, c);
    hipLaunchKernelGGL(kCgStartFinish, dim3(1), dim3(1024), 0, s, a, nb, tol);
    hipLaunchKernelGGL(kCgZeroX, dim3(nb), dim3(kB), 0, s, a, c.x);
    hipLaunchKernelGGL(kCgInitP, dim3(nb), dim3(kB), 0, s, a, c);
    hipLaunchKernelGGL(kCgInitPFinish, dim3(1), dim3(1024), 0, s, a, nb);
    const int batch = 8;
    int done = 0, iters = 0;
    if (ctx->profile && ctx->evPool.size() < 2 * batch) {
        ctx->evPool.resize(2 * batch);
        for (auto& e : ctx->evPool) MPS_HIP(hipEventCreate(&e));
    }
    std::vector<float> spmvTimes;
    for (int launched = 0; !done; launched += batch) {
        for (int q = 0; q < batch; q++) {
            if (ctx->profile) MPS_HIP(hipEventRecord(ctx->evPool[2 * q], s));
            hipLaunchKernelGGL(kCgSpmv, dim3(nb), dim3(kB), 0, s, a, c);
            if (ctx->profile) MPS_HIP(hipEventRecord(ctx->evPool[2 * q + 1], s));
            hipLaunchKernelGGL(kCgAlpha, dim3(1), dim3(1024), 0, s, a, nb);
            hipLaunchKernelGGL(kCgUpdate, dim3(nb), dim3(kB), 0, s, a, c);
            hipLaunchKernelGGL(kCgBeta, dim3(1), dim3(1024), 0, s, a, nb, maxIt);
            hipLaunchKernelGGL(kCgP, dim3(nb), dim3(kB), 0, s, a, c);
        }
        MPS_HIP(hipMemcpyAsync(ctx->hostInts, ctx->is, sizeof(int) * NIS, hipMemcpyDeviceToHost, s));
        MPS_HIP(hipStreamSynchronize(s));
        done = ctx->hostInts[I_CG_DONE];
        iters = ctx->hostInts[I_CG_ITER];
        if (ctx->profile)
            for (int q = 0; q < batch; q++) {
                float ms = 0.f;
                hipEventElapsedTime(&ms, ctx->evPool[2 * q], ctx->evPool[2 * q + 1]);
                spmvTimes.push_back(ms);
            }
        if (launched > 4 * maxIt + 64) break;   // safety net: the device flag ends every batch
    }
    hipLaunchKernelGGL(kTrueResidual, dim3(nb), dim3(kB), 0, s, a, c);
    hipLaunchKernelGGL(kTrueResidualFinish, dim3(1), dim3(1024), 0, s, a, nb);
    hipLaunchKernelGGL(kPressFromX, dim3(gridFor(N, kB)), dim3(kB), 0, s, a, c.x);
    MPS_HIP(hipEventRecord(ctx->ev1, s));
    MPS_HIP(hipMemcpyAsync(ctx->hostScalars, ctx->ds, sizeof(double) * NDS, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipStreamSynchronize(s));
    MPS_HIP(hipGetLastError());
    float ms = 0.f;
    hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    ctx->ppeMs = ms;
    const double* hs = ctx->hostScalars;
    ctx->solverIter = iters;
    if (ctx->profile) {
        // launches that did work: one per iteration plus the one that converged
        const int real = std::min((int)spmvTimes.size(), ctx->hostInts[I_CG_DONE] == 1 || ctx->hostInts[I_CG_DONE] == 3 ? iters + 1 : 0);
        const double nnz = (double)ctx->hostInts[I_NNZ];
        const double bytes = nnz * 12.0 + 28.0 * (double)ctx->nLiveHost;
        for (int q = 0; q < real; q++) {
            ctx->spmvMsTotal += spmvTimes[q];
            ctx->spmvBytesTotal += bytes;
        }
        ctx->spmvLaunches += real;
    }
    ctx->solverErr = hs[S_RHSN2] > 0 ? sqrt(hs[S_RN2] / hs[S_RHSN2]) : 0.0;
    ctx->trueResidual = hs[S_BN2] > 0 ? sqrt(hs[S_TRUE_RN2] / hs[S_BN2]) : 0.0;
    return MPS_OK;
}

}  // namespace mpsgpu
