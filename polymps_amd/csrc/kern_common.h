// Device helpers shared by the MPS kernels.  Expressions keep the
// reference's evaluation order (no FMA contraction: the library is built
// with -ffp-contract=off) so FP64 results match the CPU path bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "mps_internal.h"

namespace mpsgpu {

#define KEPS (10.0 * 2.220446049250313e-16)

// src/MpsParticle.cpp:57-72 (weight; weightGradient has the same forms)
__device__ __forceinline__ double wgt(double dst, double re, int t) {
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
__device__ __forceinline__ double delWgt(double dst, double re, int t) {
    switch (t) {
        case 0: return -re / (dst * dst);
        case 1: return -re / (dst * dst) + 1.0 / re;
        case 2: return -re / (dst * dst) - 1.0 / re;
        case 3: return -3.0 / re * (1.0 - dst / re) * (1.0 - dst / re);
        case 4: return -2.0 / re * (1.0 - dst / re);
        default: return -re / (dst * dst);
    }
}

// Particle slot k takes part in per-particle loops when its reference index
// is below the live particle count (the reference loops i < numParticles).
__device__ __forceinline__ bool live(const KArgs& a, int k) {
    return k < a.N && a.I[ICANON][k] < a.is[I_NLIVE];
}

// Stencil walk over the 3x3(x3) buckets around the bucket of (x,y,z) with the
// +eps truncation of MpsBucket::bucketCoordinates (src/MpsBucket.cpp:420-425),
// jz -> jy -> jx ascending; inside a bucket slots are in reference-index
// order, which reproduces the linked-list walk of the reference.  Cells
// outside the grid are skipped.
template <class F>
__device__ __forceinline__ void walkCells(const KArgs& a, double x, double y, double z, F&& f) {
    const mps_params& P = a.P;
    const int ix = (int)((x - P.domain_min[0]) * P.inv_bucket_side + KEPS);
    const int iy = (int)((y - P.domain_min[1]) * P.inv_bucket_side + KEPS);
    const int iz = (int)((z - P.domain_min[2]) * P.inv_bucket_side + KEPS);
    const int d3 = (int)(P.dim - 2.0);
    const int nx = P.num_buckets_x, ny = P.num_buckets_y, nz = P.num_buckets_z;
    for (int jz = (iz - 1) * d3; jz <= (iz + 1) * d3; jz++) {
        if (jz < 0 || jz >= nz) continue;
        for (int jy = iy - 1; jy <= iy + 1; jy++) {
            if (jy < 0 || jy >= ny) continue;
            const int row = jz * P.num_buckets_xy + jy * nx;
            int jx0 = ix - 1, jx1 = ix + 1;
            if (jx0 < 0) jx0 = 0;
            if (jx1 >= nx) jx1 = nx - 1;
            if (jx0 > jx1) continue;
            // buckets jx0..jx1 of one row are contiguous in slot order
            const int s = a.cellStart[row + jx0];
            const int e = a.cellStart[row + jx1 + 1];
            for (int j = s; j < e; j++) f(j);
        }
    }
}

// Iterate the per-step neighbour list of slot i (walk order).
template <class F>
__device__ __forceinline__ void listNbrs(const KArgs& a, int i, F&& f) {
    const int n = a.nbrCnt[i];
    const int* col = a.nbr + i;
    for (int c = 0; c < n; c++) f(col[(size_t)c * a.N]);
}

// Rref = I - 2 n n^T, n = normalised 0.5 (x - xm), |n| > eps test on |n|
// (src/MpsPndNeigh.cpp:116-136, src/MpsPressure.cpp:1366-1383)
__device__ __forceinline__ void reflectSqrt(const double* xi, const double* xm, double* n, double* R, double* ns) {
    n[0] = 0.5 * (xi[0] - xm[0]);
    n[1] = 0.5 * (xi[1] - xm[1]);
    n[2] = 0.5 * (xi[2] - xm[2]);
    const double s = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (s > KEPS) {
        n[0] = n[0] / s; n[1] = n[1] / s; n[2] = n[2] / s;
    } else {
        n[0] = 0; n[1] = 0; n[2] = 0;
    }
    R[0] = 1.0 - 2.0 * n[0] * n[0]; R[1] = 0.0 - 2.0 * n[0] * n[1]; R[2] = 0.0 - 2.0 * n[0] * n[2];
    R[3] = 0.0 - 2.0 * n[1] * n[0]; R[4] = 1.0 - 2.0 * n[1] * n[1]; R[5] = 0.0 - 2.0 * n[1] * n[2];
    R[6] = 0.0 - 2.0 * n[2] * n[0]; R[7] = 0.0 - 2.0 * n[2] * n[1]; R[8] = 1.0 - 2.0 * n[2] * n[2];
    *ns = s;
}
// same with the |n|^2 > eps test (src/MpsViscosity.cpp:1549-1562)
__device__ __forceinline__ void reflectMod2(const double* xi, const double* xm, double* n, double* R) {
    n[0] = 0.5 * (xi[0] - xm[0]);
    n[1] = 0.5 * (xi[1] - xm[1]);
    n[2] = 0.5 * (xi[2] - xm[2]);
    const double m2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (m2 > KEPS) {
        const double m = sqrt(m2);
        n[0] = n[0] / m; n[1] = n[1] / m; n[2] = n[2] / m;
    } else {
        n[0] = 0; n[1] = 0; n[2] = 0;
    }
    R[0] = 1.0 - 2.0 * n[0] * n[0]; R[1] = 0.0 - 2.0 * n[0] * n[1]; R[2] = 0.0 - 2.0 * n[0] * n[2];
    R[3] = 0.0 - 2.0 * n[1] * n[0]; R[4] = 1.0 - 2.0 * n[1] * n[1]; R[5] = 0.0 - 2.0 * n[1] * n[2];
    R[6] = 0.0 - 2.0 * n[2] * n[0]; R[7] = 0.0 - 2.0 * n[2] * n[1]; R[8] = 1.0 - 2.0 * n[2] * n[2];
}

// wave-level / block-level sum for fixed-order reductions
__device__ __forceinline__ double waveSum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double waveMax(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace mpsgpu
