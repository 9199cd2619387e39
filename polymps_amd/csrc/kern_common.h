// Device helpers shared by the MPS kernels.  Expressions keep the
// reference's evaluation order (no FMA contraction: the library is built
// with -ffp-contract=off) so FP64 results match the CPU path bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

#include "mps_internal.h"

namespace mpsgpu {

// PPE matrix (ELL) layout: KE entries per row (ellWidth: the row capacity K
// plus one 16-entry batch the row kernels read past a row's end), rows in
// tiles of 64 (a wavefront's rows).  Within a tile, element g of row i (a
// pair or a group of four entries of the row) is at (tile * G + g) * 64 +
// i % 64, G elements per row: one 16- or 8-byte load per lane reads a
// contiguous 1 KiB or 512 B of the tile, and a wavefront's whole matrix is
// one contiguous stretch (a column-major [G][N] layout streams ~8 % slower:
// a wavefront's elements are N apart).  Arrays hold ellRows(N) rows.
__host__ __device__ __forceinline__ int ellWidth(int K) { return K + kPpeCodePad; }
__host__ __device__ __forceinline__ size_t ellRows(size_t N) { return (N + 63) / 64 * 64; }
__host__ __device__ __forceinline__ size_t ellElem(int g, int i, int G) {
    return ((size_t)(i >> 6) * (size_t)G + (size_t)g) * 64 + (size_t)(i & 63);
}
// FP64 value of entry c of row i: entries 2m and 2m + 1 of a row are adjacent
// (one 16-byte load per lane per pair).
__host__ __device__ __forceinline__ size_t ellIdx64(int c, int i, int KE) {
    return (ellElem(c >> 1, i, KE >> 1) << 1) | (size_t)(c & 1);
}
// FP32 value and 16-bit column codes of entry c of row i: groups of four
// entries of a row are adjacent (one 16-byte / 8-byte load per lane per four
// entries; the assembly stores whole groups).
__host__ __device__ __forceinline__ size_t ellIdx4(int c, int i, int KE) {
    return (ellElem(c >> 2, i, KE >> 2) << 2) | (size_t)(c & 3);
}

#define KEPS (10.0 * 2.220446049250313e-16)

// slot bits of a neighbour-list entry (periodic domains keep the ghost shift
// in the two bits above)
constexpr unsigned kListSlotMask = 0x3fffffffu;

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

// Slot handled by this thread in a one-thread-per-slot launch.  (A
// renumbering that gives each of the 8 XCDs one contiguous slot range,
// xcdSlotId below, was measured slower than this plain mapping on the 1.73M
// dam break for every kernel but the list-based viscosity.)
__device__ __forceinline__ int slotId() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Slot of a thread when the workgroups, dealt round-robin to the 8 XCDs, are
// renumbered so that XCD x takes one contiguous range of slot blocks: the
// neighbour data a list consumer gathers for adjacent particles is then
// fetched into one XCD's L2 instead of eight.
__device__ __forceinline__ int xcdSlotId() {
    const int nb = gridDim.x, b = blockIdx.x;
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, k = b >> 3;
    return (x * q + min(x, r) + k) * blockDim.x + threadIdx.x;
}

// Particle slot k takes part in per-particle loops when its reference index
// is below the live particle count (the reference loops i < numParticles).
__device__ __forceinline__ bool live(const KArgs& a, int k) {
    return k < a.N && a.I[ICANON][k] < a.is[I_NLIVE];
}

// Density RHO, viscosity MEU and phase density (DNS_FL1 for phase 1,
// DNS_FL2 otherwise) of particle k: per particle with the granular phase
// (fluid_type 1), the reference's constant values for a Newtonian fluid
// (readMpsParticleFile, src/MpsInputOutput.cpp:832-837).
__device__ __forceinline__ double rhoOf(const KArgs& a, int k) {
    return a.P.fluid_type == 1 ? a.D[DRHO][k] : a.P.dns_fl1;
}
__device__ __forceinline__ double meuOf(const KArgs& a, int k) {
    return a.P.fluid_type == 1 ? a.D[DMEU][k] : a.P.knm_vs1 * a.P.dns_fl1;
}
__device__ __forceinline__ double phaseDensity(const KArgs& a, int k) {
    return (a.P.fluid_type == 1 && a.I[IPTYPE][k] != 1) ? a.P.dns_fl2 : a.P.dns_fl1;
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

// Periodic domains (one periodic axis d; MpsBoundaryCondition::setBucketBC,
// src/MpsBoundaryCondition.cpp:20-63): bucket layers 0 and n-1 of axis d are
// ghosts that list the particles of layers n-2 and 1
// (MpsBucket::copyDataBetweenBuckets, src/MpsBucket.cpp:356-417), seen
// shifted by bPBC * periodicLength with bPBC = -1 and +1
// (MpsParticle::getPeriodicLengths, src/MpsParticle.cpp:48-54).  The walk
// visits f(j, xj + plx, yj + ply, zj + plz, bPBC) in the reference's order,
// adding pl (0.0 outside the ghosts) as sqrDistBetweenParticles does
// (src/MpsParticle.cpp:35-45); slot i itself is skipped.
__device__ __forceinline__ bool periodicDomain(const mps_params& P) { return P.domain_type_bc == 1; }

template <class F>
__device__ void walkCellsPeriodic(const KArgs& a, int i, double x, double y, double z, F&& f) {
    const mps_params& P = a.P;
    const int c[3] = {(int)((x - P.domain_min[0]) * P.inv_bucket_side + KEPS),
                      (int)((y - P.domain_min[1]) * P.inv_bucket_side + KEPS),
                      (int)((z - P.domain_min[2]) * P.inv_bucket_side + KEPS)};
    const int n[3] = {P.num_buckets_x, P.num_buckets_y, P.num_buckets_z};
    const int d = P.periodic_dir[0] ? 0 : (P.periodic_dir[1] ? 1 : 2);
    const int d3 = (int)(P.dim - 2.0);
    const double* __restrict__ X = a.D[DX];
    const double* __restrict__ Y = a.D[DY];
    const double* __restrict__ Z = a.D[DZ];
    for (int jz = (c[2] - 1) * d3; jz <= (c[2] + 1) * d3; jz++) {
        if (jz < 0 || jz >= n[2]) continue;
        for (int jy = c[1] - 1; jy <= c[1] + 1; jy++) {
            if (jy < 0 || jy >= n[1]) continue;
            for (int jx = c[0] - 1; jx <= c[0] + 1; jx++) {
                if (jx < 0 || jx >= n[0]) continue;
                int b[3] = {jx, jy, jz};
                int pbc = 0;
                if (b[d] == 0) { b[d] = n[d] - 2; pbc = -1; }
                else if (b[d] == n[d] - 1) { b[d] = 1; pbc = 1; }
                const int key = b[2] * P.num_buckets_xy + b[1] * n[0] + b[0];
                const double plx = pbc * P.periodic_length[0];
                const double ply = pbc * P.periodic_length[1];
                const double plz = pbc * P.periodic_length[2];
                const int e = a.cellStart[key + 1];
                for (int j = a.cellStart[key]; j < e; j++)
                    if (j != i) f(j, X[j] + plx, Y[j] + ply, Z[j] + plz, pbc);
            }
        }
    }
}

// walkCells with the neighbour's position handed to f(j, xj, yj, zj); slot
// i itself is skipped.  Positions of four consecutive candidates are loaded
// before the first of them is tested, so the loads overlap.
template <class F>
__device__ __forceinline__ void walkCellsPos(const KArgs& a, int i, double x, double y, double z, F&& f) {
    const mps_params& P = a.P;
    if (periodicDomain(P)) {
        walkCellsPeriodic(a, i, x, y, z, [&](int j, double xj, double yj, double zj, int) { f(j, xj, yj, zj); });
        return;
    }
    const int ix = (int)((x - P.domain_min[0]) * P.inv_bucket_side + KEPS);
    const int iy = (int)((y - P.domain_min[1]) * P.inv_bucket_side + KEPS);
    const int iz = (int)((z - P.domain_min[2]) * P.inv_bucket_side + KEPS);
    const int d3 = (int)(P.dim - 2.0);
    const int nx = P.num_buckets_x, ny = P.num_buckets_y, nz = P.num_buckets_z;
    const double* __restrict__ X = a.D[DX];
    const double* __restrict__ Y = a.D[DY];
    const double* __restrict__ Z = a.D[DZ];
    for (int jz = (iz - 1) * d3; jz <= (iz + 1) * d3; jz++) {
        if (jz < 0 || jz >= nz) continue;
        for (int jy = iy - 1; jy <= iy + 1; jy++) {
            if (jy < 0 || jy >= ny) continue;
            const int row = jz * P.num_buckets_xy + jy * nx;
            int jx0 = ix - 1, jx1 = ix + 1;
            if (jx0 < 0) jx0 = 0;
            if (jx1 >= nx) jx1 = nx - 1;
            if (jx0 > jx1) continue;
            const int s = a.cellStart[row + jx0];
            const int e = a.cellStart[row + jx1 + 1];
            int j = s;
            for (; j + 4 <= e; j += 4) {
                const double x0 = X[j], x1 = X[j + 1], x2 = X[j + 2], x3 = X[j + 3];
                const double y0 = Y[j], y1 = Y[j + 1], y2 = Y[j + 2], y3 = Y[j + 3];
                const double z0 = Z[j], z1 = Z[j + 1], z2 = Z[j + 2], z3 = Z[j + 3];
                if (j != i) f(j, x0, y0, z0);
                if (j + 1 != i) f(j + 1, x1, y1, z1);
                if (j + 2 != i) f(j + 2, x2, y2, z2);
                if (j + 3 != i) f(j + 3, x3, y3, z3);
            }
            for (; j < e; j++)
                if (j != i) f(j, X[j], Y[j], Z[j]);
        }
    }
}

// The per-step neighbour list of slot i: column-major [K][N] slot indices in
// bucket-walk order, cnt[i] entries (the full list nbr / nbrCnt, or the
// small-radius list nbrS / nbrCntS).  Entries go in batches of four: the
// batch's packed (x, y, z, type) copies P4[j] and its extra per-slot values
// load(j) are all gathered before g(j, q, e) runs for each entry, and the
// next batch's slot indices are loaded while this batch is worked on.  (A
// two-stage pipeline that also keeps the next batch's gathers in flight
// needs twice the registers and ran up to 4x slower.)
struct NoExtra {};
template <class L, class G>
__device__ __forceinline__ void listNbrsX(const KArgs& a, const int* nbr, const int* cnt, int i, L&& load, G&& g) {
    const int n = cnt[i];
    const int* __restrict__ col = nbr + i;
    const double4* __restrict__ P4 = a.P4;
    const size_t N = a.N;
    int c = 0;
    if (periodicDomain(a.P)) {
        // entries carry the ghost shift in bits 30-31 (1: bPBC -1, 2: +1)
        const mps_params& P = a.P;
        for (; c < n; c++) {
            const unsigned e = (unsigned)col[c * N];
            const int j = (int)(e & kListSlotMask);
            const unsigned t = e >> 30;
            const int pbc = t == 1u ? -1 : (t == 2u ? 1 : 0);
            double4 q = P4[j];
            q.x = q.x + pbc * P.periodic_length[0];
            q.y = q.y + pbc * P.periodic_length[1];
            q.z = q.z + pbc * P.periodic_length[2];
            g(j, q, load(j));
        }
        return;
    }
    if (n >= 4) {
        int j0 = col[0], j1 = col[N], j2 = col[2 * N], j3 = col[3 * N];
        for (; c + 4 <= n; c += 4) {
            const double4 q0 = P4[j0], q1 = P4[j1], q2 = P4[j2], q3 = P4[j3];
            const auto e0 = load(j0), e1 = load(j1), e2 = load(j2), e3 = load(j3);
            // issued after the gathers: loads complete in issue order, so
            // waiting for the gathers does not wait for these.  Entries
            // c + 4 .. c + 7 exist when c + 8 <= n (n <= K: in bounds)
            int k0 = 0, k1 = 0, k2 = 0, k3 = 0;
            if (c + 8 <= n) {
                k0 = col[(c + 4) * N]; k1 = col[(c + 5) * N]; k2 = col[(c + 6) * N]; k3 = col[(c + 7) * N];
            }
            g(j0, q0, e0);
            g(j1, q1, e1);
            g(j2, q2, e2);
            g(j3, q3, e3);
            j0 = k0; j1 = k1; j2 = k2; j3 = k3;
        }
    }
    for (; c < n; c++) {
        const int j = col[c * N];
        g(j, P4[j], load(j));
    }
}

// Positions only: g(j, q).
template <class G>
__device__ __forceinline__ void listNbrsQ(const KArgs& a, const int* nbr, const int* cnt, int i, G&& g) {
    listNbrsX(a, nbr, cnt, i, [](int) { return NoExtra{}; }, [&](int j, const double4& q, NoExtra) { g(j, q); });
}

// The large-radius list with a second packed per-slot array: g(j, q, v)
// receives P4[j] and V[j], both gathered in the batch.
template <class G>
__device__ __forceinline__ void listNbrsQV(const KArgs& a, const double4* __restrict__ V, int i, G&& g) {
    listNbrsX(a, a.nbr, a.nbrCnt, i, [&](int j) { return V[j]; }, g);
}

// Particle type from the w component of P4: its low 10 bits, sign-extended
// (types in [-512, 511]); bits 10 and up hold the flags of update_bc.
__device__ __forceinline__ int p4Type(double w) { return ((((int)w) & 1023) ^ 512) - 512; }
__device__ __forceinline__ int p4FlagBits(double w) { return ((int)w) >> 10; }
__device__ __forceinline__ int p4TypeOf(const KArgs& a, double w) { return a.p4Flags ? p4Type(w) : (int)w; }

// Neighbours of slot i with their positions, f(j, xj, yj, zj): from the list
// (LIST) or from a bucket walk.
template <bool LIST, class F>
__device__ __forceinline__ void forNbrsPos(const KArgs& a, int i, double x, double y, double z, F&& f) {
    if (LIST) listNbrsQ(a, a.nbr, a.nbrCnt, i, [&](int j, const double4& q) { f(j, q.x, q.y, q.z); });
    else walkCellsPos(a, i, x, y, z, f);
}

// The same for operators whose radii are at most the small one (re_s and
// the collision distance): with the list, the small-radius list.
template <bool LIST, class F>
__device__ __forceinline__ void forNbrsPosS(const KArgs& a, int i, double x, double y, double z, F&& f) {
    if (LIST) listNbrsQ(a, a.nbrS, a.nbrCntS, i, [&](int j, const double4& q) { f(j, q.x, q.y, q.z); });
    else walkCellsPos(a, i, x, y, z, f);
}

// Small-radius neighbours with one more per-slot FP64 value S[j]:
// f(j, xj, yj, zj, sOf), sOf() returning S[j].  With the list S[j] is
// gathered in the batch (every entry is a valid slot); a walk reads it only
// when sOf is called.
template <bool LIST, class F>
__device__ __forceinline__ void forNbrsPosSX(const KArgs& a, int i, double x, double y, double z,
                                             const double* __restrict__ S, F&& f) {
    if (LIST) {
        listNbrsX(a, a.nbrS, a.nbrCntS, i, [&](int j) { return S[j]; },
                  [&](int j, const double4& q, double s) { f(j, q.x, q.y, q.z, [s]() { return s; }); });
    } else {
        walkCellsPos(a, i, x, y, z,
                     [&](int j, double xj, double yj, double zj) { f(j, xj, yj, zj, [S, j]() { return S[j]; }); });
    }
}

// The same with the neighbour's particle type, f(j, xj, yj, zj, tj).
template <bool LIST, class F>
__device__ __forceinline__ void forNbrsPosType(const KArgs& a, int i, double x, double y, double z, F&& f) {
    if (LIST) {
        listNbrsQ(a, a.nbr, a.nbrCnt, i, [&](int j, const double4& q) { f(j, q.x, q.y, q.z, p4TypeOf(a, q.w)); });
    } else {
        const int* __restrict__ T = a.I[ITYPE];
        walkCellsPos(a, i, x, y, z, [&](int j, double xj, double yj, double zj) { f(j, xj, yj, zj, T[j]); });
    }
}

// Neighbours with position, type and the update_bc flags of P4.w,
// f(j, xj, yj, zj, tj, flags); flags is 0 unless the list carries them.
template <bool LIST, class F>
__device__ __forceinline__ void forNbrsPosTypeFlags(const KArgs& a, int i, double x, double y, double z, F&& f) {
    if (LIST) {
        listNbrsQ(a, a.nbr, a.nbrCnt, i, [&](int j, const double4& q) {
            f(j, q.x, q.y, q.z, p4TypeOf(a, q.w), a.p4Flags ? p4FlagBits(q.w) : 0);
        });
    } else {
        const int* __restrict__ T = a.I[ITYPE];
        walkCellsPos(a, i, x, y, z, [&](int j, double xj, double yj, double zj) { f(j, xj, yj, zj, T[j], 0); });
    }
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
__device__ __forceinline__ double waveMin(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double waveMax(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int waveSumInt(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int waveMinInt(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int waveMaxInt(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Wave-cooperative stencil walk for one-wave (64-thread) workgroups.
//
// The rows (jz, jy) of the 3x3(x3) stencil are taken in walk order.  For
// each row the wave forms the union [us, ue) of its lanes' candidate ranges
// [s, e) (the three jx buckets of the row), stages it into LDS 64 slots at a
// time with one coalesced fetch per lane, and every lane then runs over the
// staged slots in ascending order, visiting those in its own range.  Each
// lane therefore visits exactly the slots of its own bucket walk, in walk
// order, while the loads are shared and the LDS reads are broadcasts.  When
// the union is longer than maxUnion (lanes of the wave sit in different
// bucket rows) the row is walked per lane with direct loads instead.
//
//   fetch(j) -> T             the candidate data a visit needs
//   begin(r, s, e)            row r starts; the lane's range (e may be lowered)
//   visit(j, const T&)        every j of the lane's range, ascending (j == i included)
//   end(r, s, e)              row r is done
template <class T, class Fetch, class Begin, class Visit, class End>
__device__ __forceinline__ void coopRows(const KArgs& a, bool act, double x, double y, double z, T* stage,
                                         int maxUnion, Fetch&& fetch, Begin&& begin, Visit&& visit, End&& end) {
    const mps_params& P = a.P;
    const int ix = (int)((x - P.domain_min[0]) * P.inv_bucket_side + KEPS);
    const int iy = (int)((y - P.domain_min[1]) * P.inv_bucket_side + KEPS);
    const int iz = (int)((z - P.domain_min[2]) * P.inv_bucket_side + KEPS);
    const int nx = P.num_buckets_x, ny = P.num_buckets_y, nz = P.num_buckets_z;
    const int rows = ((int)P.dim == 3) ? 9 : 3;
    const int t = threadIdx.x;
    for (int r = 0; r < rows; r++) {
        int s = 0, e = 0;
        if (act) {
            const int jz = ((int)P.dim == 3) ? iz - 1 + r / 3 : 0;
            const int jy = iy - 1 + r % 3;
            if (!(jz < 0 || jz >= nz || jy < 0 || jy >= ny)) {
                const int row = jz * P.num_buckets_xy + jy * nx;
                int jx0 = ix - 1, jx1 = ix + 1;
                if (jx0 < 0) jx0 = 0;
                if (jx1 >= nx) jx1 = nx - 1;
                if (jx0 <= jx1) {
                    s = a.cellStart[row + jx0];
                    e = a.cellStart[row + jx1 + 1];
                }
            }
            begin(r, s, e);
        }
        const bool has = act && e > s;
        const int us = waveMinInt(has ? s : INT_MAX);
        const int ue = waveMaxInt(has ? e : INT_MIN);
        if (us < ue) {
            if (ue - us <= maxUnion) {
                for (int cb = us; cb < ue; cb += 64) {
                    __syncthreads();
                    if (cb + t < ue) stage[t] = fetch(cb + t);
                    __syncthreads();
                    const int m = min(64, ue - cb);
                    for (int k = 0; k < m; k++) {
                        const int j = cb + k;
                        if (j >= s && j < e) visit(j, stage[k]);
                    }
                }
            } else if (has) {
                for (int j = s; j < e; j++) visit(j, fetch(j));
            }
        }
        if (act) end(r, s, e);
    }
}

}  // namespace mpsgpu
