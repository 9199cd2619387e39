// Bucket construction: MpsBucket::updateParticlesID (src/MpsBucket.cpp:299-353)
// as a counting sort.  Slots are ordered by (bucket, reference index) so a
// walk over a bucket's slot range visits particles in the order of the
// reference's linked list.  Ghosts, and particles whose bucket falls outside
// the grid, go to one trailing segment that no stencil visits.
#include "kern_common.h"

namespace mpsgpu {

// Bucket key of every slot and its rank among the slots of its bucket (any
// permutation of 0..count-1; the segment sort orders a bucket by canonical
// id).  Slots are mostly in bucket order from the previous step, so a
// wavefront holds a few runs of equal keys: one lane per run takes the run's
// ranks with a single atomic (per-lane atomics on a few addresses serialise).
// The bucket index is computed as updateParticlesID does: +eps only in 2D.
__global__ __launch_bounds__(256) void kBucketKey(KArgs a, int* key, int* rank, int* count) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = k < a.N;
    const mps_params& P = a.P;
    int ib = a.nB;
    const int t = in ? a.I[ITYPE][k] : P.ghost;
    if (in && a.I[ICANON][k] < a.is[I_NLIVE] && t != P.ghost) {
        const double x = a.D[DX][k], y = a.D[DY][k], z = a.D[DZ][k];
        if ((int)P.dim == 2) {
            const int ix = (int)((x - P.domain_min[0]) * P.inv_bucket_side + KEPS);
            const int iy = (int)((y - P.domain_min[1]) * P.inv_bucket_side + KEPS);
            if (ix >= 0 && ix < P.num_buckets_x && iy >= 0 && iy < P.num_buckets_y) ib = iy * P.num_buckets_x + ix;
        } else {
            const int ix = (int)((x - P.domain_min[0]) * P.inv_bucket_side);
            const int iy = (int)((y - P.domain_min[1]) * P.inv_bucket_side);
            const int iz = (int)((z - P.domain_min[2]) * P.inv_bucket_side);
            if (ix >= 0 && ix < P.num_buckets_x && iy >= 0 && iy < P.num_buckets_y && iz >= 0 && iz < P.num_buckets_z)
                ib = iz * P.num_buckets_xy + iy * P.num_buckets_x + ix;
        }
    }
    if (in) key[k] = ib;
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(in);
    int r = 0;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int kb = __shfl(ib, leader, 64);
        const unsigned long long run = __ballot(in && ib == kb) & pending;
        int base = 0;
        if (lane == leader) base = atomicAdd(&count[kb], __popcll(run));
        base = __shfl(base, leader, 64);
        if ((run >> lane) & 1ull) r = base + __popcll(run & ((1ull << lane) - 1ull));
        pending &= ~run;
    }
    if (in) rank[k] = r;
}

// exclusive scan, three passes: 4096 elements per block
constexpr int kScanBlock = 1024;
constexpr int kScanPer = 4;

__device__ __forceinline__ int blockExclusiveScan(int v, int* lds, int& total) {
    // v: per-thread value; returns exclusive prefix within the block
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (wid == 0) {
        const int nw = blockDim.x >> 6;
        int w = lane < nw ? lds[lane] : 0;
        for (int o = 1; o < 64; o <<= 1) {
            int y = __shfl_up(w, o, 64);
            if (lane >= o) w += y;
        }
        if (lane < nw) lds[lane] = w;
    }
    __syncthreads();
    const int waveOff = wid > 0 ? lds[wid - 1] : 0;
    total = lds[(blockDim.x >> 6) - 1];
    __syncthreads();
    return waveOff + x - v;
}

__global__ __launch_bounds__(kScanBlock) void kScanBlocks(const int* in, int* out, int* blockSums, int n) {
    __shared__ int lds[64];
    const int base = blockIdx.x * kScanBlock * kScanPer + threadIdx.x * kScanPer;
    int v[kScanPer];
    int s = 0;
    for (int q = 0; q < kScanPer; q++) {
        v[q] = (base + q < n) ? in[base + q] : 0;
        s += v[q];
    }
    int total;
    int pre = blockExclusiveScan(s, lds, total);
    for (int q = 0; q < kScanPer; q++) {
        if (base + q < n) out[base + q] = pre;
        pre += v[q];
    }
    if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void kScanSums(int* sums, int nb) {
    __shared__ int lds[64];
    // single block; nb <= kScanBlock * kScanPer
    const int base = threadIdx.x * kScanPer;
    int v[kScanPer];
    int s = 0;
    for (int q = 0; q < kScanPer; q++) {
        v[q] = (base + q < nb) ? sums[base + q] : 0;
        s += v[q];
    }
    int total;
    int pre = blockExclusiveScan(s, lds, total);
    for (int q = 0; q < kScanPer; q++) {
        if (base + q < nb) sums[base + q] = pre;
        pre += v[q];
    }
}

__global__ void kScanAdd(int* out, const int* sums, int n, int* tail, int tailValue) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += sums[i / (kScanBlock * kScanPer)];
    if (i == 0) *tail = tailValue;
}

__global__ void kBucketScatter(int n, const int* key, const int* rank, const int* start, int* order) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    order[start[key[k]] + rank[k]] = k;
}

// One wavefront per bucket: sort the bucket's slots by reference index
// (indices are unique, so each element's rank is the count of smaller keys).
constexpr int kSortLds = 1024;
__global__ __launch_bounds__(256) void kBucketSortSegments(const int* start, int nSeg, const int* canon,
                                                           const int* order, int* order2) {
    __shared__ int keys[4][kSortLds];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int b = blockIdx.x * 4 + wid;
    if (b >= nSeg) return;
    const int s = start[b], e = start[b + 1], n = e - s;
    if (n == 0) return;
    if (n == 1) {
        if (lane == 0) order2[s] = order[s];
        return;
    }
    if (n <= kSortLds) {
        for (int q = lane; q < n; q += 64) keys[wid][q] = canon[order[s + q]];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        for (int q = lane; q < n; q += 64) {
            const int kq = keys[wid][q];
            int r = 0;
            for (int m = 0; m < n; m++) r += keys[wid][m] < kq;
            order2[s + r] = order[s + q];
        }
    } else {
        for (int q = lane; q < n; q += 64) {
            const int kq = canon[order[s + q]];
            int r = 0;
            for (int m = 0; m < n; m++) r += canon[order[s + m]] < kq;
            order2[s + r] = order[s + q];
        }
    }
}

struct PermArgs {
    const double* src[NPD];
    double* dst[NPD];
    const int* isrc[NPI];
    int* idst[NPI];
};

__global__ void kPermute(PermArgs pa, int n, const int* order, int* slotOf) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int s = order[k];
#pragma unroll
    for (int f = 0; f < NPD; f++) pa.dst[f][k] = pa.src[f][s];
#pragma unroll
    for (int f = 0; f < NPI; f++) pa.idst[f][k] = pa.isrc[f][s];
    slotOf[pa.isrc[ICANON][s]] = k;
}

// the granular phase's carried per-step fields and PTYPE (kNNCarried)
struct PermArgsNN {
    const double* src[kNNCarriedCount];
    double* dst[kNNCarriedCount];
    const int* isrc;
    int* idst;
};
__global__ void kPermuteNN(PermArgsNN pa, int n, const int* order) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int s = order[k];
#pragma unroll
    for (int f = 0; f < kNNCarriedCount; f++) pa.dst[f][k] = pa.src[f][s];
    pa.idst[k] = pa.isrc[s];
}

int launchScan(Ctx* ctx, const int* in, int* out, int n, int tailValue) {
    // exclusive scan of n values into out[0..n-1], out[n] = tailValue
    const int per = kScanBlock * kScanPer;
    const int nb = gridFor(n, per);
    if (nb > per) {
        ctx->err = "bucket grid too large for the scan";
        return MPS_E_CAPACITY;
    }
    hipLaunchKernelGGL(kScanBlocks, dim3(nb), dim3(kScanBlock), 0, ctx->stream, in, out, ctx->scanTmp, n);
    hipLaunchKernelGGL(kScanSums, dim3(1), dim3(kScanBlock), 0, ctx->stream, ctx->scanTmp, nb);
    hipLaunchKernelGGL(kScanAdd, dim3(gridFor(n, 256)), dim3(256), 0, ctx->stream, out, ctx->scanTmp, n, out + n,
                       tailValue);
    return MPS_OK;
}

int launchBucketSort(Ctx* ctx) {
    ctx->asmMode = -1;   // slots move: an assembled matrix no longer matches them
    KArgs a = ctx->args();
    const int N = ctx->N, nB = ctx->nB;
    MPS_HIP(hipMemsetAsync(ctx->cellCount, 0, sizeof(int) * (nB + 1), ctx->stream));
    hipLaunchKernelGGL(kBucketKey, dim3(gridFor(N, 256)), dim3(256), 0, ctx->stream, a, ctx->key, ctx->rank,
                       ctx->cellCount);
    // start[0..nB] over nB+1 segments (the last one holds excluded slots), start[nB+1] = N
    int rc = launchScan(ctx, ctx->cellCount, ctx->cellStart, nB + 1, N);
    if (rc) return rc;
    hipLaunchKernelGGL(kBucketScatter, dim3(gridFor(N, 256)), dim3(256), 0, ctx->stream, N, ctx->key, ctx->rank,
                       ctx->cellStart, ctx->order);
    hipLaunchKernelGGL(kBucketSortSegments, dim3(gridFor(nB + 1, 4)), dim3(256), 0, ctx->stream, ctx->cellStart,
                       nB + 1, ctx->I[ICANON], ctx->order, ctx->order2);
    PermArgs pa;
    for (int f = 0; f < NPD; f++) { pa.src[f] = ctx->D[f]; pa.dst[f] = ctx->Dalt[f]; }
    for (int f = 0; f < NPI; f++) { pa.isrc[f] = ctx->I[f]; pa.idst[f] = ctx->Ialt[f]; }
    hipLaunchKernelGGL(kPermute, dim3(gridFor(N, 256)), dim3(256), 0, ctx->stream, pa, N, ctx->order2, ctx->slotOf);
    for (int f = 0; f < NPD; f++) std::swap(ctx->D[f], ctx->Dalt[f]);
    for (int f = 0; f < NPI; f++) std::swap(ctx->I[f], ctx->Ialt[f]);
    if (ctx->P.fluid_type == 1) {
        PermArgsNN pn;
        for (int q = 0; q < kNNCarriedCount; q++) { pn.src[q] = ctx->D[kNNCarried[q]]; pn.dst[q] = ctx->DaltNN[q]; }
        pn.isrc = ctx->I[IPTYPE];
        pn.idst = ctx->IaltPtype;
        hipLaunchKernelGGL(kPermuteNN, dim3(gridFor(N, 256)), dim3(256), 0, ctx->stream, pn, N, ctx->order2);
        for (int q = 0; q < kNNCarriedCount; q++) std::swap(ctx->D[kNNCarried[q]], ctx->DaltNN[q]);
        std::swap(ctx->I[IPTYPE], ctx->IaltPtype);
    }
    ctx->listValid = false;
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

}  // namespace mpsgpu
