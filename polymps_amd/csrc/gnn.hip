// Learned PPE operator on gfx950: a message-passing network over the
// particle neighbour graph (edges within reS) that outputs a sparse operator
//
//     (M r)_i = s_i r_i / a_ii + sum_j c_ij r_j / sqrt(a_ii a_jj)
//
// used as the warm start x0 = x + M (b - A x) and as the preconditioner of a
// flexible CG.  With its output weights at zero M is the Jacobi
// preconditioner.  Network (H = 32 channels, 3 layers, Q = 16 head channels):
//
//     h = relu(W0 f + b0)
//     layer: G = B h;  m_i = sum_j om_ij relu(G_j + E e_ij + c);  h += relu(U [h; m] + d)
//     head:  c_ij = v . relu(Pa h_i + Pb h_j + Pe e_ij + p);  s_i = exp(clip(vd . h_i + bd, -3, 3))
//
// Node features f: PND deviation (pndSmall - n0S) / n0S, velocity divergence
// times dt, free-surface flag, wall distance / reL (clipped at 1).  Edge
// features e_ij: (x_j - x_i) / l0 and w(r_ij) / w(l0), w the reS weight.
//
// The dense node products (B h, U [h; m], [Pa; Pb] h and the data products
// of the backward pass) run on MFMA in bfloat16 with FP32 accumulation;
// activations are stored in bfloat16, everything per edge runs in FP32.
// Training (hand-written backward pass + Adam) minimises
// |A M v - v|^2 / |v|^2 over the solver's residual and a random vector.
#include <cfloat>
#include <cmath>
#include <vector>

#include "kern_common.h"

namespace mpsgpu {

namespace gnn {
constexpr int H = 32, Q = 16, L = 3, KE = 64;
// flat parameter layout
constexpr int oW0 = 0, ob0 = 128;
constexpr int oLayer = 160, sLayer = 3264;
constexpr int oB(int l) { return oLayer + sLayer * l; }
constexpr int oE(int l) { return oB(l) + 1024; }
constexpr int oC(int l) { return oB(l) + 1152; }
constexpr int oU(int l) { return oB(l) + 1184; }
constexpr int oD(int l) { return oB(l) + 3232; }
constexpr int oPa = 9952, oPb = 10464, oPe = 10976, oP = 11040, oV = 11056, oVd = 11072, oBd = 11104;
constexpr int NPARAM = 11105;
// bf16 weight copies: per layer B [32][32], Bt [32][32], U [32][64], Ut [64][32]; Pab [32][32], Pabt [32][32]
constexpr int wB(int l) { return l * 6144; }
constexpr int wBt(int l) { return l * 6144 + 1024; }
constexpr int wU(int l) { return l * 6144 + 2048; }
constexpr int wUt(int l) { return l * 6144 + 4096; }
constexpr int wPab = 3 * 6144, wPabt = 3 * 6144 + 1024, NWB = 3 * 6144 + 2048;
}  // namespace gnn

using gnn::H;
using gnn::KE;
using gnn::Q;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short toBf16(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float fromBf16(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }
// the 32 channels of one node, rounded to bfloat16, as four 16-byte stores
__device__ __forceinline__ void storeBf16x32(unsigned short* dst, const float* v) {
    uint4* d = reinterpret_cast<uint4*>(dst);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint4 w;
        w.x = (unsigned)toBf16(v[8 * q + 0]) | ((unsigned)toBf16(v[8 * q + 1]) << 16);
        w.y = (unsigned)toBf16(v[8 * q + 2]) | ((unsigned)toBf16(v[8 * q + 3]) << 16);
        w.z = (unsigned)toBf16(v[8 * q + 4]) | ((unsigned)toBf16(v[8 * q + 5]) << 16);
        w.w = (unsigned)toBf16(v[8 * q + 6]) | ((unsigned)toBf16(v[8 * q + 7]) << 16);
        d[q] = w;
    }
}

struct Gnn {
    int N = 0;
    int step = 0;
    float* theta = nullptr;   // [NPARAM] master weights
    float* grad = nullptr;
    float* adamM = nullptr;
    float* adamV = nullptr;
    unsigned short* wb = nullptr;   // bf16 weight copies
    // graph (slot order)
    int* gcol = nullptr;      // [KE][N]
    int* gcnt = nullptr;
    float4* ge = nullptr;     // [KE][N]
    float* gw = nullptr;      // [KE][N] raw weights
    float* wsum = nullptr;
    float4* feat = nullptr;
    float* da = nullptr;      // 1 / sqrt|a_ii|
    float* ainv = nullptr;    // 1 / a_ii
    // activations
    unsigned short* Hs[gnn::L + 1] = {};
    unsigned short* Gs[gnn::L] = {};
    unsigned short* Ms[gnn::L] = {};
    unsigned short* As[gnn::L] = {};   // pre-activations of the updates
    unsigned short* pre0 = nullptr;
    unsigned short* AB = nullptr;      // [N][32]: Pa h | Pb h
    float* u = nullptr;
    float* s = nullptr;
    // operator
    float* cs = nullptr;      // [KE][N]
    float* dsi = nullptr;
    // backward / training buffers
    float* gz = nullptr;
    float* rv = nullptr;      // the vector M is applied to (float)
    float* dh = nullptr;      // [N][32]
    float* dX = nullptr;      // [N][32]
    float* dm = nullptr;      // [N][32]
    float* zt = nullptr;
    double* dv1 = nullptr;    // [N] double work vectors
    double* dv2 = nullptr;
    double* dv3 = nullptr;
    double* acc = nullptr;    // [4] reductions
    bool ready = false;       // operator built for the current matrix
};

// ------------------------------------------------------------- graph + features
// Edges i <- j for live j within reS (from the neighbour list or a bucket
// walk), their features and raw weights, node features, and the diagonal
// scalings of the assembled matrix.
template <bool LIST>
__global__ void kGnnGraph(KArgs a, const double* diag, int* gcol, int* gcnt, float4* ge, float* gw, float* wsum,
                          float4* feat, float* da, float* ainv) {
    const int i = slotId();
    if (i >= a.N) return;
    const double di = diag[i];
    da[i] = (float)(1.0 / sqrt(fabs(di)));
    ainv[i] = (float)(1.0 / di);
    if (!live(a, i)) {
        gcnt[i] = 0;
        wsum[i] = 0.f;
        feat[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        return;
    }
    const mps_params& P = a.P;
    const double xi = a.D[DX][i], yi = a.D[DY][i], zi = a.D[DZ][i];
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    const double w0 = wgt(P.part_dist, P.re_s, P.weight_type);
    const double il0 = 1.0 / P.part_dist;
    const size_t N = a.N;
    int c = 0;
    double ws = 0.0, div = 0.0;
    auto edge = [&](int j, double xj, double yj, double zj, double uj, double vj, double wj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        if (!(d2 < P.re_s2) || c >= KE) return;
        const double w = wgt(sqrt(d2), P.re_s, P.weight_type);
        const size_t o = (size_t)c * N + i;
        gcol[o] = j;
        ge[o] = make_float4((float)(v0 * il0), (float)(v1 * il0), (float)(v2 * il0), (float)(w / w0));
        gw[o] = (float)w;
        ws += w;
        div += ((uj - ux) * v0 + (vj - uy) * v1 + (wj - uz) * v2) * w / d2;
        c++;
    };
    const double* __restrict__ U = a.D[DU];
    const double* __restrict__ V = a.D[DV];
    const double* __restrict__ W = a.D[DW];
    if (LIST) {
        // the small-radius list (entries within reS, or a superset in walk
        // order), velocities gathered with each batch's positions
        listNbrsX(a, a.nbrS, a.nbrCntS, i, [&](int j) { return make_double4(U[j], V[j], W[j], 0.0); },
                  [&](int j, const double4& q, const double4& u) { edge(j, q.x, q.y, q.z, u.x, u.y, u.z); });
    } else {
        forNbrsPos<false>(a, i, xi, yi, zi,
                          [&](int j, double xj, double yj, double zj) { edge(j, xj, yj, zj, U[j], V[j], W[j]); });
    }
    gcnt[i] = c;
    wsum[i] = (float)ws;
    const double n0 = P.pnd_small_zero;
    const double f0 = (a.D[DPNDSMALL][i] - n0) / n0;
    const double f1 = P.dim / n0 * div * P.time_step;
    const double f2 = (a.I[IBC][i] == P.surface) ? 1.0 : 0.0;
    const double f3 = fmin(sqrt(a.D[DRIW2][i]) / P.re_l, 1.0);
    feat[i] = make_float4((float)f0, (float)f1, (float)f2, (float)f3);
}

// -------------------------------------------------------------------- encoder
__global__ void kGnnEncode(int n, const float4* feat, const float* th, unsigned short* h, unsigned short* pre0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 f = feat[i];
    float pv[H], hv[H];
#pragma unroll
    for (int ch = 0; ch < H; ch++) {
        const float* w = th + gnn::oW0 + ch * 4;
        const float p = w[0] * f.x + w[1] * f.y + w[2] * f.z + w[3] * f.w + th[gnn::ob0 + ch];
        pv[ch] = p;
        hv[ch] = relu(p);
    }
    storeBf16x32(pre0 + (size_t)i * H, pv);
    storeBf16x32(h + (size_t)i * H, hv);
}

// ------------------------------------------------------------------ MFMA GEMM
// Y[N][OUT] = X[N][K] W^T, W bfloat16 [OUT][K] row-major, X = [X1 | X2]
// (two [N][32] halves when K = 64), bf16 or FP32 (rounded to bf16 on load).
// One wavefront computes 16 rows with v_mfma_f32_16x16x16_bf16: lane l
// supplies A[l % 16][4 (l / 16) + 0..3] and B[4 (l / 16) + 0..3][l % 16] and
// receives D[4 (l / 16) + 0..3][l % 16].
enum GemmMode { GM_STORE = 0, GM_UPDATE = 1, GM_ADD = 2, GM_SPLIT = 3 };
struct GemmEpi {
    const float* bias;          // GM_UPDATE
    const unsigned short* hin;  // GM_UPDATE: residual input
    unsigned short* out;        // GM_STORE / GM_UPDATE
    unsigned short* aout;       // GM_UPDATE: pre-activation
    float* fout;                // GM_ADD: += ; GM_SPLIT: columns 0..31 +=
    float* fout2;               // GM_SPLIT: columns 32..63 =
};

template <typename XT>
__device__ __forceinline__ s4 loadA(const XT* X, size_t row, int k, bool ok);
template <>
__device__ __forceinline__ s4 loadA<unsigned short>(const unsigned short* X, size_t row, int k, bool ok) {
    s4 v = {0, 0, 0, 0};
    if (ok) {
        const uint2 q = *reinterpret_cast<const uint2*>(X + row * 32 + k);
        v[0] = (short)(q.x & 0xFFFF); v[1] = (short)(q.x >> 16);
        v[2] = (short)(q.y & 0xFFFF); v[3] = (short)(q.y >> 16);
    }
    return v;
}
template <>
__device__ __forceinline__ s4 loadA<float>(const float* X, size_t row, int k, bool ok) {
    s4 v = {0, 0, 0, 0};
    if (ok) {
        const float4 q = *reinterpret_cast<const float4*>(X + row * 32 + k);
        v[0] = (short)toBf16(q.x); v[1] = (short)toBf16(q.y);
        v[2] = (short)toBf16(q.z); v[3] = (short)toBf16(q.w);
    }
    return v;
}

template <int K, int OUT, int MODE, typename XT>
__global__ __launch_bounds__(256) void kGnnGemm(int n, const XT* X1, const XT* X2, const unsigned short* W,
                                                GemmEpi ep) {
    const int lane = threadIdx.x & 63;
    const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (row0 >= n) return;
    const int r16 = lane & 15, kq = (lane >> 4) * 4;
    const size_t arow = (size_t)row0 + r16;
    const bool aok = arow < (size_t)n;
    s4 av[K / 16];
#pragma unroll
    for (int ks = 0; ks < K / 16; ks++) {
        const int k = ks * 16 + kq;
        av[ks] = (k < 32) ? loadA<XT>(X1, arow, k, aok) : loadA<XT>(X2, arow, k - 32, aok);
    }
#pragma unroll
    for (int t = 0; t < OUT / 16; t++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        const int ncol = t * 16 + r16;
#pragma unroll
        for (int ks = 0; ks < K / 16; ks++) {
            const uint2 q = *reinterpret_cast<const uint2*>(W + (size_t)ncol * K + ks * 16 + kq);
            s4 bv;
            bv[0] = (short)(q.x & 0xFFFF); bv[1] = (short)(q.x >> 16);
            bv[2] = (short)(q.y & 0xFFFF); bv[3] = (short)(q.y >> 16);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av[ks], bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const size_t m = (size_t)row0 + kq + e;
            if (m >= (size_t)n) continue;
            const float y = acc[e];
            if (MODE == GM_STORE) {
                ep.out[m * OUT + ncol] = toBf16(y);
            } else if (MODE == GM_UPDATE) {
                const float av2 = y + ep.bias[ncol];
                ep.aout[m * 32 + ncol] = toBf16(av2);
                ep.out[m * 32 + ncol] = toBf16(fromBf16(ep.hin[m * 32 + ncol]) + relu(av2));
            } else if (MODE == GM_ADD) {
                ep.fout[m * 32 + ncol] += y;
            } else {
                if (ncol < 32) ep.fout[m * 32 + ncol] += y;
                else ep.fout2[m * 32 + ncol - 32] = y;
            }
        }
    }
}

// ------------------------------------------------------------------ messages
// m_i = sum_j (w_ij / wsum_i) relu(G_j + E e_ij + c), one thread per node.
__global__ __launch_bounds__(256) void kGnnMessage(int n, const int* gcol, const int* gcnt, const float4* ge,
                                                   const float* gw, const float* wsum, const unsigned short* G,
                                                   const float* Eg, const float* cg, unsigned short* M) {
    // the layer's E and c, shared by the block (read as LDS broadcasts)
    __shared__ float E[H * 4], cb[H];
    for (int k = threadIdx.x; k < H * 5; k += blockDim.x) {
        if (k < H * 4) E[k] = Eg[k];
        else cb[k - H * 4] = cg[k - H * 4];
    }
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc[H];
#pragma unroll
    for (int ch = 0; ch < H; ch++) acc[ch] = 0.f;
    const int cnt = gcnt[i];
    const float ws = wsum[i];
    const float inv = ws > 0.f ? 1.f / ws : 0.f;
    for (int e = 0; e < cnt; e++) {
        const size_t o = (size_t)e * n + i;
        const int j = gcol[o];
        const float4 ef = ge[o];
        const float om = gw[o] * inv;
        const uint4* gj = reinterpret_cast<const uint4*>(G + (size_t)j * H);
        const uint4 v0 = gj[0], v1 = gj[1], v2 = gj[2], v3 = gj[3];
        const unsigned wv[16] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w,
                                 v2.x, v2.y, v2.z, v2.w, v3.x, v3.y, v3.z, v3.w};
#pragma unroll
        for (int ch = 0; ch < H; ch++) {
            const unsigned bits = (ch & 1) ? (wv[ch >> 1] & 0xFFFF0000u) : (wv[ch >> 1] << 16);
            const float g = __uint_as_float(bits);
            const float q = g + E[ch * 4] * ef.x + E[ch * 4 + 1] * ef.y + E[ch * 4 + 2] * ef.z + E[ch * 4 + 3] * ef.w +
                            cb[ch];
            acc[ch] += om * relu(q);
        }
    }
    storeBf16x32(M + (size_t)i * H, acc);
}

// -------------------------------------------------------------------- head
__global__ __launch_bounds__(256) void kGnnHead(int n, const int* gcol, const int* gcnt, const float4* ge,
                                                const unsigned short* AB, const unsigned short* h, const float* th,
                                                const float* da, const float* ainv, float* cs, float* dsi, float* uo,
                                                float* so) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float pa[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) pa[q] = fromBf16(AB[(size_t)i * 32 + q]) + th[gnn::oP + q];
    float u = th[gnn::oBd];
#pragma unroll
    for (int ch = 0; ch < H; ch++) u += th[gnn::oVd + ch] * fromBf16(h[(size_t)i * H + ch]);
    const float s = expf(fminf(fmaxf(u, -3.f), 3.f));
    uo[i] = u;
    so[i] = s;
    dsi[i] = s * ainv[i];
    const float dai = da[i];
    const int cnt = gcnt[i];
    for (int e = 0; e < cnt; e++) {
        const size_t o = (size_t)e * n + i;
        const int j = gcol[o];
        const float4 ef = ge[o];
        float c = 0.f;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const float* pe = th + gnn::oPe + q * 4;
            const float t = pa[q] + fromBf16(AB[(size_t)j * 32 + Q + q]) + pe[0] * ef.x + pe[1] * ef.y +
                            pe[2] * ef.z + pe[3] * ef.w;
            c += th[gnn::oV + q] * relu(t);
        }
        cs[o] = c * dai * da[j];
    }
}

// z = M r (FP32)
__global__ void kGnnApply(int n, const int* gcol, const int* gcnt, const float* cs, const float* dsi, const float* r,
                          float* z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = dsi[i] * r[i];
    const int cnt = gcnt[i];
    for (int e = 0; e < cnt; e++) {
        const size_t o = (size_t)e * n + i;
        acc += cs[o] * r[gcol[o]];
    }
    z[i] = acc;
}
// x += z on live slots (the warm-start correction)
__global__ void kGnnAddFloat(KArgs a, const float* z, double* x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N || !live(a, i)) return;
    x[i] += (double)z[i];
}

// ------------------------------------------------------------------ backward
// Node terms of the operator and the diagonal head; initialises dh.
__global__ void kGnnBwdNode(int n, const float* gz, const float* r, const float* ainv, const float* s, const float* u,
                            const unsigned short* h, const float* th, float* dh, float* grad) {
    __shared__ float red[H + 1];
    if (threadIdx.x <= H) red[threadIdx.x] = 0.f;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float ds = gz[i] * r[i] * ainv[i];
        const float du = (u[i] > -3.f && u[i] < 3.f) ? ds * s[i] : 0.f;
#pragma unroll
        for (int ch = 0; ch < H; ch++) {
            dh[(size_t)i * H + ch] = du * th[gnn::oVd + ch];
            atomicAdd(&red[ch], du * fromBf16(h[(size_t)i * H + ch]));
        }
        atomicAdd(&red[H], du);
    }
    __syncthreads();
    if (threadIdx.x < H) atomicAdd(&grad[gnn::oVd + threadIdx.x], red[threadIdx.x]);
    if (threadIdx.x == H) atomicAdd(&grad[gnn::oBd], red[H]);
}

// Edge head, row side: dPA_i and the v, Pe, p gradients.  Thread = (node, q).
__global__ __launch_bounds__(256) void kGnnBwdHeadI(int n, const int* gcol, const int* gcnt, const float4* ge,
                                                    const unsigned short* AB, const float* th, const float* da,
                                                    const float* gz, const float* r, float* dAB, float* grad) {
    __shared__ float red[Q * 6];
    for (int k = threadIdx.x; k < Q * 6; k += blockDim.x) red[k] = 0.f;
    __syncthreads();
    const int q = threadIdx.x & (Q - 1);
    const int i = blockIdx.x * (blockDim.x / Q) + threadIdx.x / Q;
    float dPA = 0.f, dv = 0.f, dp = 0.f, dpe0 = 0.f, dpe1 = 0.f, dpe2 = 0.f, dpe3 = 0.f;
    if (i < n) {
        const float* pe = th + gnn::oPe + q * 4;
        const float pa = fromBf16(AB[(size_t)i * 32 + q]) + th[gnn::oP + q];
        const float vq = th[gnn::oV + q];
        const float gi = gz[i] * da[i];
        const int cnt = gcnt[i];
        for (int e = 0; e < cnt; e++) {
            const size_t o = (size_t)e * n + i;
            const int j = gcol[o];
            const float4 ef = ge[o];
            const float tp = pa + fromBf16(AB[(size_t)j * 32 + Q + q]) + pe[0] * ef.x + pe[1] * ef.y + pe[2] * ef.z +
                             pe[3] * ef.w;
            const float dc = gi * r[j] * da[j];
            dv += dc * relu(tp);
            if (tp > 0.f) {
                const float dt = dc * vq;
                dPA += dt;
                dp += dt;
                dpe0 += dt * ef.x; dpe1 += dt * ef.y; dpe2 += dt * ef.z; dpe3 += dt * ef.w;
            }
        }
        dAB[(size_t)i * 32 + q] = dPA;
    }
    atomicAdd(&red[q], dv);
    atomicAdd(&red[Q + q], dp);
    atomicAdd(&red[2 * Q + q * 4], dpe0);
    atomicAdd(&red[2 * Q + q * 4 + 1], dpe1);
    atomicAdd(&red[2 * Q + q * 4 + 2], dpe2);
    atomicAdd(&red[2 * Q + q * 4 + 3], dpe3);
    __syncthreads();
    for (int k = threadIdx.x; k < Q * 6; k += blockDim.x) {
        if (k < Q) atomicAdd(&grad[gnn::oV + k], red[k]);
        else if (k < 2 * Q) atomicAdd(&grad[gnn::oP + k - Q], red[k]);
        else atomicAdd(&grad[gnn::oPe + k - 2 * Q], red[k]);
    }
}

// Edge head, column side: dPB_j = sum over edges i <- j, visited from j's
// own list (the reS graph is symmetric; e_ij = -e_ji in position, same w).
__global__ __launch_bounds__(256) void kGnnBwdHeadJ(int n, const int* gcol, const int* gcnt, const float4* ge,
                                                    const unsigned short* AB, const float* th, const float* da,
                                                    const float* gz, const float* r, float* dAB) {
    const int q = threadIdx.x & (Q - 1);
    const int j = blockIdx.x * (blockDim.x / Q) + threadIdx.x / Q;
    if (j >= n) return;
    const float* pe = th + gnn::oPe + q * 4;
    const float pb = fromBf16(AB[(size_t)j * 32 + Q + q]) + th[gnn::oP + q];
    const float vq = th[gnn::oV + q];
    const float rj = r[j] * da[j];
    const int cnt = gcnt[j];
    float acc = 0.f;
    for (int e = 0; e < cnt; e++) {
        const size_t o = (size_t)e * n + j;
        const int i = gcol[o];
        const float4 ef = ge[o];
        const float tp = fromBf16(AB[(size_t)i * 32 + q]) + pb - pe[0] * ef.x - pe[1] * ef.y - pe[2] * ef.z +
                         pe[3] * ef.w;
        if (tp > 0.f) acc += gz[i] * da[i] * rj * vq;
    }
    dAB[(size_t)j * 32 + Q + q] = acc;
}

// dA = dh * [a > 0]
__global__ void kGnnMask(int n, const float* dh, const unsigned short* a, float* dA) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (size_t)n * H) return;
    dA[k] = (fromBf16(a[k]) > 0.f) ? dh[k] : 0.f;
}

// Messages, column side: dG_j = sum over edges i <- j of om_ij dm_i [q_ij > 0]
// and the E, c gradients.  Thread = (node, 8-channel group).
__global__ __launch_bounds__(256) void kGnnBwdMessage(int n, const int* gcol, const int* gcnt, const float4* ge,
                                                      const float* gw, const float* wsum, const unsigned short* G,
                                                      const float* E, const float* cb, const float* dm, float* dG,
                                                      float* gE, float* gc) {
    __shared__ float red[H * 5];
    for (int k = threadIdx.x; k < H * 5; k += blockDim.x) red[k] = 0.f;
    __syncthreads();
    const int grp = threadIdx.x & 3;
    const int j = blockIdx.x * (blockDim.x / 4) + threadIdx.x / 4;
    float acc[8], pE[8][4], pc[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        acc[k] = 0.f; pc[k] = 0.f;
        pE[k][0] = pE[k][1] = pE[k][2] = pE[k][3] = 0.f;
    }
    if (j < n) {
        float g[8];
#pragma unroll
        for (int k = 0; k < 8; k++) g[k] = fromBf16(G[(size_t)j * H + grp * 8 + k]);
        const int cnt = gcnt[j];
        for (int e = 0; e < cnt; e++) {
            const size_t o = (size_t)e * n + j;
            const int i = gcol[o];
            const float4 f = ge[o];
            const float ex = -f.x, ey = -f.y, ez = -f.z, ew = f.w;
            const float ws = wsum[i];
            const float om = ws > 0.f ? gw[o] / ws : 0.f;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int ch = grp * 8 + k;
                const float q = g[k] + E[ch * 4] * ex + E[ch * 4 + 1] * ey + E[ch * 4 + 2] * ez + E[ch * 4 + 3] * ew + cb[ch];
                if (q > 0.f) {
                    const float dq = om * dm[(size_t)i * H + ch];
                    acc[k] += dq;
                    pc[k] += dq;
                    pE[k][0] += dq * ex; pE[k][1] += dq * ey; pE[k][2] += dq * ez; pE[k][3] += dq * ew;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 8; k++) dG[(size_t)j * H + grp * 8 + k] = acc[k];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int ch = grp * 8 + k;
        atomicAdd(&red[ch], pc[k]);
        atomicAdd(&red[H + ch * 4], pE[k][0]);
        atomicAdd(&red[H + ch * 4 + 1], pE[k][1]);
        atomicAdd(&red[H + ch * 4 + 2], pE[k][2]);
        atomicAdd(&red[H + ch * 4 + 3], pE[k][3]);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < H * 5; k += blockDim.x) {
        if (k < H) atomicAdd(&gc[k], red[k]);
        else atomicAdd(&gE[k - H], red[k]);
    }
}

// Weight gradients gW[o][k] += sum_i dY[i][o] X[i][k] (X = [X1 | X2]) and
// gb[o] += sum_i dY[i][o].  Blocks take 32-node tiles through LDS.
template <int O, int K, typename XT>
__global__ __launch_bounds__(256) void kGnnWGrad(int n, int nodesPerBlock, const float* dY, const XT* X1, const XT* X2,
                                                 float* gW, float* gb) {
    __shared__ float sY[32][O];
    __shared__ float sX[32][K];
    constexpr int NP = O * K + O;
    constexpr int PER = (NP + 255) / 256;
    float acc[PER];
#pragma unroll
    for (int p = 0; p < PER; p++) acc[p] = 0.f;
    const int start = blockIdx.x * nodesPerBlock;
    const int stop = min(n, start + nodesPerBlock);
    for (int t0 = start; t0 < stop; t0 += 32) {
        __syncthreads();
        for (int k = threadIdx.x; k < 32 * O; k += 256) {
            const int r = k / O, c = k % O, i = t0 + r;
            sY[r][c] = (i < stop) ? dY[(size_t)i * O + c] : 0.f;
        }
        for (int k = threadIdx.x; k < 32 * K; k += 256) {
            const int r = k / K, c = k % K, i = t0 + r;
            float v = 0.f;
            if (i < stop) {
                if (K == 4) v = reinterpret_cast<const float*>(X1)[(size_t)i * 4 + c];
                else if (c < 32) v = (float)X1[(size_t)i * 32 + c];
                else v = (float)X2[(size_t)i * 32 + c - 32];
            }
            sX[r][c] = v;
        }
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int idx = threadIdx.x + p * 256;
            if (idx < O * K) {
                const int o = idx / K, k = idx % K;
                float s = 0.f;
                for (int r = 0; r < 32; r++) s += sY[r][o] * sX[r][k];
                acc[p] += s;
            } else if (idx < NP) {
                const int o = idx - O * K;
                float s = 0.f;
                for (int r = 0; r < 32; r++) s += sY[r][o];
                acc[p] += s;
            }
        }
    }
#pragma unroll
    for (int p = 0; p < PER; p++) {
        const int idx = threadIdx.x + p * 256;
        if (idx < O * K) atomicAdd(&gW[idx], acc[p]);
        else if (idx < NP && gb) atomicAdd(&gb[idx - O * K], acc[p]);
    }
}

// bf16 activations as float inputs of the weight-gradient kernel
struct Bf16In {
    unsigned short v;
    __device__ operator float() const { return fromBf16(v); }
};

// bfloat16 copies of the GEMM weights (and their transposes for the
// backward data products)
__global__ void kGnnRefreshWeights(const float* th, unsigned short* wb) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    for (int l = 0; l < gnn::L; l++) {
        if (k < 1024) {
            const int o = k / 32, c = k % 32;
            wb[gnn::wB(l) + k] = toBf16(th[gnn::oB(l) + k]);
            wb[gnn::wBt(l) + c * 32 + o] = toBf16(th[gnn::oB(l) + k]);
        }
        if (k < 2048) {
            const int o = k / 64, c = k % 64;
            wb[gnn::wU(l) + k] = toBf16(th[gnn::oU(l) + k]);
            wb[gnn::wUt(l) + c * 32 + o] = toBf16(th[gnn::oU(l) + k]);
        }
    }
    if (k < 1024) {
        const int o = k / 32, c = k % 32;
        const float w = (o < Q) ? th[gnn::oPa + o * 32 + c] : th[gnn::oPb + (o - Q) * 32 + c];
        wb[gnn::wPab + k] = toBf16(w);
        wb[gnn::wPabt + c * 32 + o] = toBf16(w);
    }
}

__global__ void kGnnAdam(int np, float* th, const float* g, float* m, float* v, float lr, float b1, float b2, float c1,
                         float c2, float eps) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= np) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    th[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// ------------------------------------------------------------ training loss
// y = A z - v (FP64 matrix, z and v FP32 slot vectors), accumulating |y|^2, |v|^2
__global__ void kGnnResid(KArgs a, const double* val, const unsigned short* col, const int* esc, const int* cnt,
                          const double* diag, const float* z, const float* v, double* y, double* acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double yy = 0.0, vv = 0.0;
    if (i < a.N) {
        const size_t N = a.N;
        const int ellW = ellWidth(a.K);
        double s = diag[i] * (double)z[i];
        int j = 0, e = 0;
        for (int c = 0; c < cnt[i]; c++) {
            const unsigned d = col[ellIdx4(c, i, ellW)];
            j = d ? j + (int)d : esc[(size_t)(e++) * N + i];
            s += val[ellIdx64(c, i, ellW)] * (double)z[j];
        }
        const double yi = s - (double)v[i];
        y[i] = yi;
        yy = yi * yi;
        vv = (double)v[i] * (double)v[i];
    }
    yy = waveSum(yy);
    vv = waveSum(vv);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&acc[0], yy);
        atomicAdd(&acc[1], vv);
    }
}
// gz = 2 A y / |v|^2 (A in place of A^T: the iteration matrix is symmetric
// up to the wall mirror cuts)
__global__ void kGnnLossGrad(KArgs a, const double* val, const unsigned short* col, const int* esc, const int* cnt,
                             const double* diag, const double* y, const double* acc, float* gz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const size_t N = a.N;
    const int ellW = ellWidth(a.K);
    double s = diag[i] * y[i];
    int j = 0, e = 0;
    for (int c = 0; c < cnt[i]; c++) {
        const unsigned d = col[ellIdx4(c, i, ellW)];
        j = d ? j + (int)d : esc[(size_t)(e++) * N + i];
        s += val[ellIdx64(c, i, ellW)] * y[j];
    }
    gz[i] = (float)(2.0 * s / acc[1]);
}
// training vectors: the current residual b - A x, or seeded Gaussian noise
// FP32 copy of an FP64 vector on live slots, 0 elsewhere
__global__ void kGnnFromDouble(KArgs a, const double* r, float* v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    v[i] = live(a, i) ? (float)r[i] : 0.f;
}
__global__ void kGnnResidualVec(KArgs a, const double* val, const unsigned short* col, const int* esc, const int* cnt,
                                const double* diag, const double* b, const double* x, float* v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    if (!live(a, i)) { v[i] = 0.f; return; }
    const size_t N = a.N;
    const int ellW = ellWidth(a.K);
    double s = diag[i] * x[i];
    int j = 0, e = 0;
    for (int c = 0; c < cnt[i]; c++) {
        const unsigned d = col[ellIdx4(c, i, ellW)];
        j = d ? j + (int)d : esc[(size_t)(e++) * N + i];
        s += val[ellIdx64(c, i, ellW)] * x[j];
    }
    v[i] = (float)(b[i] - s);
}
__global__ void kGnnNoise(KArgs a, unsigned long long seed, float* v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    if (!live(a, i)) { v[i] = 0.f; return; }
    auto mix = [](unsigned long long z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    const unsigned long long c = (unsigned long long)a.I[ICANON][i];
    const double u1 = ((mix(seed + (2 * c + 1) * 0x9E3779B97F4A7C15ull) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    const double u2 = (mix(seed + (2 * c + 2) * 0x9E3779B97F4A7C15ull) >> 11) * (1.0 / 9007199254740992.0);
    v[i] = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

// ============================================================== host side
static inline int g256(int n) { return (n + 255) / 256; }

void gnnFree(Ctx* ctx) {
    Gnn* g = ctx->gnn;
    if (!g) return;
    void* bufs[] = {g->theta, g->grad, g->adamM, g->adamV, g->wb, g->gcol, g->gcnt, g->ge, g->gw, g->wsum, g->feat,
                    g->da, g->ainv, g->pre0, g->AB, g->u, g->s, g->cs, g->dsi, g->gz, g->rv, g->dh, g->dX, g->dm,
                    g->zt, g->dv1, g->dv2, g->dv3, g->acc};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (int l = 0; l <= gnn::L; l++) if (g->Hs[l]) (void)hipFree(g->Hs[l]);
    for (int l = 0; l < gnn::L; l++) {
        if (g->Gs[l]) (void)hipFree(g->Gs[l]);
        if (g->Ms[l]) (void)hipFree(g->Ms[l]);
        if (g->As[l]) (void)hipFree(g->As[l]);
    }
    delete g;
    ctx->gnn = nullptr;
}

static std::vector<float> initParams(unsigned long long seed) {
    // Uniform +-sqrt(6 / (fan_in + fan_out)) for the weight matrices from a
    // splitmix64 stream over the flat index; biases and the output weights
    // v, vd, bd start at zero (M = Jacobi).
    struct T { int off, rows, cols; };
    const T mats[] = {{gnn::oW0, H, 4}, {gnn::oB(0), H, H}, {gnn::oE(0), H, 4}, {gnn::oU(0), H, 2 * H},
                      {gnn::oB(1), H, H}, {gnn::oE(1), H, 4}, {gnn::oU(1), H, 2 * H},
                      {gnn::oB(2), H, H}, {gnn::oE(2), H, 4}, {gnn::oU(2), H, 2 * H},
                      {gnn::oPa, Q, H}, {gnn::oPb, Q, H}, {gnn::oPe, Q, 4}};
    std::vector<float> th(gnn::NPARAM, 0.f);
    for (const T& m : mats) {
        const double lim = std::sqrt(6.0 / (m.rows + m.cols));
        for (int k = 0; k < m.rows * m.cols; k++) {
            const unsigned long long idx = (unsigned long long)(m.off + k) + 1;
            unsigned long long z = seed + idx * 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
            th[m.off + k] = (float)((2.0 * u - 1.0) * lim);
        }
    }
    return th;
}

int gnnEnsure(Ctx* ctx) {
    if (ctx->gnn) return MPS_OK;
    Gnn* g = new Gnn();
    ctx->gnn = g;
    const size_t n = ctx->N;
    g->N = ctx->N;
    MPS_HIP(hipMalloc(&g->theta, sizeof(float) * gnn::NPARAM));
    MPS_HIP(hipMalloc(&g->grad, sizeof(float) * gnn::NPARAM));
    MPS_HIP(hipMalloc(&g->adamM, sizeof(float) * gnn::NPARAM));
    MPS_HIP(hipMalloc(&g->adamV, sizeof(float) * gnn::NPARAM));
    MPS_HIP(hipMalloc(&g->wb, sizeof(unsigned short) * gnn::NWB));
    MPS_HIP(hipMalloc(&g->gcol, sizeof(int) * KE * n));
    MPS_HIP(hipMalloc(&g->gcnt, sizeof(int) * n));
    MPS_HIP(hipMalloc(&g->ge, sizeof(float4) * KE * n));
    MPS_HIP(hipMalloc(&g->gw, sizeof(float) * KE * n));
    MPS_HIP(hipMalloc(&g->wsum, sizeof(float) * n));
    MPS_HIP(hipMalloc(&g->feat, sizeof(float4) * n));
    MPS_HIP(hipMalloc(&g->da, sizeof(float) * n));
    MPS_HIP(hipMalloc(&g->ainv, sizeof(float) * n));
    for (int l = 0; l <= gnn::L; l++) MPS_HIP(hipMalloc(&g->Hs[l], sizeof(unsigned short) * H * n));
    for (int l = 0; l < gnn::L; l++) {
        MPS_HIP(hipMalloc(&g->Gs[l], sizeof(unsigned short) * H * n));
        MPS_HIP(hipMalloc(&g->Ms[l], sizeof(unsigned short) * H * n));
        MPS_HIP(hipMalloc(&g->As[l], sizeof(unsigned short) * H * n));
    }
    MPS_HIP(hipMalloc(&g->pre0, sizeof(unsigned short) * H * n));
    MPS_HIP(hipMalloc(&g->AB, sizeof(unsigned short) * 32 * n));
    MPS_HIP(hipMalloc(&g->u, sizeof(float) * n));
    MPS_HIP(hipMalloc(&g->s, sizeof(float) * n));
    MPS_HIP(hipMalloc(&g->cs, sizeof(float) * KE * n));
    MPS_HIP(hipMalloc(&g->dsi, sizeof(float) * n));
    MPS_HIP(hipMalloc(&g->gz, sizeof(float) * n));
    MPS_HIP(hipMalloc(&g->rv, sizeof(float) * n));
    MPS_HIP(hipMalloc(&g->dh, sizeof(float) * H * n));
    MPS_HIP(hipMalloc(&g->dX, sizeof(float) * H * n));
    MPS_HIP(hipMalloc(&g->dm, sizeof(float) * H * n));
    MPS_HIP(hipMalloc(&g->zt, sizeof(float) * n));
    MPS_HIP(hipMalloc(&g->dv1, sizeof(double) * n));
    MPS_HIP(hipMalloc(&g->dv2, sizeof(double) * n));
    MPS_HIP(hipMalloc(&g->dv3, sizeof(double) * n));
    MPS_HIP(hipMalloc(&g->acc, sizeof(double) * 4));
    return gnnSetParams(ctx, nullptr, 0);
}

// Install parameters (nullptr: the seeded initial values) and reset Adam.
int gnnSetParams(Ctx* ctx, const double* theta, unsigned long long seed) {
    Gnn* g = ctx->gnn;
    std::vector<float> th = theta ? std::vector<float>(gnn::NPARAM) : initParams(seed);
    if (theta)
        for (int k = 0; k < gnn::NPARAM; k++) th[k] = (float)theta[k];
    hipStream_t s = ctx->stream;
    MPS_HIP(hipMemcpyAsync(g->theta, th.data(), sizeof(float) * gnn::NPARAM, hipMemcpyHostToDevice, s));
    MPS_HIP(hipStreamSynchronize(s));
    MPS_HIP(hipMemsetAsync(g->adamM, 0, sizeof(float) * gnn::NPARAM, s));
    MPS_HIP(hipMemsetAsync(g->adamV, 0, sizeof(float) * gnn::NPARAM, s));
    hipLaunchKernelGGL(kGnnRefreshWeights, dim3(8), dim3(256), 0, s, g->theta, g->wb);
    g->step = 0;
    g->ready = false;
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

int gnnGetParams(Ctx* ctx, double* theta) {
    Gnn* g = ctx->gnn;
    std::vector<float> th(gnn::NPARAM);
    MPS_HIP(hipMemcpyAsync(th.data(), g->theta, sizeof(float) * gnn::NPARAM, hipMemcpyDeviceToHost, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < gnn::NPARAM; k++) theta[k] = th[k];
    return MPS_OK;
}

int gnnParamCount() { return gnn::NPARAM; }
float* gnnGradPtr(Ctx* ctx) { return ctx->gnn->grad; }

// Graph, features and forward pass for the current state and assembled matrix.
int gnnForward(Ctx* ctx) {
    int rc = gnnEnsure(ctx);
    if (rc) return rc;
    Gnn* g = ctx->gnn;
    KArgs a = ctx->args();
    hipStream_t s = ctx->stream;
    const int n = ctx->N, gb = g256(n);
    if (ctx->listValid)
        hipLaunchKernelGGL(kGnnGraph<true>, dim3(gb), dim3(256), 0, s, a, ctx->aDiag, g->gcol, g->gcnt, g->ge, g->gw,
                           g->wsum, g->feat, g->da, g->ainv);
    else
        hipLaunchKernelGGL(kGnnGraph<false>, dim3(gb), dim3(256), 0, s, a, ctx->aDiag, g->gcol, g->gcnt, g->ge, g->gw,
                           g->wsum, g->feat, g->da, g->ainv);
    hipLaunchKernelGGL(kGnnEncode, dim3(gb), dim3(256), 0, s, n, g->feat, g->theta, g->Hs[0], g->pre0);
    const int gg = (n + 63) / 64;
    for (int l = 0; l < gnn::L; l++) {
        GemmEpi e0{};
        e0.out = g->Gs[l];
        hipLaunchKernelGGL((kGnnGemm<32, 32, GM_STORE, unsigned short>), dim3(gg), dim3(256), 0, s, n, g->Hs[l],
                           g->Hs[l], g->wb + gnn::wB(l), e0);
        hipLaunchKernelGGL(kGnnMessage, dim3(gb), dim3(256), 0, s, n, g->gcol, g->gcnt, g->ge, g->gw, g->wsum,
                           g->Gs[l], g->theta + gnn::oE(l), g->theta + gnn::oC(l), g->Ms[l]);
        GemmEpi e1{};
        e1.bias = g->theta + gnn::oD(l);
        e1.hin = g->Hs[l];
        e1.out = g->Hs[l + 1];
        e1.aout = g->As[l];
        hipLaunchKernelGGL((kGnnGemm<64, 32, GM_UPDATE, unsigned short>), dim3(gg), dim3(256), 0, s, n, g->Hs[l],
                           g->Ms[l], g->wb + gnn::wU(l), e1);
    }
    GemmEpi e2{};
    e2.out = g->AB;
    hipLaunchKernelGGL((kGnnGemm<32, 32, GM_STORE, unsigned short>), dim3(gg), dim3(256), 0, s, n, g->Hs[gnn::L],
                       g->Hs[gnn::L], g->wb + gnn::wPab, e2);
    hipLaunchKernelGGL(kGnnHead, dim3(gb), dim3(256), 0, s, n, g->gcol, g->gcnt, g->ge, g->AB, g->Hs[gnn::L], g->theta,
                       g->da, g->ainv, g->cs, g->dsi, g->u, g->s);
    MPS_HIP(hipGetLastError());
    g->ready = true;
    return MPS_OK;
}

// grad = d/dtheta of sum_i gz_i (M r)_i for the vectors in g->gz and g->rv.
int gnnBackward(Ctx* ctx) {
    Gnn* g = ctx->gnn;
    hipStream_t s = ctx->stream;
    const int n = ctx->N, gb = g256(n), gg = (n + 63) / 64;
    const int npb = 2048, wg = (n + npb - 1) / npb;
    float* gr = g->grad;
    MPS_HIP(hipMemsetAsync(gr, 0, sizeof(float) * gnn::NPARAM, s));
    hipLaunchKernelGGL(kGnnBwdNode, dim3(gb), dim3(256), 0, s, n, g->gz, g->rv, g->ainv, g->s, g->u, g->Hs[gnn::L],
                       g->theta, g->dh, gr);
    const int hb = (n + 15) / 16;
    hipLaunchKernelGGL(kGnnBwdHeadI, dim3(hb), dim3(256), 0, s, n, g->gcol, g->gcnt, g->ge, g->AB, g->theta, g->da,
                       g->gz, g->rv, g->dX, gr);
    hipLaunchKernelGGL(kGnnBwdHeadJ, dim3(hb), dim3(256), 0, s, n, g->gcol, g->gcnt, g->ge, g->AB, g->theta, g->da,
                       g->gz, g->rv, g->dX);
    // [Pa; Pb] gradients (Pa and Pb are contiguous: one [32][32] block) and
    // dh += dAB [Pa; Pb]
    hipLaunchKernelGGL((kGnnWGrad<32, 32, Bf16In>), dim3(wg), dim3(256), 0, s, n, npb, g->dX,
                       reinterpret_cast<const Bf16In*>(g->Hs[gnn::L]), reinterpret_cast<const Bf16In*>(g->Hs[gnn::L]),
                       gr + gnn::oPa, nullptr);
    GemmEpi ea{};
    ea.fout = g->dh;
    hipLaunchKernelGGL((kGnnGemm<32, 32, GM_ADD, float>), dim3(gg), dim3(256), 0, s, n, g->dX, g->dX,
                       g->wb + gnn::wPabt, ea);
    for (int l = gnn::L - 1; l >= 0; l--) {
        const size_t tot = (size_t)n * H;
        hipLaunchKernelGGL(kGnnMask, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, g->dh, g->As[l], g->dX);
        hipLaunchKernelGGL((kGnnWGrad<32, 64, Bf16In>), dim3(wg), dim3(256), 0, s, n, npb, g->dX,
                           reinterpret_cast<const Bf16In*>(g->Hs[l]), reinterpret_cast<const Bf16In*>(g->Ms[l]),
                           gr + gnn::oU(l), gr + gnn::oD(l));
        GemmEpi es{};
        es.fout = g->dh;
        es.fout2 = g->dm;
        hipLaunchKernelGGL((kGnnGemm<32, 64, GM_SPLIT, float>), dim3(gg), dim3(256), 0, s, n, g->dX, g->dX,
                           g->wb + gnn::wUt(l), es);
        // dG into dX, E and c gradients
        const int mb = (n + 63) / 64;
        hipLaunchKernelGGL(kGnnBwdMessage, dim3(mb), dim3(256), 0, s, n, g->gcol, g->gcnt, g->ge, g->gw, g->wsum,
                           g->Gs[l], g->theta + gnn::oE(l), g->theta + gnn::oC(l), g->dm, g->dX, gr + gnn::oE(l),
                           gr + gnn::oC(l));
        hipLaunchKernelGGL((kGnnWGrad<32, 32, Bf16In>), dim3(wg), dim3(256), 0, s, n, npb, g->dX,
                           reinterpret_cast<const Bf16In*>(g->Hs[l]), reinterpret_cast<const Bf16In*>(g->Hs[l]),
                           gr + gnn::oB(l), nullptr);
        GemmEpi eb{};
        eb.fout = g->dh;
        hipLaunchKernelGGL((kGnnGemm<32, 32, GM_ADD, float>), dim3(gg), dim3(256), 0, s, n, g->dX, g->dX,
                           g->wb + gnn::wBt(l), eb);
    }
    // encoder
    const size_t tot = (size_t)n * H;
    hipLaunchKernelGGL(kGnnMask, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n, g->dh, g->pre0, g->dX);
    hipLaunchKernelGGL((kGnnWGrad<32, 4, float>), dim3(wg), dim3(256), 0, s, n, npb, g->dX,
                       reinterpret_cast<const float*>(g->feat), reinterpret_cast<const float*>(g->feat),
                       gr + gnn::oW0, gr + gnn::ob0);
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

// One training step on the current assembled matrix: vectors v = the
// residual b - A x of the current solution estimate and seeded noise; loss
// sum_v |A M v - v|^2 / |v|^2; Adam update.  lr = 0 evaluates the loss only.
int gnnTrainStep(Ctx* ctx, double lr, unsigned long long seed, double* lossOut) {
    int rc = gnnForward(ctx);
    if (rc) return rc;
    Gnn* g = ctx->gnn;
    KArgs a = ctx->args();
    hipStream_t s = ctx->stream;
    const int n = ctx->N, gb = g256(n);
    std::vector<float> gsum(gnn::NPARAM, 0.f);
    double loss = 0.0;
    for (int vsel = 0; vsel < 2; vsel++) {
        if (vsel == 0)
            hipLaunchKernelGGL(kGnnResidualVec, dim3(gb), dim3(256), 0, s, a, ctx->aVal, ctx->aCol, ctx->aEsc,
                               ctx->aCnt, ctx->aDiag, ctx->D[DPPEB], ctx->D[DPPEX], g->rv);
        else
            hipLaunchKernelGGL(kGnnNoise, dim3(gb), dim3(256), 0, s, a, seed, g->rv);
        hipLaunchKernelGGL(kGnnApply, dim3(gb), dim3(256), 0, s, n, g->gcol, g->gcnt, g->cs, g->dsi, g->rv, g->zt);
        MPS_HIP(hipMemsetAsync(g->acc, 0, sizeof(double) * 4, s));
        hipLaunchKernelGGL(kGnnResid, dim3(gb), dim3(256), 0, s, a, ctx->aVal, ctx->aCol, ctx->aEsc, ctx->aCnt,
                           ctx->aDiag, g->zt, g->rv, g->dv1, g->acc);
        double hacc[2];
        if (lr == 0.0) {
            MPS_HIP(hipMemcpyAsync(hacc, g->acc, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
            MPS_HIP(hipStreamSynchronize(s));
            if (hacc[1] > 0) loss += hacc[0] / hacc[1];
            continue;
        }
        hipLaunchKernelGGL(kGnnLossGrad, dim3(gb), dim3(256), 0, s, a, ctx->aVal, ctx->aCol, ctx->aEsc, ctx->aCnt,
                           ctx->aDiag, g->dv1, g->acc, g->gz);
        if ((rc = gnnBackward(ctx))) break;
        MPS_HIP(hipMemcpyAsync(hacc, g->acc, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
        MPS_HIP(hipStreamSynchronize(s));
        if (hacc[1] > 0) loss += hacc[0] / hacc[1];
        // accumulate the gradient of this vector
        std::vector<float> gh(gnn::NPARAM);
        MPS_HIP(hipMemcpyAsync(gh.data(), g->grad, sizeof(float) * gnn::NPARAM, hipMemcpyDeviceToHost, s));
        MPS_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < gnn::NPARAM; k++) gsum[k] += gh[k];
    }
    if (rc) return rc;
    if (lossOut) *lossOut = loss;
    if (lr == 0.0) return MPS_OK;
    MPS_HIP(hipMemcpyAsync(g->grad, gsum.data(), sizeof(float) * gnn::NPARAM, hipMemcpyHostToDevice, s));
    g->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)g->step), c2 = 1.f - std::pow(b2, (float)g->step);
    hipLaunchKernelGGL(kGnnAdam, dim3(g256(gnn::NPARAM)), dim3(256), 0, s, gnn::NPARAM, g->theta, g->grad, g->adamM,
                       g->adamV, (float)lr, b1, b2, c1, c2, 1e-8f);
    hipLaunchKernelGGL(kGnnRefreshWeights, dim3(8), dim3(256), 0, s, g->theta, g->wb);
    MPS_HIP(hipStreamSynchronize(s));
    MPS_HIP(hipGetLastError());
    g->ready = false;
    if (lossOut) *lossOut = loss;
    return MPS_OK;
}

// Warm start: x += M (b - A x) with the FP64 matrix.
int gnnWarmStart(Ctx* ctx) {
    Gnn* g = ctx->gnn;
    KArgs a = ctx->args();
    hipStream_t s = ctx->stream;
    const int n = ctx->N, gb = g256(n);
    if (ctx->r0Fresh)   // b - A x of the guess, left by the assembly
        hipLaunchKernelGGL(kGnnFromDouble, dim3(gb), dim3(256), 0, s, a, ctx->cgR, g->rv);
    else
        hipLaunchKernelGGL(kGnnResidualVec, dim3(gb), dim3(256), 0, s, a, ctx->aVal, ctx->aCol, ctx->aEsc, ctx->aCnt,
                           ctx->aDiag, ctx->D[DPPEB], ctx->D[DPPEX], g->rv);
    hipLaunchKernelGGL(kGnnApply, dim3(gb), dim3(256), 0, s, n, g->gcol, g->gcnt, g->cs, g->dsi, g->rv, g->zt);
    hipLaunchKernelGGL(kGnnAddFloat, dim3(gb), dim3(256), 0, s, a, g->zt, ctx->D[DPPEX]);
    ctx->r0Fresh = false;   // the guess has moved
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

// The graph, features, diagonal and operator in reference order (CSR over
// the live particles; columns are reference indices).  Arrays may be NULL
// to query the edge count.
int gnnDownloadGraph(Ctx* ctx, int* indptr, int* cols, double* e, double* w, double* feat, double* diag, double* cs,
                     double* dsi, int* nedges) {
    Gnn* g = ctx->gnn;
    const int n = ctx->N;
    hipStream_t s = ctx->stream;
    std::vector<int> gcol((size_t)KE * n), gcnt(n), canon(n);
    std::vector<float4> ge((size_t)KE * n), ft(n);
    std::vector<float> gw((size_t)KE * n), hcs((size_t)KE * n), hdsi(n);
    std::vector<double> hdiag(n);
    MPS_HIP(hipMemcpyAsync(gcnt.data(), g->gcnt, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipMemcpyAsync(gcol.data(), g->gcol, sizeof(int) * KE * n, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipMemcpyAsync(ge.data(), g->ge, sizeof(float4) * KE * n, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipMemcpyAsync(gw.data(), g->gw, sizeof(float) * KE * n, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipMemcpyAsync(hcs.data(), g->cs, sizeof(float) * KE * n, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipMemcpyAsync(hdsi.data(), g->dsi, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipMemcpyAsync(ft.data(), g->feat, sizeof(float4) * n, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipMemcpyAsync(hdiag.data(), ctx->aDiag, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipMemcpyAsync(canon.data(), ctx->I[ICANON], sizeof(int) * n, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipStreamSynchronize(s));
    const int nl = ctx->nLiveHost;
    std::vector<int> slot(n, -1);
    for (int k = 0; k < n; k++)
        if (canon[k] >= 0 && canon[k] < n) slot[canon[k]] = k;
    int p = 0;
    for (int r = 0; r < nl; r++) {
        const int k = slot[r];
        if (indptr) indptr[r] = p;
        for (int c = 0; c < gcnt[k]; c++, p++) {
            const size_t o = (size_t)c * n + k;
            if (cols) cols[p] = canon[gcol[o]];
            if (e) { e[4 * p] = ge[o].x; e[4 * p + 1] = ge[o].y; e[4 * p + 2] = ge[o].z; e[4 * p + 3] = ge[o].w; }
            if (w) w[p] = gw[o];
            if (cs) cs[p] = hcs[o];
        }
        if (feat) { feat[4 * r] = ft[k].x; feat[4 * r + 1] = ft[k].y; feat[4 * r + 2] = ft[k].z; feat[4 * r + 3] = ft[k].w; }
        if (diag) diag[r] = hdiag[k];
        if (dsi) dsi[r] = hdsi[k];
    }
    if (indptr) indptr[nl] = p;
    if (nedges) *nedges = p;
    return MPS_OK;
}

// r and gz (reference order, nlive entries) into the backward-pass buffers.
int gnnUploadVectors(Ctx* ctx, const double* r, const double* gz) {
    Gnn* g = ctx->gnn;
    const int n = ctx->N;
    hipStream_t s = ctx->stream;
    std::vector<int> canon(n);
    MPS_HIP(hipMemcpyAsync(canon.data(), ctx->I[ICANON], sizeof(int) * n, hipMemcpyDeviceToHost, s));
    MPS_HIP(hipStreamSynchronize(s));
    std::vector<float> hr(n, 0.f), hg(n, 0.f);
    const int nl = ctx->nLiveHost;
    for (int k = 0; k < n; k++) {
        const int c = canon[k];
        if (c >= 0 && c < nl) {
            hr[k] = (float)r[c];
            hg[k] = (float)gz[c];
        }
    }
    MPS_HIP(hipMemcpyAsync(g->rv, hr.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
    MPS_HIP(hipMemcpyAsync(g->gz, hg.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
    MPS_HIP(hipStreamSynchronize(s));
    return MPS_OK;
}

// FP32 preconditioner application z = M r for the inner solve.
int gnnApply32(Ctx* ctx, const float* r, float* z) {
    Gnn* g = ctx->gnn;
    const int n = ctx->N;
    hipLaunchKernelGGL(kGnnApply, dim3(g256(n)), dim3(256), 0, ctx->stream, n, g->gcol, g->gcnt, g->cs, g->dsi, r, z);
    return MPS_OK;
}

}  // namespace mpsgpu
