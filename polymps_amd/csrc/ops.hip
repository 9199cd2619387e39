// Particle operators of the MPS time step on gfx950.
//
// One thread per particle slot; slots are bucket-sorted, so the threads of a
// wavefront work on neighbouring particles and share the buckets they read.
// Neighbours come either from a bucket walk (walkCells) or from the per-step
// neighbour list (listNbrs); both yield the reference's visiting order, so
// FP64 sums are accumulated in the reference's order.
#include "kern_common.h"

namespace mpsgpu {

constexpr int kBlock = 256;

#define PX(k) a.D[DX][k]
#define PY(k) a.D[DY][k]
#define PZ(k) a.D[DZ][k]

// --------------------------------------------------- neighbour list build
// Candidate bitmasks for the largest operator radius, j != i, and the packed
// (x, y, z, type) copy that list consumers gather.  Word (r, w) of slot i is
// at mask[(r * maskW + w) * N + i]; bit b marks candidate slot
// rowStart(r) + 32 w + b of stencil row r as inside the radius.  Every lane
// writes its words in row order, so the stores coalesce.
//
// One wavefront per workgroup.  Per stencil row the wave stages the union of
// its lanes' candidate slots in LDS, 64 at a time, and every lane tests every
// staged candidate in a wave-uniform loop: the candidate's coordinates are
// broadcast LDS reads and each test shifts one bit into a 32-bit word of the
// union (u = 2u + bit, bit-reversed once per word).  A lane's own words are
// then cut out of two consecutive union words by one funnel shift, which
// drops the union's candidates before the lane's row start; the candidates
// after its row end and the particle itself are masked off per word.  Slots
// past the union's end are staged at infinity, so they never set a bit.
constexpr int kWave = 64;

// u = 2u + c: the lane's bit of the wave's compare mask shifted into u by
// one carry-in add (the compiler's select + shift + or takes two).
__device__ __forceinline__ unsigned shiftIn(unsigned u, bool c) {
    const unsigned long long m = __builtin_amdgcn_ballot_w64(c);
    unsigned r;
    unsigned long long carry;
    asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry) : "v"(u), "s"(m));
    return r;
}

// Word w of a lane's row from union words `hi` (later) and `lo` (earlier),
// the lane's row starting `sh` bits into `lo`.
__device__ __forceinline__ unsigned laneWord(unsigned hi, unsigned lo, int sh) {
    return __builtin_amdgcn_alignbit(hi, lo, (unsigned)sh);
}

// A second mask, maskS, marks the candidates inside the small radius r2s
// (<= r2) for the small-radius list (SMALL; without it only the large list
// is built and cntS is left as it is).
template <bool SMALL>
__global__ __launch_bounds__(kWave) void kBuildList(KArgs a, unsigned* mask, int* cnt, unsigned* maskS, int* cntS,
                                                    double r2, double r2s, const int* gate, int* fixList) {
    if (gate != nullptr && *gate == 0) return;
    __shared__ double4 stage[kWave];
    const int t = threadIdx.x;
    const int i = blockIdx.x * kWave + t;
    const bool inRange = i < a.N;
    const double xi = inRange ? PX(i) : 0.0, yi = inRange ? PY(i) : 0.0, zi = inRange ? PZ(i) : 0.0;
    if (inRange) a.P4[i] = make_double4(xi, yi, zi, (double)a.I[ITYPE][i]);
    const bool act = inRange && live(a, i);
    if (inRange && !act) {
        cnt[i] = 0;
        if (SMALL) cntS[i] = 0;
    }
    const double* __restrict__ X = a.D[DX];
    const double* __restrict__ Y = a.D[DY];
    const double* __restrict__ Z = a.D[DZ];
    const size_t N = a.N;
    const int W = a.maskW;
    const mps_params& P = a.P;
    const int ix = (int)((xi - P.domain_min[0]) * P.inv_bucket_side + KEPS);
    const int iy = (int)((yi - P.domain_min[1]) * P.inv_bucket_side + KEPS);
    const int iz = (int)((zi - P.domain_min[2]) * P.inv_bucket_side + KEPS);
    const int nx = P.num_buckets_x, ny = P.num_buckets_y, nz = P.num_buckets_z;
    const int rows = ((int)P.dim == 3) ? 9 : 3;
    const double inf = __builtin_inf();
    // the bounding box of the wave's particles: a candidate farther from it
    // than the radius (with a margin far above the rounding of either
    // distance) is in no lane's radius and is not tested
    const double bxl = waveMin(act ? xi : inf), bxh = waveMax(act ? xi : -inf);
    const double byl = waveMin(act ? yi : inf), byh = waveMax(act ? yi : -inf);
    const double bzl = waveMin(act ? zi : inf), bzh = waveMax(act ? zi : -inf);
    const double rCull2 = r2 * (1.0 + 1e-9);
    int total = 0, totalS = 0, over = 0;
    for (int r = 0; r < rows; r++) {
        // the lane's candidate slots [s, e) of row r (the walk's row order)
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
            if (e - s > 32 * W) {
                over = max(over, e - s);
                e = s + 32 * W;
            }
        }
        const int len = e - s;
        unsigned* const o1 = mask + (size_t)r * W * N + (act ? i : 0);
        unsigned* const o2 = maskS + (size_t)r * W * N + (act ? i : 0);
        int w = 0;   // words of the row stored
        // store word w: the lane's candidates only (before its row end, not itself)
        auto put = [&](unsigned v, unsigned vs) {
            const int rem = len - 32 * w;
            const int sk = i - s - 32 * w;
            const unsigned drop = (rem < 32 ? ~0u << rem : 0u) | ((unsigned)sk < 32u ? 1u << sk : 0u);
            v &= ~drop;
            o1[(size_t)w * N] = v;
            total += __builtin_popcount(v);
            if (SMALL) {
                vs &= ~drop;
                o2[(size_t)w * N] = vs;
                totalS += __builtin_popcount(vs);
            }
            w++;
        };
        const int us = __builtin_amdgcn_readfirstlane(waveMinInt(len > 0 ? s : INT_MAX));
        const int ue = __builtin_amdgcn_readfirstlane(waveMaxInt(len > 0 ? e : INT_MIN));
        if (us < ue && ue - us <= 2 * 32 * W) {
            const int off = s - us;
            const int sh = off & 31, q0 = off >> 5;
            unsigned prev = 0, prevS = 0;
            int c = 0;   // union words tested
            // the next 64 candidates are loaded while the current ones are tested
            double nx0 = inf, ny0 = inf, nz0 = inf;
            if (us + t < ue) {
                nx0 = X[us + t];
                ny0 = Y[us + t];
                nz0 = Z[us + t];
            }
            for (int cb = us; cb < ue; cb += kWave) {
                __syncthreads();
                stage[t] = make_double4(nx0, ny0, nz0, 0.0);
                // candidates within the radius of the wave's bounding box
                // (the only ones a lane can reach; staged infinities never are)
                const double ex = fmax(fmax(bxl - nx0, nx0 - bxh), 0.0);
                const double ey = fmax(fmax(byl - ny0, ny0 - byh), 0.0);
                const double ez = fmax(fmax(bzl - nz0, nz0 - bzh), 0.0);
                const unsigned long long rel = __builtin_amdgcn_ballot_w64(ex * ex + ey * ey + ez * ez < rCull2);
                __syncthreads();
                const int jn = cb + kWave + t;
                if (jn < ue) {
                    nx0 = X[jn];
                    ny0 = Y[jn];
                    nz0 = Z[jn];
                } else {
                    nx0 = inf;
                    ny0 = inf;
                    nz0 = inf;
                }
                const int m = min(kWave, ue - cb);
                for (int h = 0; h < m; h += 32) {
                    // the word's reachable candidates in ascending order; each
                    // one's bit enters at its place (u shifted past the others)
                    unsigned rb = (unsigned)(rel >> h);
                    unsigned U = 0, US = 0;
                    int pos = 0;
                    while (__builtin_popcount(rb) >= 4) {
                        int kk[4];
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            kk[u] = __builtin_ctz(rb);
                            rb &= rb - 1;
                        }
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const double4 q = stage[h + kk[u]];
                            const double v0 = q.x - xi, v1 = q.y - yi, v2 = q.z - zi;
                            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
                            U = shiftIn(U << (kk[u] - pos), d2 < r2);
                            if (SMALL) US = shiftIn(US << (kk[u] - pos), d2 < r2s);
                            pos = kk[u] + 1;
                        }
                    }
                    while (rb) {
                        const int k1 = __builtin_ctz(rb);
                        rb &= rb - 1;
                        const double4 q = stage[h + k1];
                        const double v0 = q.x - xi, v1 = q.y - yi, v2 = q.z - zi;
                        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
                        U = shiftIn(U << (k1 - pos), d2 < r2);
                        if (SMALL) US = shiftIn(US << (k1 - pos), d2 < r2s);
                        pos = k1 + 1;
                    }
                    // candidate h + q of the chunk at bit q
                    U = pos ? __builtin_bitreverse32(U << (32 - pos)) : 0u;
                    if (SMALL) US = pos ? __builtin_bitreverse32(US << (32 - pos)) : 0u;
                    if (c > q0 && 32 * w < len) put(laneWord(U, prev, sh), laneWord(US, prevS, sh));
                    prev = U;
                    prevS = US;
                    c++;
                }
            }
            if (c > q0 && 32 * w < len) put(laneWord(0u, prev, sh), laneWord(0u, prevS, sh));
        } else if (len > 0) {
            // a union too wide for the shared loop (lanes in distant rows): per lane
            unsigned v = 0, vs = 0;
            for (int j = s; j < e; j++) {
                const double v0 = X[j] - xi, v1 = Y[j] - yi, v2 = Z[j] - zi;
                const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
                const int k = (j - s) & 31;
                v |= (d2 < r2 ? 1u : 0u) << k;
                if (SMALL) vs |= (d2 < r2s ? 1u : 0u) << k;
                if (k == 31 || j == e - 1) {
                    put(v, vs);
                    v = 0;
                    vs = 0;
                }
            }
        }
        if (act) {
            for (; w < W; w++) {
                o1[(size_t)w * N] = 0u;
                if (SMALL) o2[(size_t)w * N] = 0u;
            }
        }
    }
    if (!act) return;
    // a stencil row with more candidates than the mask words hold (dense
    // clusters): the particle's lists are rebuilt by kFixList's direct walk
    if (over > 0) fixList[atomicAdd(&a.is[I_FIXCNT], 1)] = i;
    if (total > a.K) {
        atomicMax(&a.is[I_OVERFLOW], total);
        total = a.K;
    }
    if (SMALL && totalS > a.KS) {
        atomicMax(&a.is[I_OVERFLOW], totalS);
        totalS = a.KS;
    }
    cnt[i] = total;
    if (SMALL) cntS[i] = totalS;
}

// Lists of the particles whose candidate rows overflowed the mask windows,
// from a per-particle walk in walk order (no window limit), both radii.
__global__ __launch_bounds__(256) void kFixList(KArgs a, int* nbr, int* cnt, int* nbrS, int* cntS,
                                                const int* fixList, double r2, double r2s, const int* gate) {
    if (gate != nullptr && *gate == 0) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.is[I_FIXCNT]) return;
    const int i = fixList[t];
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const size_t N = a.N;
    int c = 0, cs = 0;
    walkCellsPos(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        if (!(d2 < r2)) return;
        if (c < a.K) nbr[(size_t)c * N + i] = j;
        c++;
        if (d2 < r2s) {
            if (cs < a.KS) nbrS[(size_t)cs * N + i] = j;
            cs++;
        }
    });
    if (c > a.K) { atomicMax(&a.is[I_OVERFLOW], c); c = a.K; }
    if (cs > a.KS) { atomicMax(&a.is[I_OVERFLOW], cs); cs = a.KS; }
    cnt[i] = c;
    cntS[i] = cs;
}

// Periodic domains: the list is built by a per-particle walk through the
// ghost layers (walkCellsPeriodic); each entry keeps its ghost shift in bits
// 30-31 for listNbrsQ.
__global__ __launch_bounds__(256) void kBuildListWalk(KArgs a, int* nbr, int* cnt, double r2) {
    const int i = slotId();
    if (i >= a.N) return;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    a.P4[i] = make_double4(xi, yi, zi, (double)a.I[ITYPE][i]);
    if (!live(a, i)) {
        cnt[i] = 0;
        return;
    }
    const size_t N = a.N;
    int c = 0;
    walkCellsPeriodic(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj, int pbc) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        if (!(d2 < r2)) return;
        if (c < a.K) nbr[(size_t)c * N + i] = (int)((unsigned)j | (pbc < 0 ? 1u << 30 : (pbc > 0 ? 2u << 30 : 0u)));
        c++;
    });
    if (c > a.K) {
        atomicMax(&a.is[I_OVERFLOW], c);
        c = a.K;
    }
    cnt[i] = c;
}

// Start slot of stencil row r of a particle in bucket (ix, iy, iz), or -1
// when the row lies outside the grid (the same rows and order as the walk).
__device__ __forceinline__ int rowStart(const KArgs& a, int ix, int iy, int iz, int r) {
    const mps_params& P = a.P;
    const int jz = ((int)P.dim == 3) ? iz - 1 + r / 3 : 0;
    const int jy = iy - 1 + r % 3;
    if (jz < 0 || jz >= P.num_buckets_z || jy < 0 || jy >= P.num_buckets_y) return -1;
    int jx0 = ix - 1;
    if (jx0 < 0) jx0 = 0;
    if (jx0 > ix + 1 || jx0 >= P.num_buckets_x) return -1;
    return a.cellStart[jz * P.num_buckets_xy + jy * P.num_buckets_x + jx0];
}

// Decode the candidate bitmasks into the column-major index list.  Every
// lane pulls its next neighbour in the same step and the k-th entries of all
// lanes are stored by one instruction into one list column, so the stores
// coalesce.  Each lane first packs the non-zero mask words of its stencil
// rows, in walk order and with the slot of their bit 0, into its own stretch
// of LDS; the next neighbour is then the lowest set bit of the current word,
// and a spent word is replaced by the one loaded ahead of it, so a step never
// waits on LDS or loops over empty words.  One wavefront per workgroup: a
// lane only reads what it wrote itself, so no barrier is needed.
template <int WMAX>
__global__ __launch_bounds__(64) void kMaskToList(KArgs a, const unsigned* mask, const int* cnt, int* nbr,
                                                  const int* gate) {
    if (gate != nullptr && *gate == 0) return;
    constexpr int kCap = 9 * WMAX;       // words of the 3-D stencil rows
    constexpr int kStride = kCap + 1;    // odd (in 8-byte units): distinct banks for one index
    __shared__ uint2 lds[64 * kStride];
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= a.N || !live(a, i)) return;
    const mps_params& P = a.P;
    const int n = cnt[i];
    if (n == 0) return;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const int ix = (int)((xi - P.domain_min[0]) * P.inv_bucket_side + KEPS);
    const int iy = (int)((yi - P.domain_min[1]) * P.inv_bucket_side + KEPS);
    const int iz = (int)((zi - P.domain_min[2]) * P.inv_bucket_side + KEPS);
    const int rows = ((int)P.dim == 3) ? 9 : 3;
    const int W = a.maskW;
    const size_t N = a.N;
    const unsigned* __restrict__ M = mask + i;
    uint2* const E = lds + threadIdx.x * kStride;   // (word, slot of bit 0), non-zero words only
    int m = 0;
    if constexpr (WMAX <= 8) {
        // every word of the rows in flight at once (one memory latency per lane)
        unsigned wd[9][WMAX];
        int sr[9];
#pragma unroll
        for (int r = 0; r < 9; r++) {
#pragma unroll
            for (int k = 0; k < WMAX; k++) wd[r][k] = (r < rows && k < W) ? M[(size_t)(r * W + k) * N] : 0u;
            sr[r] = (r < rows) ? rowStart(a, ix, iy, iz, r) : 0;
        }
#pragma unroll
        for (int r = 0; r < 9; r++)
#pragma unroll
            for (int k = 0; k < WMAX; k++)
                if (wd[r][k] != 0u) E[m++] = make_uint2(wd[r][k], (unsigned)(sr[r] + 32 * k));
    } else {
        for (int r = 0; r < rows; r++) {
            unsigned wd[WMAX];
#pragma unroll
            for (int k = 0; k < WMAX; k++) wd[k] = (k < W) ? M[(size_t)(r * W + k) * N] : 0u;
            const int sr = rowStart(a, ix, iy, iz, r);
#pragma unroll
            for (int k = 0; k < WMAX; k++)
                if (wd[k] != 0u) E[m++] = make_uint2(wd[k], (unsigned)(sr + 32 * k));
        }
    }
    // m >= 1 words hold the n set bits
    int q = 0;
    uint2 cur = E[0];
    uint2 ahead = E[m > 1 ? 1 : 0];
    auto next = [&]() -> int {
        if (cur.x == 0u) {
            cur = ahead;
            ++q;
            ahead = E[q + 1 < m ? q + 1 : q];
        }
        const int b = __builtin_ctz(cur.x);
        cur.x &= cur.x - 1u;
        return (int)cur.y + b;
    };
    constexpr int kChunk = 8;
    for (int c0 = 0; c0 < n; c0 += kChunk) {
        int e[kChunk];
#pragma unroll
        for (int k = 0; k < kChunk; k++) e[k] = (c0 + k < n) ? next() : 0;
#pragma unroll
        for (int k = 0; k < kChunk; k++)
            if (c0 + k < n) nbr[(size_t)(c0 + k) * N + i] = e[k];
    }
}

// ------------------------------------------------------------- viscosity
// MpsViscosity::calcViscosity (src/MpsViscosity.cpp:20-102)
template <bool LIST>
__global__ void kViscosity(KArgs a) {
    // gathers 64 B per neighbour: the XCD-contiguous order keeps them in one
    // L2 (HBM fetch 7.0 -> 1.9 GB per launch at 1.73M, 1.65 -> 1.56 ms)
    const int i = xcdSlotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double meu = P.knm_vs1 * P.dns_fl1, rho = P.dns_fl1;
    double nuF;
    if ((meu + meu) > KEPS) nuF = 2.0 * meu * meu / (meu + meu);
    else nuF = 0.0;
    nuF = nuF / rho;
    const double nuW = 2.0 * meu / rho;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double ax = 0, ay = 0, az = 0;
    auto term = [&](double xj, double yj, double zj, int tj, double uj, double vj, double wj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_l2 && (d2 < dm2 || pw)) {
            const double wL = wgt(sqrt(d2), P.re_l, P.weight_type);
            const double nu = (tj == P.wall) ? nuW : nuF;
            ax += (uj - ux) * wL * nu;
            ay += (vj - uy) * wL * nu;
            az += (wj - uz) * wL * nu;
        }
    };
    if (LIST) {
        // positions and packed velocities (V4) of a batch gathered together
        listNbrsQV(a, a.V4, i, [&](int, const double4& q, const double4& v) {
            term(q.x, q.y, q.z, p4TypeOf(a, q.w), v.x, v.y, v.z);
        });
    } else {
        forNbrsPosType<false>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj, int tj) {
            term(xj, yj, zj, tj, a.D[DU][j], a.D[DV][j], a.D[DW][j]);
        });
    }
    a.D[DAX][i] = P.coeff_visc_multiphase * ax;
    a.D[DAY][i] = P.coeff_visc_multiphase * ay;
    a.D[DAZ][i] = P.coeff_visc_multiphase * az;
}

__global__ void kPackVel(KArgs a, double4* v4) {
    const int i = slotId();
    if (i >= a.N) return;
    v4[i] = make_double4(a.D[DU][i], a.D[DV][i], a.D[DW][i], 0.0);
}

// The same from a cooperative bucket walk (one-wave workgroups): the
// candidates' position, velocity and type are staged in LDS.
struct ViscCand {
    double x, y, z, u, v, w, wall;
};
__global__ __launch_bounds__(kWave) void kViscosityWalk(KArgs a) {
    __shared__ ViscCand stage[kWave];
    const int i = blockIdx.x * kWave + threadIdx.x;
    const bool act = i < a.N && live(a, i);
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double meu = P.knm_vs1 * P.dns_fl1, rho = P.dns_fl1;
    double nuF;
    if ((meu + meu) > KEPS) nuF = 2.0 * meu * meu / (meu + meu);
    else nuF = 0.0;
    nuF = nuF / rho;
    const double nuW = 2.0 * meu / rho;
    const int k = act ? i : 0;
    const double xi = PX(k), yi = PY(k), zi = PZ(k);
    const double ux = a.D[DU][k], uy = a.D[DV][k], uz = a.D[DW][k];
    const double mx = a.D[DMX][k], my = a.D[DMY][k], mz = a.D[DMZ][k];
    const double* __restrict__ U = a.D[DU];
    const double* __restrict__ V = a.D[DV];
    const double* __restrict__ Wv = a.D[DW];
    const int* __restrict__ T = a.I[ITYPE];
    double ax = 0, ay = 0, az = 0;
    coopRows(
        a, act, xi, yi, zi, stage, 2 * 32 * a.maskW,
        [&](int j) {
            return ViscCand{PX(j), PY(j), PZ(j), U[j], V[j], Wv[j], (T[j] == P.wall) ? 1.0 : 0.0};
        },
        [&](int, int&, int&) {},
        [&](int j, const ViscCand& q) {
            if (j == i) return;
            const double v0 = q.x - xi, v1 = q.y - yi, v2 = q.z - zi;
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = q.x - mx, w1 = q.y - my, w2 = q.z - mz;
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_l2 && (d2 < dm2 || pw)) {
                const double wL = wgt(sqrt(d2), P.re_l, P.weight_type);
                const double nu = (q.wall != 0.0) ? nuW : nuF;
                ax += (q.u - ux) * wL * nu;
                ay += (q.v - uy) * wL * nu;
                az += (q.w - uz) * wL * nu;
            }
        },
        [&](int, int, int) {});
    if (!act) return;
    a.D[DAX][i] = P.coeff_visc_multiphase * ax;
    a.D[DAY][i] = P.coeff_visc_multiphase * ay;
    a.D[DAZ][i] = P.coeff_visc_multiphase * az;
}

// MpsParticle::addGravity (src/MpsParticle.cpp:111-130)
__global__ void kAddGravity(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const double ax = a.D[DAX][i] + a.P.gravity[0];
    const double ay = a.D[DAY][i] + a.P.gravity[1];
    const double az = a.D[DAZ][i] + a.P.gravity[2];
    a.D[DAX][i] = ax; a.D[DAY][i] = ay; a.D[DAZ][i] = az;
    a.D[DASX][i] = ax; a.D[DASY][i] = ay; a.D[DASZ][i] = az;
}

// MpsParticleVelPos::predictionVelocityPosition (src/MpsParticleVelPos.cpp:20-50)
__global__ void kPrediction(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const double far = 10e8 * P.part_dist;
    if (a.I[ITYPE][i] == P.fluid) {
        const double u = a.D[DU][i] + a.D[DAX][i] * P.time_step;
        const double v = a.D[DV][i] + a.D[DAY][i] * P.time_step;
        const double w = a.D[DW][i] + a.D[DAZ][i] * P.time_step;
        a.D[DU][i] = u; a.D[DV][i] = v; a.D[DW][i] = w;
        a.D[DX][i] = a.D[DX][i] + u * P.time_step;
        a.D[DY][i] = a.D[DY][i] + v * P.time_step;
        a.D[DZ][i] = a.D[DZ][i] + w * P.time_step;
    }
    a.D[DAX][i] = 0.0; a.D[DAY][i] = 0.0; a.D[DAZ][i] = 0.0;
    a.D[DDVX][i] = 0.0; a.D[DDVY][i] = 0.0; a.D[DDVZ][i] = 0.0;
    a.D[DNPX][i] = 0.0; a.D[DNPY][i] = 0.0; a.D[DNPZ][i] = 0.0;
    a.I[INNWALL][i] = 0;
    a.I[INEARWALL][i] = 0;
    a.D[DRIW2][i] = far;
    if (P.wall_type == MPS_WALL_POLYGON) {
        a.D[DMX][i] = far; a.D[DMY][i] = far; a.D[DMZ][i] = far;
    }
}

// MpsParticleVelPos::correctionVelocityPosition (src/MpsParticleVelPos.cpp:53-97);
// velMax = sqrt(max |u|^2) over fluid particles, reduced with an integer max
// on the (non-negative) bit pattern so the result does not depend on order.
// The largest |u|^2 of each block goes to red[block]; kVelMaxFinish takes
// the maximum of the partials (one atomic per wavefront on a single address
// serialises device-wide and tripled the kernel's time).
__global__ __launch_bounds__(kBlock) void kCorrection(KArgs a) {
    const int i = slotId();
    double v2 = 0.0;
    if (live(a, i) && a.I[ITYPE][i] == a.P.fluid) {
        const mps_params& P = a.P;
        const double ax = a.D[DAX][i], ay = a.D[DAY][i], az = a.D[DAZ][i];
        const double u = a.D[DU][i] + ax * P.time_step;
        const double v = a.D[DV][i] + ay * P.time_step;
        const double w = a.D[DW][i] + az * P.time_step;
        a.D[DU][i] = u; a.D[DV][i] = v; a.D[DW][i] = w;
        a.D[DX][i] = a.D[DX][i] + ax * P.time_step * P.time_step;
        a.D[DY][i] = a.D[DY][i] + ay * P.time_step * P.time_step;
        a.D[DZ][i] = a.D[DZ][i] + az * P.time_step * P.time_step;
        a.D[DAX][i] = 0.0; a.D[DAY][i] = 0.0; a.D[DAZ][i] = 0.0;
        v2 = u * u + v * v + w * w;
    }
    __shared__ double wm[kBlock / 64];
    v2 = waveMax(v2);
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v2;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wm[0];
        for (int q = 1; q < kBlock / 64; q++) m = fmax(m, wm[q]);
        a.red[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(1024) void kVelMaxFinish(KArgs a, int nb) {
    __shared__ double wm[16];
    double m = 0.0;
    for (int q = threadIdx.x; q < nb; q += blockDim.x) m = fmax(m, a.red[q]);
    m = waveMax(m);
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = fmax(m, wm[q]);
    a.ds[S_VELMAX] = sqrt(m);
}

// --------------------------------------------------------------- domain
// MpsParticle::checkParticleOutDomain (src/MpsParticle.cpp:234-402).  Flag
// pass collects the reference indices of fluid particles outside the domain
// less one bucket; one thread then replays the reference's compaction in
// ascending reference order (a slot receives the data of the last live
// particle, the last particle becomes a ghost).
// Periodic wrap of MpsParticle::checkParticleOutDomain
// (src/MpsParticle.cpp:236-269): every particle past the physical domain of
// the periodic axis moves by one periodic length.
__global__ void kPeriodicWrap(KArgs a) {
    const int i = slotId();
    if (i >= a.N || !live(a, i)) return;
    const mps_params& P = a.P;
    double* const X[3] = {a.D[DX], a.D[DY], a.D[DZ]};
    for (int k = 0; k < 3; k++) {
        if (!P.periodic_dir[k]) continue;
        double v = X[k][i];
        if (v > P.phys_max[k]) v -= P.periodic_length[k];
        if (v < P.phys_min[k]) v += P.periodic_length[k];
        X[k][i] = v;
    }
}

__global__ void kOutFlag(KArgs a, int* outList, int cap) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (a.I[ITYPE][i] != P.fluid) return;
    const double x = PX(i), y = PY(i), z = PZ(i);
    const bool out = x > P.domain_max[0] - P.bucket_side || x < P.domain_min[0] + P.bucket_side ||
                     y > P.domain_max[1] - P.bucket_side || y < P.domain_min[1] + P.bucket_side ||
                     (P.dim == 3 && (z > P.domain_max[2] - P.bucket_side || z < P.domain_min[2] + P.bucket_side));
    if (!out) return;
    const int s = atomicAdd(&a.is[I_OUTCNT], 1);
    if (s < cap) outList[s] = a.I[ICANON][i];
    else a.is[I_COMPACT_OVF] = 1;
}

__global__ void kCompact(KArgs a, int* outList, int cap, const int* slotOf) {
    int n = a.is[I_OUTCNT];
    if (n == 0) return;
    a.is[I_REBUILD] = 1;   // particles change slots: a list kept over the step is void
    if (n > cap) n = cap;
    for (int p = 1; p < n; p++) {   // insertion sort (few entries)
        const int v = outList[p];
        int q = p - 1;
        while (q >= 0 && outList[q] > v) { outList[q + 1] = outList[q]; q--; }
        outList[q + 1] = v;
    }
    const mps_params& P = a.P;
    int newN = a.is[I_NLIVE];
    for (int p = 0; p < n; p++) {
        const int c = outList[p];
        if (c >= newN) break;
        const int L = newN - 1;
        const int kc = slotOf[c], kl = slotOf[L];
        if (kc != kl) {
            for (int f = 0; f < ND; f++) a.D[f][kc] = a.D[f][kl];
            for (int f = 0; f < NI; f++)
                if (f != ICANON) a.I[f][kc] = a.I[f][kl];
        }
        a.I[ITYPE][kl] = P.ghost;
        a.I[IBC][kl] = P.other;
        a.I[INEARWALL][kl] = 0;
        a.I[INEARMESH][kl] = 0;
        a.D[DRIW2][kl] = 10e8 * P.part_dist;
        a.D[DU][kl] = 0.0; a.D[DV][kl] = 0.0; a.D[DW][kl] = 0.0;
        a.D[DPRESS][kl] = 0.0;
        a.D[DX][kl] = P.domain_max[0] - P.part_dist;
        a.D[DY][kl] = P.domain_max[1] - P.part_dist;
        a.D[DZ][kl] = (P.dim == 2) ? 0.0 : P.domain_max[2] - P.part_dist;
        newN--;
    }
    a.is[I_NLIVE] = newN;
    a.is[I_OUTCNT] = 0;
}

// ------------------------------------------------------------ collisions
// MpsParticleCollision::checkParticleCollisions (src/MpsParticleCollision.cpp:22-119)
template <bool LIST>
__global__ void kCollisionPC(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (a.I[ITYPE][i] != P.fluid) return;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double mi = phaseDensity(a, i);
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double dux = 0, duy = 0, duz = 0;
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.dist_collision_limit2 && (d2 < dm2 || pw)) {
            double fDT = (ux - a.D[DU][j]) * v0 + (uy - a.D[DV][j]) * v1 + (uz - a.D[DW][j]) * v2;
            if (fDT > 0.0) {
                const bool fj = a.I[ITYPE][j] == P.fluid;
                const double mj = fj ? phaseDensity(a, j) : P.dns_wall;
                fDT *= P.restitution_collision * mj / (mi + mj) / d2;
                if (fj) {
                    dux -= v0 * fDT; duy -= v1 * fDT; duz -= v2 * fDT;
                } else {
                    dux -= 2.0 * v0 * fDT; duy -= 2.0 * v1 * fDT; duz -= 2.0 * v2 * fDT;
                }
            }
        }
    });
    a.D[DDVX][i] = dux; a.D[DDVY][i] = duy; a.D[DDVZ][i] = duz;
}

// checkDynamicParticleCollisions (src/MpsParticleCollision.cpp:124-292)
template <bool LIST>
__global__ void kCollisionDPC(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (a.I[ITYPE][i] != P.fluid) return;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double Wij5 = 0.5 * 0.5 * 0.5 * 0.5 * 3.0;
    const double gmod = sqrt(P.gravity[0] * P.gravity[0] + P.gravity[1] * P.gravity[1] + P.gravity[2] * P.gravity[2]);
    const double mi = phaseDensity(a, i), rhoi = rhoOf(a, i);
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    const double pi = a.D[DPRESS][i];
    double dux = 0, duy = 0, duz = 0;
    forNbrsPos<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.part_dist * P.part_dist && (d2 < dm2 || pw)) {
            const bool fj = a.I[ITYPE][j] == P.fluid;
            const double mj = fj ? phaseDensity(a, j) : P.dns_wall;
            const double dst = sqrt(d2);
            double Wij = 0.0;
            if (dst > KEPS) {
                const double q1 = 1.0 - dst * P.inv_part_dist;
                const double q2 = 4.0 * dst * P.inv_part_dist + 1.0;
                Wij = q1 * q1 * q1 * q1 * q2;
            }
            const double chi = sqrt(Wij / Wij5);
            const double kappa = (dst < 0.5 * P.part_dist) ? 1.0 : chi;
            double fDT = (ux - a.D[DU][j]) * v0 + (uy - a.D[DV][j]) * v1 + (uz - a.D[DW][j]) * v2;
            if (fDT > 0.0) {
                fDT *= kappa * 2.0 * mj / (mi + mj) / d2;
                if (fj) {
                    dux -= v0 * fDT; duy -= v1 * fDT; duz -= v2 * fDT;
                } else {
                    dux -= 2.0 * v0 * fDT; duy -= 2.0 * v1 * fDT; duz -= 2.0 * v2 * fDT;
                }
            } else {
                const double pmax = 2.0 / 3.0 * rhoi * gmod * 0.2;
                const double pmin = rhoi * gmod * P.part_dist;
                const double ptil = fmax(fmin(P.lambda_collision * fabs(pi + a.D[DPRESS][j]), P.lambda_collision * pmax), pmin);
                const double pb = ptil * chi;
                const double rep = P.time_step / mi * chi * pb / d2;
                if (fj) {
                    dux -= v0 * rep; duy -= v1 * rep; duz -= v2 * rep;
                } else {
                    dux -= 2.0 * v0 * rep; duy -= 2.0 * v1 * rep; duz -= 2.0 * v2 * rep;
                }
            }
        }
    });
    a.D[DDVX][i] = dux; a.D[DDVY][i] = duy; a.D[DDVZ][i] = duz;
}

__global__ void kApplyDvel(KArgs a, int movePos) {
    const int i = slotId();
    if (!live(a, i)) return;
    if (a.I[ITYPE][i] == a.P.fluid) {
        const double dx = a.D[DDVX][i], dy = a.D[DDVY][i], dz = a.D[DDVZ][i];
        a.D[DU][i] = a.D[DU][i] + dx;
        a.D[DV][i] = a.D[DV][i] + dy;
        a.D[DW][i] = a.D[DW][i] + dz;
        if (movePos) {
            a.D[DX][i] = a.D[DX][i] + dx * a.P.time_step;
            a.D[DY][i] = a.D[DY][i] + dy * a.P.time_step;
            a.D[DZ][i] = a.D[DZ][i] + dz * a.P.time_step;
        }
    }
    a.D[DDVX][i] = 0.0; a.D[DDVY][i] = 0.0; a.D[DDVZ][i] = 0.0;
}

// ------------------------------------------------------------ mesh walls
// Closest point on triangle abc to p (Ericson, Real-Time Collision Detection 5.1.5)
__device__ __forceinline__ void closestPtTri(const double* p, const double* A, const double* B, const double* C, double* q) {
    const double ab[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    const double ac[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
    const double ap[3] = {p[0] - A[0], p[1] - A[1], p[2] - A[2]};
    const double d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
    const double d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
    if (d1 <= 0.0 && d2 <= 0.0) { q[0] = A[0]; q[1] = A[1]; q[2] = A[2]; return; }
    const double bp[3] = {p[0] - B[0], p[1] - B[1], p[2] - B[2]};
    const double d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
    const double d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
    if (d3 >= 0.0 && d4 <= d3) { q[0] = B[0]; q[1] = B[1]; q[2] = B[2]; return; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double v = d1 / (d1 - d3);
        q[0] = A[0] + v * ab[0]; q[1] = A[1] + v * ab[1]; q[2] = A[2] + v * ab[2];
        return;
    }
    const double cp[3] = {p[0] - C[0], p[1] - C[1], p[2] - C[2]};
    const double d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
    const double d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
    if (d6 >= 0.0 && d5 <= d6) { q[0] = C[0]; q[1] = C[1]; q[2] = C[2]; return; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double w = d2 / (d2 - d6);
        q[0] = A[0] + w * ac[0]; q[1] = A[1] + w * ac[1]; q[2] = A[2] + w * ac[2];
        return;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        q[0] = B[0] + w * (C[0] - B[0]); q[1] = B[1] + w * (C[1] - B[1]); q[2] = B[2] + w * (C[2] - B[2]);
        return;
    }
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    q[0] = A[0] + ab[0] * v + ac[0] * w;
    q[1] = A[1] + ab[1] * v + ac[1] * w;
    q[2] = A[2] + ab[2] * v + ac[2] * w;
}

// PolygonMesh::closestPointPNDBoundaryAABB (src/PolygonMesh.cpp:585-715),
// brute force over the faces; ties keep the lowest face index.
__global__ void kClosestPoint(KArgs a, int meshType) {
    const int i = slotId();
    if (!live(a, i)) return;
    const double p[3] = {PX(i), PY(i), PZ(i)};
    double best = 1.79769313486231570e308, q[3], bq[3] = {0, 0, 0};
    int bf = -1;
    for (int f = 0; f < a.nf; f++) {
        const int t0 = a.mF[f * 3], t1 = a.mF[f * 3 + 1], t2 = a.mF[f * 3 + 2];
        closestPtTri(p, a.mV + t0 * 3, a.mV + t1 * 3, a.mV + t2 * 3, q);
        const double e0 = p[0] - q[0], e1 = p[1] - q[1], e2 = p[2] - q[2];
        const double d2 = e0 * e0 + e1 * e1 + e2 * e2;
        if (d2 < best) { best = d2; bf = f; bq[0] = q[0]; bq[1] = q[1]; bq[2] = q[2]; }
    }
    if (bf < 0) return;
    if (best < a.D[DRIW2][i]) {
        a.D[DRIW2][i] = best;
        a.D[DWPX][i] = bq[0]; a.D[DWPY][i] = bq[1]; a.D[DWPZ][i] = bq[2];
        a.D[DMX][i] = p[0] + 2.0 * (bq[0] - p[0]);
        a.D[DMY][i] = p[1] + 2.0 * (bq[1] - p[1]);
        a.D[DMZ][i] = p[2] + 2.0 * (bq[2] - p[2]);
        a.D[DPNX][i] = a.mN[bf * 3]; a.D[DPNY][i] = a.mN[bf * 3 + 1]; a.D[DPNZ][i] = a.mN[bf * 3 + 2];
        a.I[INEARMESH][i] = meshType;
    }
}

// The same query through the triangle BVH: nearer child first, a subtree is
// skipped when its box is farther than the best distance (with a relative
// margin far above rounding, so that no subtree holding an equally close face
// is skipped), and among equally close faces the lowest index wins -- the
// result of the brute-force loop above.
__device__ __forceinline__ double boxDist2(const BvhNode& n, const double* p) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double d = fmax(fmax(n.lo[k] - p[k], p[k] - n.hi[k]), 0.0);
        s += d * d;
    }
    return s;
}
constexpr int kBvhStack = 64;
__global__ void kClosestPointBvh(KArgs a, int meshType) {
    const int i = slotId();
    if (!live(a, i)) return;
    const double p[3] = {PX(i), PY(i), PZ(i)};
    double best = 1.79769313486231570e308, q[3], bq[3] = {0, 0, 0};
    int bf = -1;
    int stack[kBvhStack];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const BvhNode n = a.bvh[stack[--sp]];
        if (boxDist2(n, p) > best * (1.0 + 1e-9)) continue;
        if (n.b < 0) {
            for (int t = n.a; t < n.a - n.b; t++) {
                const int f = a.bvhTri[t];
                const int t0 = a.mF[f * 3], t1 = a.mF[f * 3 + 1], t2 = a.mF[f * 3 + 2];
                closestPtTri(p, a.mV + t0 * 3, a.mV + t1 * 3, a.mV + t2 * 3, q);
                const double e0 = p[0] - q[0], e1 = p[1] - q[1], e2 = p[2] - q[2];
                const double d2 = e0 * e0 + e1 * e1 + e2 * e2;
                if (d2 < best || (d2 == best && f < bf)) { best = d2; bf = f; bq[0] = q[0]; bq[1] = q[1]; bq[2] = q[2]; }
            }
        } else if (sp + 2 <= kBvhStack) {   // the depth of a median-split tree is far below the stack
            const double da = boxDist2(a.bvh[n.a], p), db = boxDist2(a.bvh[n.b], p);
            if (da <= db) { stack[sp++] = n.b; stack[sp++] = n.a; }
            else { stack[sp++] = n.a; stack[sp++] = n.b; }
        }
    }
    if (bf < 0) return;
    if (best < a.D[DRIW2][i]) {
        a.D[DRIW2][i] = best;
        a.D[DWPX][i] = bq[0]; a.D[DWPY][i] = bq[1]; a.D[DWPZ][i] = bq[2];
        a.D[DMX][i] = p[0] + 2.0 * (bq[0] - p[0]);
        a.D[DMY][i] = p[1] + 2.0 * (bq[1] - p[1]);
        a.D[DMZ][i] = p[2] + 2.0 * (bq[2] - p[2]);
        a.D[DPNX][i] = a.mN[bf * 3]; a.D[DPNY][i] = a.mN[bf * 3 + 1]; a.D[DPNZ][i] = a.mN[bf * 3 + 2];
        a.I[INEARMESH][i] = meshType;
    }
}

// PolygonMesh::interpolateWij / interpolateNumNeighWall (src/PolygonMesh.cpp:533-581)
__device__ __forceinline__ double interpZ(const double* xd, const double* yd, double re, double x) {
    const double r = x / re;
    int i = 0;
    if (r >= xd[kWallTable - 2]) i = kWallTable - 2;
    else while (r > xd[i + 1]) i++;
    const double xL = xd[i], yL = yd[i], xR = xd[i + 1], yR = yd[i + 1];
    const double dydx = (yR - yL) / (xR - xL);
    return yL + dydx * (r - xL);
}

// PolygonMesh::updateParticlesNearPolygonMesh (src/PolygonMesh.cpp:718-795)
__global__ void kUpdateNear(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    double niw = 0.0;
    int nnw = 0;
    const double x2 = a.D[DRIW2][i];
    if (x2 < P.re_l2) {
        a.I[INEARWALL][i] = 1;
        const double x = sqrt(x2);
        niw = interpZ(a.wtab, a.wtab + 2 * kWallTable, sqrt(P.re_s2), x);
        nnw = (int)interpZ(a.wtab + kWallTable, a.wtab + 3 * kWallTable, sqrt(P.re_l2), x);
    }
    a.D[DPNDWALL][i] = niw;
    a.I[INNWALL][i] = nnw;
}

// MpsPndNeigh::calcWallNPCD (src/MpsPndNeigh.cpp:105-195)
template <bool LIST>
__global__ void kWallNPCD(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (!(a.I[ITYPE][i] == P.fluid && a.I[INEARWALL][i])) return;
    const double xi[3] = {PX(i), PY(i), PZ(i)};
    const double xm[3] = {a.D[DMX][i], a.D[DMY][i], a.D[DMZ][i]};
    double n[3], R[9], ns;
    reflectSqrt(xi, xm, n, R, &ns);
    double nx = 0, ny = 0, nz = 0;
    forNbrsPosS<LIST>(a, i, xi[0], xi[1], xi[2], [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi[0], v1 = yj - xi[1], v2 = zj - xi[2];
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - xm[0], w1 = yj - xm[1], w2 = zj - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2) {
            const double wS = wgt(sqrt(dm2), P.re_s, P.weight_type);
            nx += w0 * wS; ny += w1 * wS; nz += w2 * wS;
        }
    });
    const double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
    const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
    if (dm2 < P.re_s2) {
        const double wS = wgt(sqrt(dm2), P.re_s, P.weight_type);
        nx += w0 * wS; ny += w1 * wS; nz += w2 * wS;
    }
    a.D[DNPX][i] = a.D[DNPX][i] + (R[0] * nx + R[1] * ny + R[2] * nz);
    a.D[DNPY][i] = a.D[DNPY][i] + (R[3] * nx + R[4] * ny + R[5] * nz);
    a.D[DNPZ][i] = a.D[DNPZ][i] + (R[6] * nx + R[7] * ny + R[8] * nz);
}

// ------------------------------------------------------ correction matrix
// MpsVectorMatrix::correctionMatrix + inverseMatrix (src/MpsVectorMatrix.cpp:20-165)
__device__ __forceinline__ void inverse3(double* M, int dim) {
    double A[9], B[9];
    for (int q = 0; q < 9; q++) A[q] = M[q];
    if (dim == 2) { A[2] = A[5] = A[6] = A[7] = 0.0; A[8] = 1.0; }
    for (int q = 0; q < 9; q++) B[q] = (q % 4 == 0) ? 1.0 : 0.0;
    const double det = (A[0] * A[4] * A[8] + A[1] * A[5] * A[6] + A[2] * A[3] * A[7]) -
                       (A[2] * A[4] * A[6] + A[1] * A[3] * A[8] + A[0] * A[5] * A[7]);
    if (det <= KEPS) {
        for (int q = 0; q < 9; q++) M[q] = (q % 4 == 0) ? 1.0 : 0.0;
        return;
    }
    for (int k = 0; k < dim; k++)
        for (int r = 0; r < dim; r++) {
            if (r == k) continue;
            const double m = A[r * 3 + k] / A[k * 3 + k];
            for (int c = 0; c < dim; c++) {
                A[r * 3 + c] -= m * A[k * 3 + c];
                B[r * 3 + c] -= m * B[k * 3 + c];
            }
        }
    for (int r = 0; r < dim; r++)
        for (int c = 0; c < dim; c++) B[r * 3 + c] /= A[r * 3 + r];
    for (int q = 0; q < 9; q++) M[q] = B[q];
}

template <bool LIST>
__global__ void kCorrectionMatrix(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
            const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
            const double inv = 1.0 / d2;
            M[0] += wS * v0 * v0 * inv; M[1] += wS * v0 * v1 * inv; M[2] += wS * v0 * v2 * inv;
            M[3] += wS * v1 * v0 * inv; M[4] += wS * v1 * v1 * inv; M[5] += wS * v1 * v2 * inv;
            M[6] += wS * v2 * v0 * inv; M[7] += wS * v2 * v1 * inv; M[8] += wS * v2 * v2 * inv;
        }
    });
    for (int q = 0; q < 9; q++) M[q] *= -P.coeff_press_grad;
    inverse3(M, (int)P.dim);
    if (a.I[INNEIGH][i] < 4)
        for (int q = 0; q < 9; q++) M[q] = (q % 4 == 0) ? 1.0 : 0.0;
    for (int q = 0; q < 9; q++) a.D[DCM0 + q][i] = M[q];
}

// ------------------------------------------------------------------ PND
// MpsPndNeigh::setInitialPndNumberOfNeighNPCD (src/MpsPndNeigh.cpp:21-101)
template <bool LIST>
__global__ void kInitialPnd(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double pnd = a.D[DPNDI][i], wSum = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
    int nn = 0;
    forNbrsPos<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_l2 && (d2 < dm2 || pw)) {
            nn += 1;
            if (d2 < P.re_s2) {
                const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
                pnd += wS;
                c0 += v0 * wS; c1 += v1 * wS; c2 += v2 * wS;
                wSum += wS;
            }
        }
    });
    pnd += a.D[DPNDWALL][i];
    if (a.I[ITYPE][i] == P.wall) pnd = P.pnd_small_zero;
    a.D[DPNDI][i] = pnd;
    a.D[DPNDSMALL][i] = pnd;
    a.D[DPNDKI][i] = pnd;
    a.I[INNEIGH][i] = nn + a.I[INNWALL][i];
    if (wSum > KEPS) { c0 /= pnd; c1 /= pnd; c2 /= pnd; }
    a.D[DNPX][i] = c0; a.D[DNPY][i] = c1; a.D[DNPZ][i] = c2;
    a.D[DNPCD2][i] = c0 * c0 + c1 * c1 + c2 * c2;
    a.D[DDEVN][i] = (c0 * a.D[DPNX][i] + c1 * a.D[DPNY][i] + c2 * a.D[DPNZ][i] < 0.0) ? 1.0 : -1.0;
}

// MpsPndNeigh::calcPndnNeighNPCD (src/MpsPndNeigh.cpp:199-331)
template <bool LIST>
__global__ void kPndNeighNPCD(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double ni = 0.0, wSum = 0.0;
    double c0 = a.D[DNPX][i], c1 = a.D[DNPY][i], c2 = a.D[DNPZ][i];
    int nn = a.I[INNWALL][i];
    forNbrsPos<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_l2 && (d2 < dm2 || pw)) {
            nn += 1;
            if (d2 < P.re_s2) {
                const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
                ni += wS;
                c0 += v0 * wS; c1 += v1 * wS; c2 += v2 * wS;
                wSum += wS;
            }
        }
    });
    a.I[INNEIGH][i] = nn;
    const double small = ni + a.D[DPNDWALL][i];
    if (P.pnd_type == MPS_PND_SUM_WIJ || P.pnd_type == MPS_PND_MEAN_SUM_WIJ) {
        a.D[DPNDKI][i] = a.D[DPNDI][i];
        a.D[DPNDI][i] = small;
    }
    a.D[DPNDSMALL][i] = small;
    if (wSum > KEPS) { c0 /= small; c1 /= small; c2 /= small; }
    a.D[DNPX][i] = c0; a.D[DNPY][i] = c1; a.D[DNPZ][i] = c2;
    a.D[DNPCD2][i] = c0 * c0 + c1 * c1 + c2 * c2;
    a.D[DDEVN][i] = (c0 * a.D[DPNX][i] + c1 * a.D[DPNY][i] + c2 * a.D[DPNZ][i] < 0.0) ? 1.0 : -1.0;
}

// calcPndDiffusiveTerm (src/MpsPndNeigh.cpp:334-489)
template <bool LIST>
__global__ void kPndDiffusive(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double ni = a.D[DPNDI][i];
    if (ni < KEPS) return;
    const double rho = rhoOf(a, i);
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    const double Pi = a.D[DPRESS][i];
    const bool fi = a.I[ITYPE][i] == P.fluid;
    double MC[9];
    for (int q = 0; q < 9; q++) MC[q] = a.D[DCM0 + q][i];
    double Di = 0.0, DivV = 0.0;
    const double flagDi = 1.0;
    forNbrsPos<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_l2 && (d2 < dm2 || pw)) {
            const double dst = sqrt(d2);
            const double wL = wgt(dst, P.re_l, P.weight_type);
            if (fi && a.I[ITYPE][j] == P.fluid) {
                const double rhoj = rhoOf(a, j);
                const double pgh = -2.0 * (rho * rhoj / (rho + rhoj)) *
                                   (P.gravity[0] * v0 + P.gravity[1] * v1 + P.gravity[2] * v2);
                const double Pj = a.D[DPRESS][j];
                Di += P.time_step / rho * P.coeff_visc_multiphase * (Pj - Pi - pgh) * wL;
            }
            if (d2 < P.re_s2) {
                const double vx = a.D[DU][j] - ux, vy = a.D[DV][j] - uy, vz = a.D[DW][j] - uz;
                const double wS = wgt(dst, P.re_s, P.weight_type);
                if (!P.gradient_correction) {
                    if (ni > KEPS)
                        DivV += (P.dim / P.pnd_small_zero) * (a.D[DPNDI][j] / ni) * (vx * v0 + vy * v1 + vz * v2) * wS / d2;
                } else {
                    const double q0 = (v0 * MC[0] + v1 * MC[1] + v2 * MC[2]);
                    const double q1 = (v0 * MC[3] + v1 * MC[4] + v2 * MC[5]);
                    const double q2 = (v0 * MC[6] + v1 * MC[7] + v2 * MC[8]);
                    DivV += (P.dim / P.pnd_small_zero) * (vx * q0 + vy * q1 + vz * q2) * wS / d2;
                }
            }
        }
    });
    a.D[DAUX1][i] = ni * (1.0 + P.time_step * (-DivV + Di * flagDi));
    a.D[DDIVU][i] = DivV;
    a.D[DDIFF][i] = Di;
}

// Velocity divergence over the mirror neighbourhood of near-wall fluid
// particle i (pndi = ni > eps), shared by calcWall{Slip,NoSlip}PndDiffusiveTerm
// (src/MpsPndNeigh.cpp:492-746) and calcWall{Slip,NoSlip}VelDivergence
// (src/MpsPressure.cpp:554-788), which form it identically.
template <bool LIST>
__device__ double wallDivV(const KArgs& a, int i, double ni) {
    const mps_params& P = a.P;
    const bool slip = P.slip_condition == MPS_FREE_SLIP;
    const double xi[3] = {PX(i), PY(i), PZ(i)};
    const double xm[3] = {a.D[DMX][i], a.D[DMY][i], a.D[DMZ][i]};
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    double n[3], R[9];
    reflectMod2(xi, xm, n, R);
    double umx, umy, umz;
    if (slip) {
        umx = R[0] * ux + R[1] * uy + R[2] * uz;
        umy = R[3] * ux + R[4] * uy + R[5] * uz;
        umz = R[6] * ux + R[7] * uy + R[8] * uz;
    } else {
        // wall velocity: uniformVelWall near the forced mesh, 0 otherwise
        const bool forced = a.I[INEARMESH][i] == 2;
        const double vw0 = forced ? P.uniform_vel_wall[0] : 0.0;
        const double vw1 = forced ? P.uniform_vel_wall[1] : 0.0;
        const double vw2 = forced ? P.uniform_vel_wall[2] : 0.0;
        const double dotnv = n[0] * vw0 + n[1] * vw1 + n[2] * vw2;
        const double t0 = ux - 2.0 * (vw0 - dotnv * n[0]);
        const double t1 = uy - 2.0 * (vw1 - dotnv * n[1]);
        const double t2 = uz - 2.0 * (vw2 - dotnv * n[2]);
        umx = -1.0 * t0 + 0.0 * t1 + 0.0 * t2;
        umy = 0.0 * t0 + -1.0 * t1 + 0.0 * t2;
        umz = 0.0 * t0 + 0.0 * t1 + -1.0 * t2;
    }
    double DivV = 0.0;
    auto term = [&](double vjx, double vjy, double vjz, double nj, double w0, double w1, double w2, double dm2) {
        const double wS = wgt(sqrt(dm2), P.re_s, P.weight_type);
        if (slip) {
            const double vx = vjx - umx, vy = vjy - umy, vz = vjz - umz;
            DivV += (P.dim / P.pnd_small_zero) * (nj / ni) * (vx * w0 + vy * w1 + vz * w2) * wS / dm2;
        } else {
            const double vx = -(vjx - umx), vy = -(vjy - umy), vz = -(vjz - umz);
            const double m0 = (R[0] * w0 + R[1] * w1 + R[2] * w2);
            const double m1 = (R[3] * w0 + R[4] * w1 + R[5] * w2);
            const double m2 = (R[6] * w0 + R[7] * w1 + R[8] * w2);
            DivV += (P.dim / P.pnd_small_zero) * (nj / ni) * (vx * m0 + vy * m1 + vz * m2) * wS / dm2;
        }
    };
    forNbrsPos<LIST>(a, i, xi[0], xi[1], xi[2], [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi[0], v1 = yj - xi[1], v2 = zj - xi[2];
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - xm[0], w1 = yj - xm[1], w2 = zj - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2)
            term(a.D[DU][j], a.D[DV][j], a.D[DW][j], a.D[DPNDI][j], w0, w1, w2, dm2);
    });
    const double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
    const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
    if (dm2 < P.re_s2) term(ux, uy, uz, ni, w0, w1, w2, dm2);
    return DivV;
}

// calcWall{Slip,NoSlip}PndDiffusiveTerm: pndAuxVar1 += -pndi dt DivV
template <bool LIST>
__global__ void kWallPndDiffusive(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const double ni = a.D[DPNDI][i];
    if (!(a.I[ITYPE][i] == P.fluid && a.I[INEARWALL][i] && ni > KEPS)) return;
    const double DivV = wallDivV<LIST>(a, i, ni);
    a.D[DAUX1][i] = a.D[DAUX1][i] + -ni * P.time_step * DivV;
}

// calcWall{Slip,NoSlip}VelDivergence (src/MpsPressure.cpp:554-788):
// velDivergence += DivV
template <bool LIST>
__global__ void kWallVelDivergence(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const double ni = a.D[DPNDI][i];
    if (!(a.I[ITYPE][i] == P.fluid && a.I[INEARWALL][i] && ni > KEPS)) return;
    a.D[DDIVU][i] = a.D[DDIVU][i] + wallDivV<LIST>(a, i, ni);
}

// updatePnd (src/MpsPndNeigh.cpp:747-774)
__global__ void kUpdatePnd(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    a.D[DPNDI][i] = a.D[DAUX1][i];
    a.D[DAUX1][i] = 0.0;
}

// meanPndParticlesWallDummySurface (src/MpsPndNeigh.cpp:777-852), first pass
template <bool LIST>
__global__ void kMeanPndWDS(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    a.D[DAUX1][i] = 0.0;
    a.D[DAUX2][i] = 0.0;
    if (!(a.I[ITYPE][i] == P.wall || a.I[IBC][i] == P.surface)) return;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double up = 0, dn = 0;
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
            const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
            up += a.D[DPNDI][j] * wS;
            dn += wS;
        }
    });
    a.D[DAUX1][i] = up;
    a.D[DAUX2][i] = dn;
}
__global__ void kMeanPndWDSFinish(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (a.I[ITYPE][i] == P.wall || a.I[IBC][i] == P.surface) {
        if (a.D[DAUX2][i] < KEPS) a.D[DPNDI][i] = a.D[DAUX1][i];
        else a.D[DPNDI][i] = a.D[DAUX1][i] / (a.D[DAUX2][i]);
    }
    a.D[DAUX1][i] = 0.0;
    a.D[DAUX2][i] = 0.0;
}

// meanWallPnd (src/MpsPndNeigh.cpp:923-994)
template <bool LIST>
__global__ void kMeanWallPnd(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    a.D[DAUX1][i] = 0.0;
    a.D[DAUX2][i] = 0.0;
    if (!a.I[INEARWALL][i]) return;
    const double xi[3] = {PX(i), PY(i), PZ(i)};
    const double xm[3] = {a.D[DMX][i], a.D[DMY][i], a.D[DMZ][i]};
    double up = 0, dn = 0;
    forNbrsPosS<LIST>(a, i, xi[0], xi[1], xi[2], [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi[0], v1 = yj - xi[1], v2 = zj - xi[2];
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - xm[0], w1 = yj - xm[1], w2 = zj - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2) {
            const double wS = wgt(sqrt(dm2), P.re_s, P.weight_type);
            up += a.D[DPNDI][j] * wS;
            dn += wS;
        }
    });
    const double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
    const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
    if (dm2 < P.re_s2) {
        const double wS = wgt(sqrt(dm2), P.re_s, P.weight_type);
        up += a.D[DPNDI][i] * wS;
        dn += wS;
    }
    a.D[DAUX1][i] = up;
    a.D[DAUX2][i] = dn;
}

// meanPnd (src/MpsPndNeigh.cpp:855-920)
template <bool LIST>
__global__ void kMeanPnd(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double up = 0, dn = 0;
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
            const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
            up += a.D[DPNDI][j] * wS;
            dn += wS;
        }
    });
    a.D[DAUX1][i] = a.D[DAUX1][i] + up;
    a.D[DAUX2][i] = a.D[DAUX2][i] + dn;
}
__global__ void kMeanPndFinish(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    if (a.I[INNEIGH][i] < 1) a.D[DPNDI][i] = a.D[DAUX1][i];
    else a.D[DPNDI][i] = a.D[DAUX1][i] / a.D[DAUX2][i];
    a.D[DAUX1][i] = 0.0;
    a.D[DAUX2][i] = 0.0;
}

// meanNeighFluidPnd (src/MpsPndNeigh.cpp:997-1050)
template <bool LIST>
__global__ void kMeanNeighFluidPnd(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double up = 0, dn = 0;
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw) && a.I[ITYPE][j] == P.fluid) {
            const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
            up += a.D[DPNDKI][j] * wS;
            dn += wS;
        }
    });
    a.D[DPNDSKI][i] = (dn > KEPS) ? up / dn : up;
}

// calcNormalParticles (src/MpsShifting.cpp:106-175): n_i = -sum_j r_ij w / |r_ij| / n0S
template <bool LIST>
__global__ void kNormalParticles(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double dx = 0.0, dy = 0.0, dz = 0.0;
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
            const double dst = sqrt(d2);
            const double wS = wgt(dst, P.re_s, P.weight_type);
            dx += v0 * wS / dst;
            dy += v1 * wS / dst;
            dz += v2 * wS / dst;
        }
    });
    a.D[DNRMX][i] = -dx / P.pnd_small_zero;
    a.D[DNRMY][i] = -dy / P.pnd_small_zero;
    a.D[DNRMZ][i] = -dz / P.pnd_small_zero;
}

// calcWallNormalParticles (src/MpsShifting.cpp:319-430): the mirror
// neighbourhood's part, reflected back by R
template <bool LIST>
__global__ void kWallNormalParticles(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (!(a.I[ITYPE][i] == P.fluid && a.I[INEARWALL][i])) return;
    const double xi[3] = {PX(i), PY(i), PZ(i)};
    const double xm[3] = {a.D[DMX][i], a.D[DMY][i], a.D[DMZ][i]};
    double n[3], R[9], ns;
    reflectSqrt(xi, xm, n, R, &ns);
    double dx = 0.0, dy = 0.0, dz = 0.0;
    forNbrsPosS<LIST>(a, i, xi[0], xi[1], xi[2], [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi[0], v1 = yj - xi[1], v2 = zj - xi[2];
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - xm[0], w1 = yj - xm[1], w2 = zj - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2) {
            const double dst = sqrt(dm2);
            const double wS = wgt(dst, P.re_s, P.weight_type);
            dx += w0 * wS / dst;
            dy += w1 * wS / dst;
            dz += w2 * wS / dst;
        }
    });
    const double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
    const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
    if (dm2 < P.re_s2) {
        const double dst = sqrt(dm2);
        const double wS = wgt(dst, P.re_s, P.weight_type);
        dx += w0 * wS / dst;
        dy += w1 * wS / dst;
        dz += w2 * wS / dst;
    }
    const double rx = R[0] * dx + R[1] * dy + R[2] * dz;
    const double ry = R[3] * dx + R[4] * dy + R[5] * dz;
    const double rz = R[6] * dx + R[7] * dy + R[8] * dz;
    a.D[DNRMX][i] = a.D[DNRMX][i] + -rx / P.pnd_small_zero;
    a.D[DNRMY][i] = a.D[DNRMY][i] + -ry / P.pnd_small_zero;
    a.D[DNRMZ][i] = a.D[DNRMZ][i] + -rz / P.pnd_small_zero;
}

// The ARC rule of updateParticleBC (src/MpsBoundaryCondition.cpp:205-392):
// in walk order, the first neighbour j (or j seen from the mirror point with
// polygon walls) that lies far but close to the point one l0 along the
// normal, or near within thetaArc of the normal, makes the particle inner and
// ends the walk; otherwise every neighbour examined leaves it at the surface.
template <bool LIST>
__device__ int arcRule(const KArgs& a, int i, int bc) {
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    double nx = a.D[DNRMX][i], ny = a.D[DNRMY][i], nz = a.D[DNRMZ][i];
    const double norm2 = nx * nx + ny * ny + nz * nz;
    if (a.D[DPNDI][i] >= P.beta_pnd && a.I[INNEIGH][i] >= P.beta_neigh && norm2 <= P.norm_threshold2) return P.inner;
    const double xi[3] = {PX(i), PY(i), PZ(i)};
    const double xm[3] = {a.D[DMX][i], a.D[DMY][i], a.D[DMZ][i]};
    if (norm2 > KEPS) {
        const double nn = sqrt(norm2);
        nx /= nn; ny /= nn; nz /= nn;
    } else {
        nx = 0.0; ny = 0.0; nz = 0.0;
    }
    double n[3], R[9];
    reflectMod2(xi, xm, n, R);
    double mx = R[0] * nx + R[1] * ny + R[2] * nz;
    double my = R[3] * nx + R[4] * ny + R[5] * nz;
    double mz = R[6] * nx + R[7] * ny + R[8] * nz;
    const double mn2 = mx * mx + my * my + mz * mz;
    if (mn2 > KEPS) {
        const double mn = sqrt(mn2);
        mx /= mn; my /= mn; mz /= mn;
    } else {
        mx = 0.0; my = 0.0; mz = 0.0;
    }
    const double l0 = P.part_dist;
    const double ax = xi[0] + l0 * nx, ay = xi[1] + l0 * ny, az = xi[2] + l0 * nz;
    const double bx = xm[0] + l0 * mx, by = xm[1] + l0 * my, bz = xm[2] + l0 * mz;
    bool done = false;
    forNbrsPos<LIST>(a, i, xi[0], xi[1], xi[2], [&](int j, double xj, double yj, double zj) {
        if (done) return;
        const double v0 = xj - xi[0], v1 = yj - xi[1], v2 = zj - xi[2];
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - xm[0], w1 = yj - xm[1], w2 = zj - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
            const double u0 = xj - ax, u1 = yj - ay, u2 = zj - az;
            const double dn2 = u0 * u0 + u1 * u1 + u2 * u2;
            const double rn = v0 * nx + v1 * ny + v2 * nz;
            const double ang = acos(rn / sqrt(d2));
            if (d2 >= P.dst_threshold2 && dn2 < P.h_threshold2) { bc = P.inner; done = true; return; }
            else if (d2 < P.dst_threshold2 && ang < P.theta_arc) { bc = P.inner; done = true; return; }
            else bc = P.surface;
        }
        if (!pw && d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2) {
            const double u0 = xj - bx, u1 = yj - by, u2 = zj - bz;
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

// MpsBoundaryCondition::updateParticleBC (src/MpsBoundaryCondition.cpp:66-400)
template <bool LIST>
__global__ void kUpdateBC(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (a.I[ITYPE][i] == P.dummy_wall) { a.I[IBC][i] = P.other; return; }
    const double small = a.D[DPNDSMALL][i];
    const int nn = a.I[INNEIGH][i];
    int bc = (small < P.beta_pnd) ? P.surface : P.inner;
    if (P.free_surf_type == MPS_FS_PND_NEIGH) {
        bc = (small < P.beta_pnd && nn < P.beta_neigh) ? P.surface : P.inner;
    } else if (P.free_surf_type == MPS_FS_PND_NPCD) {
        if (bc == P.surface && nn > 4 && a.D[DNPCD2][i] < P.delta2) bc = P.inner;
    } else if (P.free_surf_type == MPS_FS_PND_ARC) {
        bc = arcRule<LIST>(a, i, bc);
    }
    a.I[IBC][i] = bc;
}

// --------------------------------------------------------------- pressure
// calcPressEMPS / calcPressWCMPS (src/MpsPressure.cpp:62-163)
__global__ void kPressExplicit(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const double mi = phaseDensity(a, i);
    double pr = 0.0;
    if (a.I[IBC][i] == P.inner) {
        if (P.mps_type == MPS_EXPLICIT) pr = (a.D[DPNDI][i] - P.pnd_small_zero) * P.coeff_press_emps * mi;
        else pr = (mi * P.coeff_press_wcmps / P.gamma) * (pow(a.D[DPNDI][i] / P.pnd_small_zero, P.gamma) - 1);
    }
    if (pr < 0.0) pr = 0.0;
    a.D[DPRESS][i] = pr;
}

// calcVelDivergence (src/MpsPressure.cpp:479-551)
template <bool LIST>
__global__ void kVelDivergence(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    a.D[DDIVU][i] = 0.0;
    if (a.I[ITYPE][i] != P.fluid) return;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double ni = a.D[DPNDI][i];
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double MC[9];
    for (int q = 0; q < 9; q++) MC[q] = a.D[DCM0 + q][i];
    double DivV = 0.0;
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
            const double vx = a.D[DU][j] - ux, vy = a.D[DV][j] - uy, vz = a.D[DW][j] - uz;
            const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
            if (!P.gradient_correction) {
                if (ni > KEPS)
                    DivV += (P.dim / P.pnd_small_zero) * (a.D[DPNDI][j] / ni) * (vx * v0 + vy * v1 + vz * v2) * wS / d2;
            } else {
                const double q0 = (v0 * MC[0] + v1 * MC[1] + v2 * MC[2]);
                const double q1 = (v0 * MC[3] + v1 * MC[4] + v2 * MC[5]);
                const double q2 = (v0 * MC[6] + v1 * MC[7] + v2 * MC[8]);
                DivV += (P.dim / P.pnd_small_zero) * (vx * q0 + vy * q1 + vz * q2) * wS / d2;
            }
        }
    });
    a.D[DDIVU][i] = DivV;
}

// calcPressGradient (src/MpsPressure.cpp:1217-1338)
// calcPressGradient (src/MpsPressure.cpp:1217-1338): acc = relax grad p / rho,
// and with PRED predictionPressGradient (:932-1028): acc += (1 - relax)
// grad p / rho without the gradient correction
template <bool LIST, bool PRED>
__device__ __forceinline__ void pressGradientBody(const KArgs& a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const int gt = P.gradient_type;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    const double Pi = a.D[DPRESS][i], ni = a.D[DPNDI][i];
    double pmin = Pi;
    double MC[9];
    for (int q = 0; q < 9; q++) MC[q] = PRED ? 0.0 : a.D[DCM0 + q][i];
    const double* __restrict__ PR = a.D[DPRESS];
    if (gt == 0 || gt == 2)
        forNbrsPosSX<LIST>(a, i, xi, yi, zi, PR, [&](int j, double xj, double yj, double zj, auto pOf) {
            const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
                const double pj = pOf();
                if (pmin > pj) pmin = pj;
            }
        });
    double ax = 0, ay = 0, az = 0;
    forNbrsPosSX<LIST>(a, i, xi, yi, zi, PR, [&](int j, double xj, double yj, double zj, auto pOf) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
            double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
            const double pj = pOf();
            if (gt == 0) wS *= (pj - pmin) / d2;
            else if (gt == 1) wS *= (pj + Pi) / d2;
            else if (gt == 2) wS *= (pj + Pi - 2.0 * pmin) / d2;
            else if (gt == 3) {
                const double nj = a.D[DPNDI][j];
                if (ni > KEPS && nj > KEPS) wS *= (ni * pj / nj + nj * Pi / ni) / d2;
            }
            if (PRED || !P.gradient_correction) {
                ax += v0 * wS; ay += v1 * wS; az += v2 * wS;
            } else {
                ax += (v0 * MC[0] + v1 * MC[1] + v2 * MC[2]) * wS;
                ay += (v0 * MC[3] + v1 * MC[4] + v2 * MC[5]) * wS;
                az += (v0 * MC[6] + v1 * MC[7] + v2 * MC[8]) * wS;
            }
        }
    });
    const double rho = rhoOf(a, i);
    if (PRED) {
        const double f = 1.0 - P.relax_press;
        a.D[DAX][i] = a.D[DAX][i] + f * ax * P.coeff_press_grad / rho;
        a.D[DAY][i] = a.D[DAY][i] + f * ay * P.coeff_press_grad / rho;
        a.D[DAZ][i] = a.D[DAZ][i] + f * az * P.coeff_press_grad / rho;
    } else {
        a.D[DAX][i] = P.relax_press * ax * P.coeff_press_grad / rho;
        a.D[DAY][i] = P.relax_press * ay * P.coeff_press_grad / rho;
        a.D[DAZ][i] = P.relax_press * az * P.coeff_press_grad / rho;
    }
}
template <bool LIST>
__global__ void kPressGradient(KArgs a) { pressGradientBody<LIST, false>(a); }
template <bool LIST>
__global__ void kPredPressGradient(KArgs a) { pressGradientBody<LIST, true>(a); }

// calcWallPressGradient + repulsive forces (src/MpsPressure.cpp:1341-1611);
// with PRED predictionWallPressGradient (:1031-1214): factor 1 - relax and
// no wall force
template <bool LIST, bool PRED>
__device__ __forceinline__ void wallPressGradientBody(const KArgs& a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (!(a.I[ITYPE][i] == P.fluid && a.I[INEARWALL][i])) return;
    const int gt = P.gradient_type;
    const double vol = a.volForce;
    const double vm = fmin(a.ds[S_VELMAX], P.expect_max_velocity);
    const double vmax2 = vm * vm;
    const double rho = rhoOf(a, i);
    const double xi[3] = {PX(i), PY(i), PZ(i)};
    const double xm[3] = {a.D[DMX][i], a.D[DMY][i], a.D[DMZ][i]};
    const double Pi = a.D[DPRESS][i], ni = a.D[DPNDI][i];
    double pmin = Pi;
    double n[3], R[9], ns;
    reflectSqrt(xi, xm, n, R, &ns);
    if (gt == 0 || gt == 2)
        forNbrsPos<LIST>(a, i, xi[0], xi[1], xi[2], [&](int j, double xj, double yj, double zj) {
            const double v0 = xj - xi[0], v1 = yj - xi[1], v2 = zj - xi[2];
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = xj - xm[0], w1 = yj - xm[1], w2 = zj - xm[2];
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && d2 < dm2) {
                const double pj = a.D[DPRESS][j];
                if (pmin > pj) pmin = pj;
            }
        });
    double ax = 0, ay = 0, az = 0;
    auto term = [&](double Pj, double nj, double w0, double w1, double w2, double dm2) {
        double wS = wgt(sqrt(dm2), P.re_s, P.weight_type);
        if (gt == 0) wS *= (Pj - pmin) / dm2;
        else if (gt == 1) wS *= (Pj + Pi) / dm2;
        else if (gt == 2) wS *= (Pj + Pi - 2.0 * pmin) / dm2;
        else if (gt == 3) {
            if (ni > KEPS && nj > KEPS) wS *= (ni * Pj / nj + nj * Pi / ni) / dm2;
        }
        ax += w0 * wS; ay += w1 * wS; az += w2 * wS;
    };
    forNbrsPos<LIST>(a, i, xi[0], xi[1], xi[2], [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi[0], v1 = yj - xi[1], v2 = zj - xi[2];
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - xm[0], w1 = yj - xm[1], w2 = zj - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2) term(a.D[DPRESS][j], a.D[DPNDI][j], w0, w1, w2, dm2);
    });
    {
        const double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (dm2 < P.re_s2) term(Pi, ni, w0, w1, w2, dm2);
    }
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    if (ns < P.re_repulsive_force && ns > KEPS) {
        double w;
        if (P.repulsive_force_type == MPS_REP_HARADA) {
            w = rho / (P.time_step * P.time_step) * (P.re_repulsive_force - ns);
        } else if (P.repulsive_force_type == MPS_REP_MITSUME) {
            w = P.rep_coef_mitsume * wgt(ns, P.re_repulsive_force, P.weight_type);
        } else if (P.repulsive_force_type == MPS_REP_LENNARD_JONES) {
            const double R1 = (P.re_repulsive_force / ns) * (P.re_repulsive_force / ns);
            const double R2 = R1 * R1;
            w = (P.rep_coef_lennard_jones * vmax2 / ns) * (R2 - R1) * rho;
        } else {
            const double W1 = (1.0 + 3.0 * 0.5 * ns / (P.re_repulsive_force));
            const double W2 = (1.0 - ns / (P.re_repulsive_force)) * (1.0 - ns / (P.re_repulsive_force)) *
                              (1.0 - ns / (P.re_repulsive_force));
            w = (P.rep_coef_monaghan_kajtar * vmax2 / (ns - 0.0 * P.part_dist)) * (1.0 / 8.0) * (W1) * (W2) * rho;
        }
        f0 = -w * n[0]; f1 = -w * n[1]; f2 = -w * n[2];
    }
    const double fr = PRED ? (1.0 - P.relax_press) : P.relax_press;
    const double g0 = fr * (R[0] * ax + R[1] * ay + R[2] * az) * P.coeff_press_grad - f0;
    const double g1 = fr * (R[3] * ax + R[4] * ay + R[5] * az) * P.coeff_press_grad - f1;
    const double g2 = fr * (R[6] * ax + R[7] * ay + R[8] * az) * P.coeff_press_grad - f2;
    a.D[DAX][i] = a.D[DAX][i] + g0 / rho;
    a.D[DAY][i] = a.D[DAY][i] + g1 / rho;
    a.D[DAZ][i] = a.D[DAZ][i] + g2 / rho;
    if (!PRED) {
        a.D[DFWX][i] = a.D[DFWX][i] + -g0 * vol;
        a.D[DFWY][i] = a.D[DFWY][i] + -g1 * vol;
        a.D[DFWZ][i] = a.D[DFWZ][i] + -g2 * vol;
    }
}
template <bool LIST>
__global__ void kWallPressGradient(KArgs a) { wallPressGradientBody<LIST, false>(a); }
template <bool LIST>
__global__ void kPredWallPressGradient(KArgs a) { wallPressGradientBody<LIST, true>(a); }

// calcWall{Slip,NoSlip}Viscosity (src/MpsViscosity.cpp:1527-1910)
template <bool LIST>
__global__ void kWallViscosity(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (!(a.I[ITYPE][i] == P.fluid && a.I[INEARWALL][i])) return;
    const bool slip = P.slip_condition == MPS_FREE_SLIP;
    const double vol = a.volForce;
    const double meu = meuOf(a, i), rho = rhoOf(a, i), dnsI = phaseDensity(a, i);
    const bool nn = P.fluid_type == 1;   // per-neighbour viscosity of the granular phase
    const double xi[3] = {PX(i), PY(i), PZ(i)};
    const double xm[3] = {a.D[DMX][i], a.D[DMY][i], a.D[DMZ][i]};
    const double ux = a.D[DU][i], uy = a.D[DV][i], uz = a.D[DW][i];
    double n[3], R[9];
    reflectMod2(xi, xm, n, R);
    double umx, umy, umz, Rm[9];
    if (slip) {
        umx = R[0] * ux + R[1] * uy + R[2] * uz;
        umy = R[3] * ux + R[4] * uy + R[5] * uz;
        umz = R[6] * ux + R[7] * uy + R[8] * uz;
        for (int q = 0; q < 9; q++) Rm[q] = R[q];
    } else {
        // wall velocity: uniformVelWall near the forced mesh, 0 otherwise
        const bool forced = a.I[INEARMESH][i] == 2;
        const double vw0 = forced ? P.uniform_vel_wall[0] : 0.0;
        const double vw1 = forced ? P.uniform_vel_wall[1] : 0.0;
        const double vw2 = forced ? P.uniform_vel_wall[2] : 0.0;
        const double dotnv = n[0] * vw0 + n[1] * vw1 + n[2] * vw2;
        const double t0 = ux - 2.0 * (vw0 - dotnv * n[0]);
        const double t1 = uy - 2.0 * (vw1 - dotnv * n[1]);
        const double t2 = uz - 2.0 * (vw2 - dotnv * n[2]);
        umx = -1.0 * t0 + 0.0 * t1 + 0.0 * t2;
        umy = 0.0 * t0 + -1.0 * t1 + 0.0 * t2;
        umz = 0.0 * t0 + 0.0 * t1 + -1.0 * t2;
        for (int q = 0; q < 9; q++) Rm[q] = (q % 4 == 0) ? -1.0 : 0.0;
    }
    double nuJ;
    if ((meu + meu) > KEPS) nuJ = 2.0 * meu * meu / (meu + meu);
    else nuJ = 0.0;
    nuJ = nuJ / dnsI;
    double nuI;
    if (meu > KEPS) nuI = 2.0 * meu * meu / (meu + meu);
    else nuI = 0.0;
    nuI = nuI / dnsI;
    double ax = 0, ay = 0, az = 0;
    forNbrsPos<LIST>(a, i, xi[0], xi[1], xi[2], [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi[0], v1 = yj - xi[1], v2 = zj - xi[2];
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - xm[0], w1 = yj - xm[1], w2 = zj - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_l2 && dm2 < P.re_l2 && d2 < dm2) {
            const double wL = wgt(sqrt(dm2), P.re_l, P.weight_type);
            double nu = nuJ;
            if (nn) {
                const double mj = a.D[DMEU][j];
                if ((meu + mj) > KEPS) nu = 2.0 * meu * mj / (meu + mj);
                else nu = 0.0;
                nu = nu / dnsI;
            }
            ax += (a.D[DU][j] - umx) * wL * nu;
            ay += (a.D[DV][j] - umy) * wL * nu;
            az += (a.D[DW][j] - umz) * wL * nu;
        }
    });
    {
        const double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (dm2 < P.re_l2) {
            const double wL = wgt(sqrt(dm2), P.re_l, P.weight_type);
            ax += (ux - umx) * wL * nuI;
            ay += (uy - umy) * wL * nuI;
            az += (uz - umz) * wL * nuI;
        }
    }
    const double r0 = (Rm[0] * ax + Rm[1] * ay + Rm[2] * az) * P.coeff_visc_multiphase;
    const double r1 = (Rm[3] * ax + Rm[4] * ay + Rm[5] * az) * P.coeff_visc_multiphase;
    const double r2 = (Rm[6] * ax + Rm[7] * ay + Rm[8] * az) * P.coeff_visc_multiphase;
    a.D[DAX][i] = a.D[DAX][i] + r0;
    a.D[DAY][i] = a.D[DAY][i] + r1;
    a.D[DAZ][i] = a.D[DAZ][i] + r2;
    a.D[DFWX][i] = a.D[DFWX][i] + -(Rm[0] * ax + Rm[1] * ay + Rm[2] * az) * P.coeff_visc_multiphase * vol * rho;
    a.D[DFWY][i] = a.D[DFWY][i] + -(Rm[3] * ax + Rm[4] * ay + Rm[5] * az) * P.coeff_visc_multiphase * vol * rho;
    a.D[DFWZ][i] = a.D[DFWZ][i] + -(Rm[6] * ax + Rm[7] * ay + Rm[8] * az) * P.coeff_visc_multiphase * vol * rho;
}

// ------------------------------------------------------------- shifting
// calcShifting (src/MpsShifting.cpp:23-104), read-old / write-new: the new
// velocity is staged in the dvel slots, then copied.
template <bool LIST>
__global__ void kShift1(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    double nu = a.D[DU][i], nv = a.D[DV][i], nw = a.D[DW][i];
    if (a.I[ITYPE][i] == P.fluid) {
        const bool pw = P.wall_type == MPS_WALL_PARTICLE;
        const double xi = PX(i), yi = PY(i), zi = PZ(i);
        const double ux = nu, uy = nv, uz = nw;
        const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
        double dx = 0, dy = 0, dz = 0;
        forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
            const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
                const double dw = delWgt(sqrt(d2), P.re_s, P.weight_type);
                dx += dw * (a.D[DU][j] - ux);
                dy += dw * (a.D[DV][j] - uy);
                dz += dw * (a.D[DW][j] - uz);
            }
        });
        if (a.I[IBC][i] == P.inner) {
            nu = nu - P.coeff_shifting1 * dx;
            nv = nv - P.coeff_shifting1 * dy;
            nw = nw - P.coeff_shifting1 * dz;
        }
    }
    a.D[DDVX][i] = nu; a.D[DDVY][i] = nv; a.D[DDVZ][i] = nw;
}
// copy staged values into (dst x,y,z) and clear the staging slots
__global__ void kCommitStaged(KArgs a, int dst0) {
    const int i = slotId();
    if (!live(a, i)) return;
    a.D[dst0][i] = a.D[DDVX][i];
    a.D[dst0 + 1][i] = a.D[DDVY][i];
    a.D[dst0 + 2][i] = a.D[DDVZ][i];
    a.D[DDVX][i] = 0.0; a.D[DDVY][i] = 0.0; a.D[DDVZ][i] = 0.0;
}

// calcConcAndConcGradient (src/MpsShifting.cpp:432-576), concentration pass
template <bool LIST>
__global__ void kShift2Conc(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const bool pw = P.wall_type == MPS_WALL_PARTICLE;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
    double c = 0.0;
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (d2 < P.re_s2 && (d2 < dm2 || pw)) c += wgt(sqrt(d2), P.re_s, P.weight_type);
    });
    c += a.D[DPNDWALL][i];
    a.D[DCONC][i] = c / P.pnd_small_zero;
}
// gradient pass: new positions staged in dvel slots
template <bool LIST>
__global__ void kShift2Grad(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    double nx = PX(i), ny = PY(i), nz = PZ(i);
    if (a.I[ITYPE][i] == P.fluid) {
        const bool pw = P.wall_type == MPS_WALL_PARTICLE;
        const double xi = nx, yi = ny, zi = nz;
        const double mx = a.D[DMX][i], my = a.D[DMY][i], mz = a.D[DMZ][i];
        const double ci = a.D[DCONC][i];
        double g0 = 0, g1 = 0, g2 = 0;
        forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
            const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = xj - mx, w1 = yj - my, w2 = zj - mz;
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && (d2 < dm2 || pw)) {
                const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
                const double cj = a.D[DCONC][j];
                g0 += (ci + cj) * v0 * wS / d2;
                g1 += (ci + cj) * v1 * wS / d2;
                g2 += (ci + cj) * v2 * wS / d2;
            }
        });
        g0 *= -P.coeff_press_grad; g1 *= -P.coeff_press_grad; g2 *= -P.coeff_press_grad;
        a.D[DGCX][i] = g0; a.D[DGCY][i] = g1; a.D[DGCZ][i] = g2;
        if (a.I[IBC][i] == P.inner) {
            nx = nx - P.coeff_shifting2 * g0;
            ny = ny - P.coeff_shifting2 * g1;
            nz = nz - P.coeff_shifting2 * g2;
        }
    }
    a.D[DDVX][i] = nx; a.D[DDVY][i] = ny; a.D[DDVZ][i] = nz;
}
// calcWallConcAndConcGradient (src/MpsShifting.cpp:578-692)
template <bool LIST>
__global__ void kShift2Wall(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    double nx = PX(i), ny = PY(i), nz = PZ(i);
    if (a.I[ITYPE][i] == P.fluid && a.I[INEARWALL][i]) {
        const double xi[3] = {nx, ny, nz};
        const double xm[3] = {a.D[DMX][i], a.D[DMY][i], a.D[DMZ][i]};
        const double ci = a.D[DCONC][i];
        double n[3], R[9], ns;
        reflectSqrt(xi, xm, n, R, &ns);
        double d0 = 0, d1 = 0, dd2 = 0;
        forNbrsPosS<LIST>(a, i, xi[0], xi[1], xi[2], [&](int j, double xj, double yj, double zj) {
            const double v0 = xj - xi[0], v1 = yj - xi[1], v2 = zj - xi[2];
            const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
            const double w0 = xj - xm[0], w1 = yj - xm[1], w2 = zj - xm[2];
            const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
            if (d2 < P.re_s2 && dm2 < P.re_s2 && d2 < dm2) {
                const double wS = wgt(sqrt(dm2), P.re_s, P.weight_type);
                const double cj = a.D[DCONC][j];
                d0 += (ci + cj) * w0 * wS / dm2;
                d1 += (ci + cj) * w1 * wS / dm2;
                dd2 += (ci + cj) * w2 * wS / dm2;
            }
        });
        const double w0 = xi[0] - xm[0], w1 = xi[1] - xm[1], w2 = xi[2] - xm[2];
        const double dm2 = w0 * w0 + w1 * w1 + w2 * w2;
        if (dm2 < P.re_s2) {
            const double wS = wgt(sqrt(dm2), P.re_s, P.weight_type);
            d0 += (ci + ci) * w0 * wS / dm2;
            d1 += (ci + ci) * w1 * wS / dm2;
            dd2 += (ci + ci) * w2 * wS / dm2;
        }
        const double gx = -(R[0] * d0 + R[1] * d1 + R[2] * dd2) * P.coeff_press_grad;
        const double gy = -(R[3] * d0 + R[4] * d1 + R[5] * dd2) * P.coeff_press_grad;
        const double gz = -(R[6] * d0 + R[7] * d1 + R[8] * dd2) * P.coeff_press_grad;
        a.D[DGCX][i] = a.D[DGCX][i] + gx;
        a.D[DGCY][i] = a.D[DGCY][i] + gy;
        a.D[DGCZ][i] = a.D[DGCZ][i] + gz;
        if (a.I[IBC][i] == P.inner) {
            nx = nx - P.coeff_shifting2 * gx;
            ny = ny - P.coeff_shifting2 * gy;
            nz = nz - P.coeff_shifting2 * gz;
        }
    }
    a.D[DDVX][i] = nx; a.D[DDVY][i] = ny; a.D[DDVZ][i] = nz;
}

// updateVelocityParticlesWallDummy (src/MpsParticleVelPos.cpp:100-158)
template <bool LIST>
__global__ void kWallDummyVelocity(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    if (a.I[ITYPE][i] != P.dummy_wall) return;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    double dx = 0, dy = 0, dz = 0, ni = 0;
    forNbrsPos<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        if (d2 < P.re_l2 && a.I[ITYPE][j] == P.fluid) {
            const double wL = wgt(sqrt(d2), P.re_l, P.weight_type);
            ni += wL;
            dx += a.D[DU][j] * wL;
            dy += a.D[DV][j] * wL;
            dz += a.D[DW][j] * wL;
        }
    });
    if (ni > KEPS) {
        a.D[DU][i] = 2.0 * 0.0 - dx / ni;
        a.D[DV][i] = 2.0 * 0.0 - dy / ni;
        a.D[DW][i] = 2.0 * 0.0 - dz / ni;
    } else {
        a.D[DU][i] = 2.0 * 0.0 - dx;
        a.D[DV][i] = 2.0 * 0.0 - dy;
        a.D[DW][i] = 2.0 * 0.0 - dz;
    }
}

// ---------------------------------------------------------------- launch
#define L1(K) hipLaunchKernelGGL(K, dim3(g), dim3(kBlock), 0, ctx->stream, a)
#define LT(K)                                                                       \
    do {                                                                            \
        if (ctx->listValid) hipLaunchKernelGGL(K<true>, dim3(g), dim3(kBlock), 0, ctx->stream, a); \
        else hipLaunchKernelGGL(K<false>, dim3(g), dim3(kBlock), 0, ctx->stream, a); \
    } while (0)
#define LW(K) hipLaunchKernelGGL(K<false>, dim3(g), dim3(kBlock), 0, ctx->stream, a)

static double listRadius2(const mps_params& P) {
    double r2 = std::max(P.re_l2, P.re_s2);
    r2 = std::max(r2, P.dist_collision_limit2);
    return std::max(r2, P.part_dist * P.part_dist);
}
static double listRadiusS2(const mps_params& P) { return std::max(P.re_s2, P.dist_collision_limit2); }

// P4.w = type & 1023 | (bc == inner) << 10 | (riw2 >= 4 reL^2) << 11: the two
// neighbour properties the PPE assembly tests, gathered with the position.
__global__ void kPackP4Flags(KArgs a) {
    const int i = slotId();
    if (i >= a.N) return;
    const mps_params& P = a.P;
    const int t = a.I[ITYPE][i];
    const int inner = a.I[IBC][i] == P.inner ? 1 : 0;
    const int far = a.D[DRIW2][i] >= 4.0 * P.re_l2 ? 1 : 0;
    a.P4[i].w = (double)((t & 1023) | (inner << 10) | (far << 11));
}

int launchBuildList(Ctx* ctx, double skin, const int* gate, bool small) {
    ctx->p4Flags = false;
    KArgs a = ctx->args();
    const int g = gridFor(ctx->N, kBlock);
    const mps_params& P = ctx->P;
    double r2 = listRadius2(P), r2s = listRadiusS2(P);
    if (skin > 0.0) {
        r2 = (std::sqrt(r2) + skin) * (std::sqrt(r2) + skin);
        r2s = (std::sqrt(r2s) + skin) * (std::sqrt(r2s) + skin);
    }
    if (P.domain_type_bc == 1) {
        hipLaunchKernelGGL(kBuildListWalk, dim3(g), dim3(kBlock), 0, ctx->stream, a, ctx->nbr, ctx->nbrCnt, r2);
        MPS_HIP(hipGetLastError());
        ctx->listValid = true;
        return MPS_OK;
    }
    const int gw = gridFor(ctx->N, 64);
    if (!ctx->fixList) MPS_HIP(hipMalloc(&ctx->fixList, sizeof(int) * (size_t)ctx->N));
    MPS_HIP(hipMemsetAsync(ctx->is + I_FIXCNT, 0, sizeof(int), ctx->stream));
    if (small)
        hipLaunchKernelGGL(kBuildList<true>, dim3(gridFor(ctx->N, kWave)), dim3(kWave), 0, ctx->stream, a,
                           ctx->nbrMask, ctx->nbrCnt, ctx->nbrMaskS, ctx->nbrCntS, r2, r2s, gate, ctx->fixList);
    else
        hipLaunchKernelGGL(kBuildList<false>, dim3(gridFor(ctx->N, kWave)), dim3(kWave), 0, ctx->stream, a,
                           ctx->nbrMask, ctx->nbrCnt, ctx->nbrMaskS, ctx->nbrCntS, r2, r2s, gate, ctx->fixList);
#define M2L(WM)                                                                                              \
    do {                                                                                                     \
        hipLaunchKernelGGL(kMaskToList<WM>, dim3(gw), dim3(64), 0, ctx->stream, a, ctx->nbrMask, ctx->nbrCnt,  \
                           ctx->nbr, gate);                                                                  \
        if (small)                                                                                           \
            hipLaunchKernelGGL(kMaskToList<WM>, dim3(gw), dim3(64), 0, ctx->stream, a, ctx->nbrMaskS,         \
                               ctx->nbrCntS, ctx->nbrS, gate);                                               \
    } while (0)
    if (ctx->maskW <= 4) M2L(4);
    else if (ctx->maskW <= 8) M2L(8);
    else if (ctx->maskW <= 16) M2L(16);
    else {
        ctx->err = "neighbour bitmask rows wider than 512 candidates are not supported";
        return MPS_E_UNSUPPORTED;
    }
#undef M2L
    hipLaunchKernelGGL(kFixList, dim3(g), dim3(kBlock), 0, ctx->stream, a, ctx->nbr, ctx->nbrCnt, ctx->nbrS,
                       ctx->nbrCntS, ctx->fixList, r2, r2s, gate);
    MPS_HIP(hipGetLastError());
    ctx->listValid = true;
    return MPS_OK;
}

// ------------------------------------------------- list kept over the step
// The list of a time step can be built once, right after the bucket sort, at
// the radius plus a skin, and serve the viscosity at the old positions as well
// as every operator after the prediction:
//  * every consumer tests its own radius on the positions it is given, so a
//    superset of the neighbours, in the same slot order, gives bit-identical
//    sums;
//  * candidates come from the stencil of the particle's current bucket over
//    the buckets of the sort (src/MpsBucket.cpp:299-353 runs once per step),
//    so a list stays a superset while neither end of a pair moved more than
//    half the skin and the particle kept its stencil centre.
// After the prediction kListRefresh copies the new positions into P4, raises
// the rebuild flag when a particle moved farther (or when the compaction moved
// particles between slots) and queues the particles whose stencil centre
// changed; the full rebuild then runs only under the flag, and kPatchList
// rebuilds the queued particles' lists alone.
__global__ void kListFlagsReset(int* is) {
    is[I_REBUILD] = 0;
    is[I_PATCHCNT] = 0;
}

__device__ __forceinline__ int3 stencilCentre(const mps_params& P, double x, double y, double z) {
    return make_int3((int)((x - P.domain_min[0]) * P.inv_bucket_side + KEPS),
                     (int)((y - P.domain_min[1]) * P.inv_bucket_side + KEPS),
                     ((int)P.dim == 3) ? (int)((z - P.domain_min[2]) * P.inv_bucket_side + KEPS) : 0);
}

__global__ __launch_bounds__(256) void kListRefresh(KArgs a, int* patch, double half2) {
    const int i = slotId();
    if (i >= a.N) return;
    const double x = PX(i), y = PY(i), z = PZ(i);
    const double4 q = a.P4[i];
    a.P4[i] = make_double4(x, y, z, (double)a.I[ITYPE][i]);
    if (!live(a, i)) return;
    const double dx = x - q.x, dy = y - q.y, dz = z - q.z;
    if (dx * dx + dy * dy + dz * dz > half2) a.is[I_REBUILD] = 1;
    const int3 c1 = stencilCentre(a.P, x, y, z), c0 = stencilCentre(a.P, q.x, q.y, q.z);
    if (c1.x != c0.x || c1.y != c0.y || c1.z != c0.z) {
        const int s = atomicAdd(&a.is[I_PATCHCNT], 1);
        patch[s] = i;
    }
}

// The list of each queued particle from its own stencil walk, in walk order
// (the order of kBuildList + kMaskToList).
__global__ __launch_bounds__(256) void kPatchList(KArgs a, int* nbr, int* cnt, int* nbrS, int* cntS, const int* patch,
                                                  double r2, double r2s) {
    if (a.is[I_REBUILD]) return;   // the whole list was rebuilt
    const int t = slotId();
    if (t >= a.is[I_PATCHCNT]) return;
    const int i = patch[t];
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    const size_t N = a.N;
    int c = 0, cs = 0;
    walkCellsPos(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        if (!(d2 < r2)) return;
        if (c < a.K) nbr[(size_t)c * N + i] = j;
        c++;
        if (d2 < r2s) {
            if (cs < a.KS) nbrS[(size_t)cs * N + i] = j;
            cs++;
        }
    });
    if (c > a.K) {
        atomicMax(&a.is[I_OVERFLOW], c);
        c = a.K;
    }
    if (cs > a.KS) {
        atomicMax(&a.is[I_OVERFLOW], cs);
        cs = a.KS;
    }
    cnt[i] = c;
    cntS[i] = cs;
}

// Drops from a kept row the entries that lie outside the list radius at the
// new positions, in place and in order (entry k is written only after entries
// 0..k were read).  The result is the list kBuildList would build now: the
// kept candidates of a particle that kept its stencil centre are the
// candidates of that stencil, and kPatchList rows are already tight.  Without
// it the consumers after the prediction walk the skin's extra entries, whose
// radius test leaves the lanes of a wave at different entry counts, so the
// PPE assembly's per-entry matrix stores no longer coalesce.
__device__ __forceinline__ void tightenRow(const double4* P4, int* col, int* cnt, size_t N, int i, double xi,
                                           double yi, double zi, double r2) {
    const int n = cnt[i];
    int k = 0, c = 0;
    for (; c + 4 <= n; c += 4) {
        int j[4];
        double4 q[4];
#pragma unroll
        for (int u = 0; u < 4; u++) j[u] = col[(size_t)(c + u) * N];
#pragma unroll
        for (int u = 0; u < 4; u++) q[u] = P4[j[u]];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double v0 = q[u].x - xi, v1 = q[u].y - yi, v2 = q[u].z - zi;
            if (v0 * v0 + v1 * v1 + v2 * v2 < r2) {
                if (k != c + u) col[(size_t)k * N] = j[u];   // entries before the first drop stay put
                k++;
            }
        }
    }
    for (; c < n; c++) {
        const int j = col[(size_t)c * N];
        const double4 q = P4[j];
        const double v0 = q.x - xi, v1 = q.y - yi, v2 = q.z - zi;
        if (v0 * v0 + v1 * v1 + v2 * v2 < r2) {
            if (k != c) col[(size_t)k * N] = j;
            k++;
        }
    }
    if (k != n) cnt[i] = k;
}

__global__ __launch_bounds__(256) void kListTighten(KArgs a, int* nbr, int* cnt, int* nbrS, int* cntS, double r2,
                                                    double r2s) {
    if (a.is[I_REBUILD]) return;   // the whole list was rebuilt at the radius
    const int i = slotId();
    if (i >= a.N || !live(a, i)) return;
    const double4 qi = a.P4[i];
    tightenRow(a.P4, nbr + i, cnt, a.N, i, qi.x, qi.y, qi.z, r2);
    tightenRow(a.P4, nbrS + i, cntS, a.N, i, qi.x, qi.y, qi.z, r2s);
}

int launchListKeepStart(Ctx* ctx, double skin) {
    hipLaunchKernelGGL(kListFlagsReset, dim3(1), dim3(1), 0, ctx->stream, ctx->is);
    MPS_HIP(hipGetLastError());
    return launchBuildList(ctx, skin, nullptr);
}

constexpr double kTightenSkin = 0.04;   // l0

int launchListKeepRefresh(Ctx* ctx, double skin) {
    ctx->p4Flags = false;
    if (!ctx->patchList) MPS_HIP(hipMalloc(&ctx->patchList, sizeof(int) * (size_t)ctx->N));
    KArgs a = ctx->args();
    const int g = gridFor(ctx->N, kBlock);
    const double h = 0.5 * skin;
    hipLaunchKernelGGL(kListRefresh, dim3(g), dim3(kBlock), 0, ctx->stream, a, ctx->patchList, h * h);
    MPS_HIP(hipGetLastError());
    int rc = launchBuildList(ctx, 0.0, ctx->is + I_REBUILD);
    if (rc) return rc;
    hipLaunchKernelGGL(kPatchList, dim3(g), dim3(kBlock), 0, ctx->stream, a, ctx->nbr, ctx->nbrCnt, ctx->nbrS,
                       ctx->nbrCntS, ctx->patchList, listRadius2(ctx->P), listRadiusS2(ctx->P));
    // Below a skin of 0.04 l0 the extra entries are too few to upset the
    // consumers (PPE assembly 2.12 ms untightened at 0.04 l0 against 4.7 ms at
    // 0.06 l0) and the tightening's read of the whole list would cost more.
    if (skin > kTightenSkin * ctx->P.part_dist)
        hipLaunchKernelGGL(kListTighten, dim3(g), dim3(kBlock), 0, ctx->stream, a, ctx->nbr, ctx->nbrCnt, ctx->nbrS,
                           ctx->nbrCntS, listRadius2(ctx->P), listRadiusS2(ctx->P));
    MPS_HIP(hipGetLastError());
    ctx->listValid = true;
    return MPS_OK;
}

// Sum or maximum of n counts: grid-stride partial per thread, a block
// reduction, then one 64-bit atomic per workgroup.
__global__ __launch_bounds__(256) void kReduceCounts(const int* __restrict__ cnt, int n, unsigned long long* out,
                                                     int isMax) {
    __shared__ unsigned long long sh[4];
    unsigned long long v = 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const unsigned long long c = (unsigned long long)max(cnt[k], 0);
        v = isMax ? (c > v ? c : v) : v + c;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long w = __shfl_xor(v, o, 64);
        v = isMax ? (w > v ? w : v) : v + w;
    }
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = sh[0];
        for (int q = 1; q < 4; q++) t = isMax ? (sh[q] > t ? sh[q] : t) : t + sh[q];
        if (isMax) atomicMax(out, t);
        else atomicAdd(out, t);
    }
}

__global__ void kCountToScalar(const unsigned long long* v, double* out) { *out = (double)*v; }

// The sum of cnt[0 .. n) into the device scalar ds[slot], without waiting:
// the host reads it with the next status copy.
int launchCountToScalar(Ctx* ctx, const int* cnt, int n, int slot) {
    if (!ctx->red64) MPS_HIP(hipMalloc(&ctx->red64, sizeof(unsigned long long)));
    MPS_HIP(hipMemsetAsync(ctx->red64, 0, sizeof(unsigned long long), ctx->stream));
    if (cnt != nullptr && n > 0)
        hipLaunchKernelGGL(kReduceCounts, dim3(std::min(gridFor(n, 256), 1024)), dim3(256), 0, ctx->stream, cnt, n,
                           ctx->red64, 0);
    hipLaunchKernelGGL(kCountToScalar, dim3(1), dim3(1), 0, ctx->stream, ctx->red64, ctx->ds + slot);
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

int reduceCounts(Ctx* ctx, const int* cnt, int n, bool max, double* out) {
    if (!ctx->red64) MPS_HIP(hipMalloc(&ctx->red64, sizeof(unsigned long long)));
    MPS_HIP(hipMemsetAsync(ctx->red64, 0, sizeof(unsigned long long), ctx->stream));
    if (cnt != nullptr && n > 0)
        hipLaunchKernelGGL(kReduceCounts, dim3(std::min(gridFor(n, 256), 1024)), dim3(256), 0, ctx->stream, cnt, n,
                           ctx->red64, max ? 1 : 0);
    MPS_HIP(hipGetLastError());
    unsigned long long h = 0;
    MPS_HIP(hipMemcpyAsync(&h, ctx->red64, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    *out = (double)h;
    return MPS_OK;
}

int launchCompaction(Ctx* ctx) {
    ctx->asmMode = -1;   // slots move: an assembled matrix no longer matches them
    KArgs a = ctx->args();
    const int g = gridFor(ctx->N, kBlock);
    const int cap = 4096;
    if (ctx->P.domain_type_bc == 1) hipLaunchKernelGGL(kPeriodicWrap, dim3(g), dim3(kBlock), 0, ctx->stream, a);
    hipLaunchKernelGGL(kOutFlag, dim3(g), dim3(kBlock), 0, ctx->stream, a, ctx->outList, cap);
    hipLaunchKernelGGL(kCompact, dim3(1), dim3(1), 0, ctx->stream, a, ctx->outList, cap, ctx->slotOf);
    MPS_HIP(hipGetLastError());
    ctx->listValid = false;   // a compaction may move particles between slots
    ctx->p4Flags = false;
    return MPS_OK;
}

// extrapolatePressParticlesWallDummy (src/MpsPressure.cpp:792-846); in place
// is safe: only wall / dummy slots are written and only fluid slots are read.
template <bool LIST>
__global__ void kExtrapolateWallDummy2(KArgs a) {
    const int i = slotId();
    if (!live(a, i)) return;
    const mps_params& P = a.P;
    const int t = a.I[ITYPE][i];
    if (!(t == P.dummy_wall || (P.mps_type == MPS_WEAKLY && t == P.wall))) return;
    const double xi = PX(i), yi = PY(i), zi = PZ(i);
    double ni = 0.0, pr = 0.0;
    forNbrsPosS<LIST>(a, i, xi, yi, zi, [&](int j, double xj, double yj, double zj) {
        const double v0 = xj - xi, v1 = yj - yi, v2 = zj - zi;
        const double d2 = v0 * v0 + v1 * v1 + v2 * v2;
        if (d2 < P.re_s2 && a.I[ITYPE][j] == P.fluid) {
            const double wS = wgt(sqrt(d2), P.re_s, P.weight_type);
            ni += wS;
            pr += (a.D[DPRESS][j] - rhoOf(a, j) * (P.gravity[0] * v0 + P.gravity[1] * v1 + P.gravity[2] * v2)) * wS;
        }
    });
    if (pr < 0.0) return;
    a.D[DPRESS][i] = (ni > 0) ? pr / ni : pr;
}

// PolygonMesh::closestPointPNDBoundaryAABB for one mesh: the kernels keep a
// strictly closer point than the particle's current one and record the type.
int launchClosestPoint(Ctx* ctx, int type) {
    KArgs a = ctx->args();
    const int g = gridFor(ctx->N, kBlock);
    if (type == 0) {
        if (!ctx->hasMesh) { ctx->err = "no wall mesh installed"; return MPS_E_ARG; }
    } else if (type == 1 || type == 2) {
        const MeshSlot& m = ctx->extraMesh[type - 1];
        if (!m.on) { ctx->err = "mesh type not installed"; return MPS_E_ARG; }
        a.mV = m.V; a.mF = m.F; a.mN = m.N; a.nf = m.nf;
        a.bvh = (ctx->bvhMode == 1 || (ctx->bvhMode < 0 && m.nf > 64)) ? m.bvh : nullptr;
        a.bvhTri = m.bvhTri;
    } else {
        ctx->err = "mesh type must be 0, 1 or 2";
        return MPS_E_ARG;
    }
    ctx->p4Flags = false;   // the wall distances change
    if (a.bvh) hipLaunchKernelGGL(kClosestPointBvh, dim3(g), dim3(kBlock), 0, ctx->stream, a, type);
    else hipLaunchKernelGGL(kClosestPoint, dim3(g), dim3(kBlock), 0, ctx->stream, a, type);
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

int launchOp(Ctx* ctx, const char* op) {
    const std::string o(op);
    // operators that rewrite the wall distance or the boundary flags that
    // update_bc packed into P4.w
    if (o == "prediction" || o == "closest_point") ctx->p4Flags = false;
    KArgs a = ctx->args();
    const int g = gridFor(ctx->N, kBlock);
    const mps_params& P = ctx->P;
    if (o == "viscosity" && P.fluid_type == 1) {
        return launchViscosityNN(ctx);
    } else if (o == "viscosity") {
        if (ctx->listValid) {
            if (!ctx->V4) MPS_HIP(hipMalloc(&ctx->V4, sizeof(double4) * (size_t)ctx->N));
            a.V4 = ctx->V4;
            hipLaunchKernelGGL(kPackVel, dim3(g), dim3(kBlock), 0, ctx->stream, a, ctx->V4);
            L1(kViscosity<true>);
        }
        else if (ctx->P.domain_type_bc == 1) LW(kViscosity);   // walk through the periodic ghosts
        else hipLaunchKernelGGL(kViscosityWalk, dim3(gridFor(ctx->N, kWave)), dim3(kWave), 0, ctx->stream, a);
    }
    else if (o == "add_gravity") L1(kAddGravity);
    else if (o == "prediction") L1(kPrediction);
    else if (o == "collision") {
        if (P.collision_type == MPS_COLLISION_PC) LT(kCollisionPC);
        else LT(kCollisionDPC);
        hipLaunchKernelGGL(kApplyDvel, dim3(g), dim3(kBlock), 0, ctx->stream, a, P.collision_type == MPS_COLLISION_DPC ? 1 : 0);
        if (P.collision_type == MPS_COLLISION_DPC) ctx->listValid = false;
    } else if (o == "closest_point") {
        // every installed mesh in type order (src/main.cpp:296-300)
        for (int t = 0; t < 3; t++) {
            if (t > 0 && !ctx->extraMesh[t - 1].on) continue;
            int rc = launchClosestPoint(ctx, t);
            if (rc) return rc;
        }
    } else if (o == "update_near_mesh") L1(kUpdateNear);
    else if (o == "wall_npcd") LT(kWallNPCD);
    else if (o == "correction_matrix") LT(kCorrectionMatrix);
    else if (o == "initial_pnd") LT(kInitialPnd);
    else if (o == "pnd_neigh_npcd") LT(kPndNeighNPCD);
    else if (o == "pnd_diffusive") {
        LT(kPndDiffusive);
        if (P.wall_type == MPS_WALL_POLYGON) LT(kWallPndDiffusive);
        L1(kUpdatePnd);
        LT(kMeanPndWDS);
        L1(kMeanPndWDSFinish);
    } else if (o == "mean_pnd") {
        if (P.wall_type == MPS_WALL_POLYGON) LT(kMeanWallPnd);
        LT(kMeanPnd);
        L1(kMeanPndFinish);
    } else if (o == "mean_pnd_only") {
        LT(kMeanPnd);
        L1(kMeanPndFinish);
    } else if (o == "mean_neigh_fluid_pnd") LT(kMeanNeighFluidPnd);
    else if (o == "update_bc") {
        LT(kUpdateBC);
        // the assembly reads the neighbours' inner / far-from-wall flags with their positions
        const bool typesFit = std::max({P.fluid, P.wall, P.dummy_wall, P.ghost}) < 512 &&
                              std::min({P.fluid, P.wall, P.dummy_wall, P.ghost}) >= -512;
        if (ctx->listValid && typesFit) {
            hipLaunchKernelGGL(kPackP4Flags, dim3(g), dim3(kBlock), 0, ctx->stream, a);
            ctx->p4Flags = true;
        }
    }
    else if (o == "normal_particles") LT(kNormalParticles);
    else if (o == "wall_normal_particles") LT(kWallNormalParticles);
    else if (o == "press_explicit") L1(kPressExplicit);
    else if (o == "vel_divergence") LT(kVelDivergence);
    else if (o == "wall_vel_divergence") LT(kWallVelDivergence);
    else if (o == "extrapolate_wall_dummy") LT(kExtrapolateWallDummy2); else if (o == "press_gradient") LT(kPressGradient);
    else if (o == "wall_press_gradient") LT(kWallPressGradient);
    else if (o == "prediction_press_gradient") LT(kPredPressGradient);
    else if (o == "prediction_wall_press_gradient") LT(kPredWallPressGradient);
    else if (o == "wall_viscosity") LT(kWallViscosity);
    else if (o == "correction") {
        hipLaunchKernelGGL(kCorrection, dim3(g), dim3(kBlock), 0, ctx->stream, a);   // g <= redCap partials
        hipLaunchKernelGGL(kVelMaxFinish, dim3(1), dim3(1024), 0, ctx->stream, a, g);
        ctx->listValid = false;
    } else if (o == "shifting") {
        if (P.shifting_type == 1) {
            LW(kShift1);
            hipLaunchKernelGGL(kCommitStaged, dim3(g), dim3(kBlock), 0, ctx->stream, a, (int)DU);
        } else if (P.shifting_type == 2) {
            LW(kShift2Conc);
            LW(kShift2Grad);
            hipLaunchKernelGGL(kCommitStaged, dim3(g), dim3(kBlock), 0, ctx->stream, a, (int)DX);
            if (P.wall_type == MPS_WALL_POLYGON) {
                LW(kShift2Wall);
                hipLaunchKernelGGL(kCommitStaged, dim3(g), dim3(kBlock), 0, ctx->stream, a, (int)DX);
            }
            ctx->listValid = false;
        }
    } else if (o == "wall_dummy_velocity") LW(kWallDummyVelocity);
    else {
        ctx->err = "unknown operator " + o;
        return MPS_E_ARG;
    }
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

}  // namespace mpsgpu
