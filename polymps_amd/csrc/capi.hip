// Context management and the C ABI of the MI355X MPS library.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "kern_common.h"

namespace mpsgpu {

static std::mutex gErrMu;
static std::string gErr;
void setGlobalError(const std::string& e) {
    std::lock_guard<std::mutex> lk(gErrMu);
    gErr = e;
}
std::string globalError() {
    std::lock_guard<std::mutex> lk(gErrMu);
    return gErr;
}

KArgs Ctx::args() const {
    KArgs a;
    a.P = P;
    a.N = N;
    a.nB = nB;
    for (int f = 0; f < ND; f++) a.D[f] = D[f];
    for (int f = 0; f < NI; f++) a.I[f] = I[f];
    a.cellStart = cellStart;
    a.nbrMask = nbrMask;
    a.maskW = maskW;
    a.nbr = nbr;
    a.nbrCnt = nbrCnt;
    // periodic domains build the full list only, which serves as the small one
    const bool oneList = P.domain_type_bc == 1 || nbrS == nullptr;
    a.nbrS = oneList ? nbr : nbrS;
    a.nbrCntS = oneList ? nbrCnt : nbrCntS;
    a.KS = oneList ? K : KS;
    a.P4 = P4;
    a.p4Flags = p4Flags ? 1 : 0;
    a.V4 = V4;
    a.K = K;
    a.mV = mV;
    a.mF = mF;
    a.mN = mN;
    a.nf = nf;
    a.bvh = (bvhMode == 1 || (bvhMode < 0 && nf > 64)) ? bvh : nullptr;
    a.bvhTri = bvhTri;
    a.wtab = wtab;
    a.ds = ds;
    a.is = is;
    a.red = red;
    a.volForce = pow(P.part_dist, P.dim);
    return a;
}

// ------------------------------------------------------------ field map
struct FieldDesc {
    const char* name;
    int kind;      // 0: double (1 comp), 1: double vec (3 comps), 2: int
    int f0;        // first field index
};
static const FieldDesc kFields[] = {
    {"pos", 1, DX}, {"vel", 1, DU}, {"acc", 1, DAX}, {"acc_star", 1, DASX}, {"npcd", 1, DNPX},
    {"mirror", 1, DMX}, {"wall_pos", 1, DWPX}, {"polygon_normal", 1, DPNX}, {"force_wall", 1, DFWX},
    {"cm1", 1, DCM0}, {"cm2", 1, DCM3}, {"cm3", 1, DCM6}, {"grad_conc", 1, DGCX}, {"normal", 1, DNRMX},
    {"press", 0, DPRESS}, {"press_avg", 0, DPAVG}, {"pndi", 0, DPNDI}, {"pndki", 0, DPNDKI},
    {"pndski", 0, DPNDSKI}, {"pnd_small", 0, DPNDSMALL}, {"npcd2", 0, DNPCD2}, {"conc", 0, DCONC},
    {"div_u", 0, DDIVU}, {"diffusive", 0, DDIFF}, {"pnd_wall", 0, DPNDWALL}, {"riw2", 0, DRIW2},
    {"ppe_x", 0, DPPEX}, {"ppe_b", 0, DPPEB}, {"dev_dot_n", 0, DDEVN},
    {"type", 2, ITYPE}, {"bc", 2, IBC}, {"num_neigh", 2, INNEIGH}, {"num_neigh_wall", 2, INNWALL},
    {"near_wall", 2, INEARWALL}, {"near_mesh_type", 2, INEARMESH},
    // granular phase (fluid_type 1)
    {"rho", 0, DRHO}, {"meu", 0, DMEU}, {"cv", 0, DCV}, {"ii", 0, DII}, {"meu_y", 0, DMEUY},
    {"inertia", 0, DINERTIA}, {"p_rheo", 0, DPRHEO}, {"p_smooth", 0, DPSMOOTH}, {"vf", 0, DVF},
    {"s11", 0, DS11}, {"s12", 0, DS12}, {"s13", 0, DS13}, {"s22", 0, DS22}, {"s23", 0, DS23}, {"s33", 0, DS33},
    {"ptype", 2, IPTYPE},
};
static const FieldDesc* findField(const char* name) {
    for (const auto& f : kFields)
        if (strcmp(f.name, name) == 0) return &f;
    return nullptr;
}

// slot -> reference order (AoS for vectors)
__global__ void kToCanonD(int n, const int* canon, const double* s0, const double* s1, const double* s2, double* dst,
                          int comps) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = canon[k];
    if (c < 0 || c >= n) return;
    if (comps == 1) {
        dst[c] = s0[k];
    } else {
        dst[c * 3] = s0[k]; dst[c * 3 + 1] = s1[k]; dst[c * 3 + 2] = s2[k];
    }
}
__global__ void kToCanonI(int n, const int* canon, const int* s, int* dst) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = canon[k];
    if (c < 0 || c >= n) return;
    dst[c] = s[k];
}
__global__ void kFromCanonD(int n, const int* canon, const double* src, double* d0, double* d1, double* d2, int comps) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = canon[k];
    if (c < 0 || c >= n) return;
    if (comps == 1) {
        d0[k] = src[c];
    } else {
        d0[k] = src[c * 3]; d1[k] = src[c * 3 + 1]; d2[k] = src[c * 3 + 2];
    }
}
__global__ void kFromCanonI(int n, const int* canon, const int* src, int* d) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = canon[k];
    if (c < 0 || c >= n) return;
    d[k] = src[c];
}

__global__ void kInitState(KArgs a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.N) return;
    for (int f = DAX; f < ND; f++) a.D[f][k] = 0.0;
    a.D[DMX][k] = 0.0; a.D[DMY][k] = 0.0; a.D[DMZ][k] = 0.0;
    a.D[DPNDI][k] = 0.0; a.D[DPPEX][k] = 0.0; a.D[DFWX][k] = 0.0; a.D[DFWY][k] = 0.0; a.D[DFWZ][k] = 0.0;
    for (int f = DPNDSKI; f < NPD; f++) a.D[f][k] = 0.0;
    a.D[DRIW2][k] = 10e8 * a.P.part_dist;
    a.I[ICANON][k] = k;
    a.I[IBC][k] = 0;
    a.I[INNEIGH][k] = 0;
    a.I[INNWALL][k] = 0;
    a.I[INEARWALL][k] = 0;
    a.I[INEARMESH][k] = 0;
}

// wall tables of PolygonMesh::initWijnNeigh (src/PolygonMesh.cpp:377-527)
static void buildWallTables(Ctx* ctx) {
    const mps_params& P = ctx->P;
    const int dim = (int)P.dim, wt = P.weight_type;
    const double lo = P.part_dist, reL = P.re_l, reS = P.re_s;
    const double reS2 = reS * reS, reL2 = reL * reL;
    const int n = kWallTable;
    ctx->hxPND.clear(); ctx->hxNeigh.clear(); ctx->hZij.clear(); ctx->hnNeigh.clear();
    const int limS = (int)ceil(reS / lo), limL = (int)ceil(reL / lo);
    for (int i = 1; i <= n; i++) {
        const double xi = 0.0, yi = i * reS / n, zi = 0.0;
        const double riw = (dim == 2) ? sqrt((0.0 - xi) * (0.0 - xi) + (0.0 - yi) * (0.0 - yi))
                                      : sqrt((0.0 - xi) * (0.0 - xi) + (0.0 - yi) * (0.0 - yi) + (0.0 - zi) * (0.0 - zi));
        ctx->hxPND.push_back(riw / reS);
        ctx->hxNeigh.push_back(riw / reL);
        double n0 = 0.0;
        for (int ix = -limS; ix <= limS; ix++)
            for (int iy = -limS; iy <= 0; iy++)
                for (int iz = (dim == 2 ? 0 : -limS); iz <= (dim == 2 ? 0 : limS); iz++) {
                    const double xj = lo * (double)ix, yj = lo * (double)(iy - 0.5), zj = lo * (double)iz;
                    const double r2 = (dim == 2) ? (xj - xi) * (xj - xi) + (yj - yi) * (yj - yi)
                                                 : (xj - xi) * (xj - xi) + (yj - yi) * (yj - yi) + (zj - zi) * (zj - zi);
                    if (!(r2 < reS2) || r2 == 0.0) continue;
                    const double r = sqrt(r2);
                    if (wt == 1) n0 += reS / r + r / reS - 2.0;
                    else if (wt == 2) n0 += reS / r - r / reS;
                    else if (wt == 3) n0 += pow(1.0 - r / reS, 3.0);
                    else if (wt == 4) n0 += pow(1.0 - r / reS, 2.0);
                    else n0 += reS / r - 1.0;
                }
        ctx->hZij.push_back(n0);
        int nn = 0;
        for (int ix = -limL; ix <= limL; ix++)
            for (int iy = -limL; iy <= 0; iy++)
                for (int iz = (dim == 2 ? 0 : -limL); iz <= (dim == 2 ? 0 : limL); iz++) {
                    const double xj = lo * (double)ix, yj = lo * (double)(iy - 0.5), zj = lo * (double)iz;
                    const double r2 = (dim == 2) ? (xj - xi) * (xj - xi) + (yj - yi) * (yj - yi)
                                                 : (xj - xi) * (xj - xi) + (yj - yi) * (yj - yi) + (zj - zi) * (zj - zi);
                    if (r2 < reL2 && r2 != 0.0) nn += 1;
                }
        ctx->hnNeigh.push_back(nn);
    }
}

static int syncStatus(Ctx* ctx) {
    MPS_HIP(hipMemcpyAsync(ctx->hostInts, ctx->is, sizeof(int) * NIS, hipMemcpyDeviceToHost, ctx->stream));
    MPS_HIP(hipMemcpyAsync(ctx->hostScalars, ctx->ds, sizeof(double) * NDS, hipMemcpyDeviceToHost, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    MPS_HIP(hipGetLastError());
    ctx->nLiveHost = ctx->hostInts[I_NLIVE];
    ctx->velMaxHost = ctx->hostScalars[S_VELMAX];
    if (ctx->hostInts[I_OVERFLOW] > 0) {
        char buf[160];
        snprintf(buf, sizeof(buf),
                 "neighbour table overflow: %d entries where %d (stencil row) / %d (PPE row) fit; "
                 "particles are packed far denser than the particle distance",
                 ctx->hostInts[I_OVERFLOW], 32 * ctx->maskW, ctx->K);
        ctx->err = buf;
        return MPS_E_CAPACITY;
    }
    if (ctx->hostInts[I_COMPACT_OVF]) {
        ctx->err = "more out-of-domain particles in one call than the compaction list holds";
        return MPS_E_CAPACITY;
    }
    return MPS_OK;
}

// A pressure solve that stops above its tolerance still leaves a pressure
// field, and the reference's step goes on after printing "Failed solving"
// (src/MpsPressure.cpp:440-442): such a failure lets the step finish and is
// returned at its end.  A non-finite residual ends the call at once.
static bool softSolverFailure(const Ctx* ctx, int rc) { return rc == MPS_E_SOLVER && ctx->solverFailed == 1; }

// calcNormalParticles, then the polygon-wall part (src/main.cpp order before
// updateParticleBC when the ARC rule is used)
static int arcNormals(Ctx* ctx) {
    int rc = launchOp(ctx, "normal_particles");
    if (rc == MPS_OK && ctx->P.wall_type == MPS_WALL_POLYGON) rc = launchOp(ctx, "wall_normal_particles");
    return rc;
}

// calcVelDivergence, then the polygon-wall term (src/MpsPressure.cpp:46-53)
static int velDivergence(Ctx* ctx) {
    int rc = launchOp(ctx, "vel_divergence");
    if (rc == MPS_OK && ctx->P.wall_type == MPS_WALL_POLYGON) rc = launchOp(ctx, "wall_vel_divergence");
    return rc;
}

static int runInit(Ctx* ctx) {
    const mps_params& P = ctx->P;
    int rc;
    if ((rc = launchBucketSort(ctx))) return rc;
    if ((rc = launchCompaction(ctx))) return rc;
    if (P.wall_type == MPS_WALL_POLYGON) {
        if (!ctx->hasMesh) { ctx->err = "polygon walls need a mesh (mps_set_mesh)"; return MPS_E_ARG; }
        if (P.fem_on && !ctx->extraMesh[0].on) {
            ctx->err = "fem_MESH needs the deformable mesh (mps_set_mesh_type)";
            return MPS_E_ARG;
        }
        if (P.forced_on && !ctx->extraMesh[1].on) {
            ctx->err = "forced_MESH needs the forced mesh (mps_set_mesh_type)";
            return MPS_E_ARG;
        }
        if ((rc = launchOp(ctx, "closest_point"))) return rc;
        if ((rc = launchOp(ctx, "update_near_mesh"))) return rc;
    }
    if ((rc = launchBuildList(ctx))) return rc;
    if (P.wall_type == MPS_WALL_POLYGON)
        if ((rc = launchOp(ctx, "wall_npcd"))) return rc;
    if ((rc = launchOp(ctx, "initial_pnd"))) return rc;
    if ((rc = launchOp(ctx, "mean_pnd"))) return rc;
    if ((rc = launchOp(ctx, "mean_neigh_fluid_pnd"))) return rc;
    if (P.free_surf_type == MPS_FS_PND_ARC && (rc = arcNormals(ctx))) return rc;
    if ((rc = launchOp(ctx, "update_bc"))) return rc;
    if ((rc = syncStatus(ctx))) return rc;
    if (P.mps_type == MPS_EXPLICIT || P.mps_type == MPS_WEAKLY) {
        if ((rc = launchOp(ctx, "press_explicit"))) return rc;
    } else {
        if (P.mps_type == MPS_IMPLICIT_PND_DIVU && (rc = velDivergence(ctx))) return rc;
        if ((rc = ppeWarmStart(ctx))) return rc;
        if ((rc = ppeAssemble(ctx))) return rc;
        rc = ppeSolve(ctx);
        if (rc && !softSolverFailure(ctx, rc)) return rc;
    }
    const int rs = syncStatus(ctx);
    return rs ? rs : rc;
}

static int calcPress(Ctx* ctx) {
    const mps_params& P = ctx->P;
    int rc;
    if (P.mps_type == MPS_EXPLICIT || P.mps_type == MPS_WEAKLY) return launchOp(ctx, "press_explicit");
    if (P.mps_type == MPS_IMPLICIT_PND_DIVU && (rc = velDivergence(ctx))) return rc;
    // the guess first: the mixed assembly computes the first residual from it
    if ((rc = ppeWarmStart(ctx))) return rc;
    if ((rc = ppeAssemble(ctx))) return rc;
    return ppeSolve(ctx);
}

static int forcedMeshUpdate(Ctx* ctx, double dt, double time);
static void freeSlot(MeshSlot& m);

// Skin (m) of a neighbour list kept over the step, 0 to build the list after
// the prediction as usual.  Automatic: list_skin_factor (3) times the largest
// displacement the last velocity maximum and gravity give in one step, plus
// 0.002 l0; a faster flow (skin above list_skin_cap, 0.08 l0) builds after
// the prediction, where the larger list and its tightening would cost more
// than the second build saves.
// The refresh checks every particle's actual displacement, so the choice only
// affects speed.
static double keptListSkin(const Ctx* ctx) {
    const mps_params& P = ctx->P;
    if (ctx->listSkin == 0.0 || P.domain_type_bc == 1 || P.collision_type != MPS_COLLISION_PC) return 0.0;
    if (ctx->listSkin > 0.0) return ctx->listSkin * P.part_dist;
    const double g = std::sqrt(P.gravity[0] * P.gravity[0] + P.gravity[1] * P.gravity[1] + P.gravity[2] * P.gravity[2]);
    const double skin =
        ctx->listSkinFactor * P.time_step * (ctx->velMaxHost + P.time_step * g) + 0.002 * P.part_dist;
    return skin > ctx->listSkinCap * P.part_dist ? 0.0 : skin;
}

static int runStep(Ctx* ctx) {
    const mps_params& P = ctx->P;
    int rc;
#define OP(name)                                   \
    do {                                           \
        if ((rc = launchOp(ctx, name))) return rc; \
    } while (0)
    if ((rc = launchBucketSort(ctx))) return rc;
    const double skin = keptListSkin(ctx);
    ctx->lastSkin = skin;
    if (skin > 0.0 && (rc = launchListKeepStart(ctx, skin))) return rc;
    // A flow too fast to keep the list (automatic skin above its cap): the
    // operators before the prediction take their neighbours from a list
    // built at the step's start positions, with no skin, instead of bucket
    // walks; only the large radius unless an operator before the prediction
    // reads the small one.  The list after the prediction is built as usual.
    const bool smallEarly = P.fluid_type == 1 || P.relax_press < 1.0;
    ctx->lastStartList = skin == 0.0 && ctx->listSkin < 0.0 && P.domain_type_bc != 1;
    if (ctx->lastStartList && (rc = launchBuildList(ctx, 0.0, nullptr, smallEarly))) return rc;
    if (P.fluid_type == 1) {   // src/main.cpp:259-262
        if ((rc = launchGranular(ctx, "volume_fraction"))) return rc;
        if ((rc = launchGranular(ctx, "viscosity_interaction"))) return rc;
    }
    OP("viscosity");
    if (ctx->lastStartList && !smallEarly) ctx->listValid = false;   // its small-radius list is stale
    OP("add_gravity");
    if (P.relax_press < 1.0) {
        OP("prediction_press_gradient");
        if (P.wall_type == MPS_WALL_POLYGON) OP("prediction_wall_press_gradient");
    }
    OP("prediction");
    if ((rc = launchCompaction(ctx))) return rc;
    if (skin > 0.0) {
        if ((rc = launchListKeepRefresh(ctx, skin))) return rc;
        OP("collision");
    } else if (P.collision_type == MPS_COLLISION_PC) {
        if ((rc = launchBuildList(ctx))) return rc;
        OP("collision");
    } else {
        OP("collision");
        if ((rc = launchBuildList(ctx))) return rc;
    }
    if (P.wall_type == MPS_WALL_POLYGON) {
        OP("closest_point");
        OP("update_near_mesh");
        OP("wall_npcd");
    }
    // The first correction matrix of the step (src/main.cpp:317-319) is read
    // before the second call rewrites every entry (:375-377) only by the
    // diffusive PND and the velocity divergence of the DivU source term;
    // otherwise it is never read and is skipped.
    const bool cmReadEarly = P.pnd_type == MPS_PND_DIFFUSIVE || P.mps_type == MPS_IMPLICIT_PND_DIVU;
    if (P.gradient_correction && cmReadEarly) OP("correction_matrix");
    OP("pnd_neigh_npcd");
    if (P.pnd_type == MPS_PND_DIFFUSIVE) OP("pnd_diffusive");
    else if (P.pnd_type == MPS_PND_MEAN_SUM_WIJ) OP("mean_pnd");
    if (P.free_surf_type == MPS_FS_PND_ARC && (rc = arcNormals(ctx))) return rc;
    OP("update_bc");
    const int solveRc = calcPress(ctx);
    if (solveRc && !softSolverFailure(ctx, solveRc)) return solveRc;
    if (P.wall_type == MPS_WALL_PARTICLE) OP("extrapolate_wall_dummy");
    if (P.gradient_correction) OP("correction_matrix");
    OP("press_gradient");
    if (P.wall_type == MPS_WALL_POLYGON) {
        OP("wall_press_gradient");
    }
    if (P.fluid_type == 1) {   // src/main.cpp:387-398
        if ((rc = launchGranular(ctx, "volume_fraction"))) return rc;
        if ((rc = launchGranular(ctx, "viscosity_interaction"))) return rc;
        if (P.wall_type == MPS_WALL_POLYGON && (rc = launchGranular(ctx, "wall_viscosity_interaction"))) return rc;
    }
    if (P.wall_type == MPS_WALL_POLYGON) OP("wall_viscosity");
    // src/main.cpp:410-414: the forced mesh moves before the corrector
    if (P.forced_on && (rc = forcedMeshUpdate(ctx, P.time_step, ctx->timeCurrent))) return rc;
    OP("correction");
    if ((rc = launchCompaction(ctx))) return rc;
    OP("shifting");
    if (P.wall_type == MPS_WALL_PARTICLE) OP("wall_dummy_velocity");
    if ((rc = launchCompaction(ctx))) return rc;
#undef OP
    ctx->numIter++;
    ctx->timeCurrent += P.time_step;
    rc = syncStatus(ctx);
    return rc ? rc : solveRc;
}

}  // namespace mpsgpu

using namespace mpsgpu;

struct mps_ctx : public Ctx {};

// Staging area of mps_download_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory: host code uses only the device
// address hipGetSymbolAddress returns, looked up once per device with that
// device current and kept in gScratchPtr.
constexpr size_t kScratchBytes = size_t(64) << 20;   // 64 MiB: 8M doubles (a vector field of 2.79M particles)
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex gScratchMu;
static void* gScratchPtr[kMaxDevices] = {};

// The current device's scratch address (ctx->dev must be current).
static int scratchAddress(Ctx* ctx, void** out) {
    if (ctx->dev < 0 || ctx->dev >= kMaxDevices) {
        ctx->err = "device index out of range for the staging area";
        return MPS_E_ARG;
    }
    std::lock_guard<std::mutex> lk(gScratchMu);
    if (!gScratchPtr[ctx->dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return MPS_E_HIP;
        }
        gScratchPtr[ctx->dev] = ptr;
    }
    *out = gScratchPtr[ctx->dev];
    return MPS_OK;
}

#define CHECK_CTX(c) \
    if (!(c)) return MPS_E_ARG

extern "C" {

int mps_create(mps_ctx** out, const mps_params* p, int n, int device) {
    if (!out || !p || n <= 0) return MPS_E_ARG;
    if (!p->derived) { setGlobalError("parameters are not derived (call mps_params_derive)"); return MPS_E_ARG; }
    if ((int)p->dim != 2 && (int)p->dim != 3) { setGlobalError("dimension must be 2 or 3"); return MPS_E_ARG; }
    mps_ctx* ctx = new mps_ctx();
    *out = ctx;
    ctx->P = *p;
    ctx->N = n;
    ctx->nLiveHost = n;
    ctx->nB = p->num_buckets_xyz;
    ctx->dev = device;
    MPS_HIP(hipSetDevice(device));
    MPS_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    {   // the device's staging area for mps_download_noalloc, looked up once per device
        void* st = nullptr;
        const int rc = scratchAddress(ctx, &st);
        if (rc) return rc;
    }
    MPS_HIP(hipEventCreate(&ctx->ev0));
    MPS_HIP(hipEventCreate(&ctx->ev1));
    const size_t nd = sizeof(double) * (size_t)n, ni = sizeof(int) * (size_t)n;
    for (int f = 0; f < ND; f++) MPS_HIP(hipMalloc(&ctx->D[f], nd));
    for (int f = 0; f < NPD; f++) MPS_HIP(hipMalloc(&ctx->Dalt[f], nd));
    for (int f = 0; f < NI; f++) MPS_HIP(hipMalloc(&ctx->I[f], ni));
    for (int f = 0; f < NPI; f++) MPS_HIP(hipMalloc(&ctx->Ialt[f], ni));
    MPS_HIP(hipMalloc(&ctx->cellStart, sizeof(int) * (ctx->nB + 2)));
    MPS_HIP(hipMalloc(&ctx->cellCount, sizeof(int) * (ctx->nB + 1)));
    MPS_HIP(hipMalloc(&ctx->key, ni));
    MPS_HIP(hipMalloc(&ctx->rank, ni));
    MPS_HIP(hipMalloc(&ctx->order, ni));
    MPS_HIP(hipMalloc(&ctx->order2, ni));
    MPS_HIP(hipMalloc(&ctx->slotOf, ni));
    MPS_HIP(hipMalloc(&ctx->scanTmp, sizeof(int) * 8192));
    MPS_HIP(hipMalloc(&ctx->outList, sizeof(int) * 4096));
    // neighbour-list capacity: the reL ball holds ~122 particles at rest in
    // 3D (~28 in 2D at reL = 3.1); the capacity leaves room for compression
    const double rr = std::max(p->re_l, p->re_s) / p->part_dist;
    const double ball = ((int)p->dim == 3) ? 4.18879 * rr * rr * rr : 3.14159 * rr * rr;
    ctx->K = std::max(32, ((int)(ball * 2.0) + 31) / 32 * 32);
    // the small-radius list (re_s, collision distance)
    const double rs = std::max(p->re_s / p->part_dist, p->dist_limit_ratio);
    const double ballS = ((int)p->dim == 3) ? 4.18879 * rs * rs * rs : 3.14159 * rs * rs;
    ctx->KS = std::min(ctx->K, std::max(32, ((int)(ballS * 2.0) + 31) / 32 * 32));
    // neighbour bitmasks: a stencil row spans 3 buckets of side bucket_side,
    // (bucket_side / l0)^dim particles each at rest; 1.6x that leaves room for
    // compression (the build reports an overflow as MPS_E_CAPACITY)
    const double perRow = 3.0 * std::pow(p->bucket_side / p->part_dist, p->dim);
    ctx->maskW = std::max(1, (int)std::ceil(1.6 * perRow / 32.0));
    ctx->maskWAlloc = ctx->maskW;
    const int rows = ((int)p->dim == 3) ? 9 : 3;
    MPS_HIP(hipMalloc(&ctx->nbrMask, sizeof(unsigned) * (size_t)rows * ctx->maskW * n));
    MPS_HIP(hipMalloc(&ctx->nbr, sizeof(int) * (size_t)ctx->K * n));
    MPS_HIP(hipMalloc(&ctx->nbrCnt, ni));
    MPS_HIP(hipMemsetAsync(ctx->nbrCnt, 0, ni, ctx->stream));
    MPS_HIP(hipMalloc(&ctx->nbrMaskS, sizeof(unsigned) * (size_t)rows * ctx->maskW * n));
    MPS_HIP(hipMalloc(&ctx->nbrS, sizeof(int) * (size_t)ctx->KS * n));
    MPS_HIP(hipMalloc(&ctx->nbrCntS, ni));
    MPS_HIP(hipMemsetAsync(ctx->nbrCntS, 0, ni, ctx->stream));
    MPS_HIP(hipMalloc(&ctx->P4, sizeof(double4) * (size_t)n));
    // ELL arrays: ellWidth(K) entries per row (K, a multiple of 32, and one
    // 16-entry batch the row kernels read past a row's end), 64-row tiles
    MPS_HIP(hipMalloc(&ctx->aVal, sizeof(double) * (size_t)ellWidth(ctx->K) * ellRows(n)));
    MPS_HIP(hipMalloc(&ctx->aCol, sizeof(unsigned short) * (size_t)ellWidth(ctx->K) * ellRows(n)));
    MPS_HIP(hipMalloc(&ctx->aEsc, sizeof(int) * (size_t)kPpeEsc * n));
    MPS_HIP(hipMalloc(&ctx->aCnt, ni));
    MPS_HIP(hipMemsetAsync(ctx->aCnt, 0, ni, ctx->stream));   // nnz reads 0 before the first assembly
    MPS_HIP(hipMalloc(&ctx->aDiag, nd));
    MPS_HIP(hipMalloc(&ctx->cgR, nd));
    MPS_HIP(hipMalloc(&ctx->cgP, nd));
    MPS_HIP(hipMalloc(&ctx->cgZ, nd));
    MPS_HIP(hipMalloc(&ctx->cgQ, nd));
    MPS_HIP(hipMalloc(&ctx->cgInvD, nd));
    MPS_HIP(hipMalloc(&ctx->ds, sizeof(double) * NDS));
    MPS_HIP(hipMalloc(&ctx->is, sizeof(int) * NIS));
    ctx->redCap = std::max(kRedBlocks, gridFor(n, 256));
    MPS_HIP(hipMalloc(&ctx->red, sizeof(double) * (3 * ctx->redCap + 16)));
    MPS_HIP(hipMalloc(&ctx->asmRed, sizeof(double) * (3 * ctx->redCap + 16)));
    if (p->fluid_type == 1) {   // granular phase: carried per-step fields, bed-level grid
        for (int q = 0; q < kNNCarriedCount; q++) MPS_HIP(hipMalloc(&ctx->DaltNN[q], nd));
        MPS_HIP(hipMalloc(&ctx->IaltPtype, ni));
        granularGridSize(*p, &ctx->bedKx, &ctx->bedKy);
        const size_t cells = (size_t)ctx->bedKx * ctx->bedKy;
        MPS_HIP(hipMalloc(&ctx->bedKeys, sizeof(unsigned long long) * (2 * cells + 1)));
        MPS_HIP(hipMalloc(&ctx->bedIdx, sizeof(int) * cells));
    }
    MPS_HIP(hipMalloc(&ctx->wtab, sizeof(double) * 4 * kWallTable));
    MPS_HIP(hipHostMalloc(&ctx->hostScalars, sizeof(double) * NDS));
    MPS_HIP(hipHostMalloc(&ctx->hostInts, sizeof(int) * NIS));
    MPS_HIP(hipMemsetAsync(ctx->ds, 0, sizeof(double) * NDS, ctx->stream));
    MPS_HIP(hipMemsetAsync(ctx->is, 0, sizeof(int) * NIS, ctx->stream));
    buildWallTables(ctx);
    std::vector<double> wt(4 * kWallTable);
    for (int q = 0; q < kWallTable; q++) {
        wt[q] = ctx->hxPND[q];
        wt[kWallTable + q] = ctx->hxNeigh[q];
        wt[2 * kWallTable + q] = ctx->hZij[q];
        wt[3 * kWallTable + q] = (double)ctx->hnNeigh[q];
    }
    MPS_HIP(hipMemcpyAsync(ctx->wtab, wt.data(), sizeof(double) * wt.size(), hipMemcpyHostToDevice, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    return MPS_OK;
}

int mps_destroy(mps_ctx* ctx) {
    if (!ctx) return MPS_OK;
    hipSetDevice(ctx->dev);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    vtuFree(ctx);
    for (int f = 0; f < ND; f++) hipFree(ctx->D[f]);
    for (int f = 0; f < NPD; f++) hipFree(ctx->Dalt[f]);
    for (int f = 0; f < NI; f++) hipFree(ctx->I[f]);
    for (int f = 0; f < NPI; f++) hipFree(ctx->Ialt[f]);
    for (int q = 0; q < kNNCarriedCount; q++)
        if (ctx->DaltNN[q]) hipFree(ctx->DaltNN[q]);
    void* bufs[] = {ctx->cellStart, ctx->cellCount, ctx->key, ctx->rank, ctx->order, ctx->order2, ctx->slotOf,
                    ctx->bvh, ctx->bvhTri, ctx->scanTmp, ctx->outList, ctx->nbrMask, ctx->nbr, ctx->nbrCnt, ctx->P4, ctx->mV, ctx->mF, ctx->mN, ctx->wtab,
                    ctx->aVal, ctx->aCol, ctx->aEsc, ctx->aCnt, ctx->aDiag, ctx->aVal32, ctx->aLc8, ctx->aLcEsc, ctx->aWin,
                    ctx->aDiagInv32,
                    ctx->aDiag32, ctx->cgR, ctx->cgP,
                    ctx->cgZ, ctx->cgQ, ctx->cgInvD, ctx->cgX, ctx->fR, ctx->fP, ctx->fZ, ctx->fQ, ctx->fX,
                    ctx->bcgR0, ctx->bcgY, ctx->bcgS, ctx->bcgT, ctx->patchList, ctx->fixList, ctx->V4, ctx->nbrMaskS, ctx->nbrS, ctx->nbrCntS,
                    ctx->ds, ctx->is, ctx->red, ctx->asmRed, ctx->red64, ctx->IaltPtype, ctx->bedKeys, ctx->bedIdx};
    for (void* b : bufs)
        if (b) hipFree(b);
    freeSlot(ctx->extraMesh[0]);
    freeSlot(ctx->extraMesh[1]);
    if (ctx->hostScalars) hipHostFree(ctx->hostScalars);
    if (ctx->hostInts) hipHostFree(ctx->hostInts);
    if (ctx->ev0) hipEventDestroy(ctx->ev0);
    if (ctx->ev1) hipEventDestroy(ctx->ev1);
    for (auto e : ctx->evPool) hipEventDestroy(e);
    gnnFree(ctx);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return MPS_OK;
}

const char* mps_last_error(const mps_ctx* ctx) {
    static thread_local std::string tl;
    if (ctx && !ctx->err.empty()) return ctx->err.c_str();
    tl = globalError();
    return tl.c_str();
}

int mps_get_params(const mps_ctx* ctx, mps_params* out) {
    CHECK_CTX(ctx);
    if (!out) return MPS_E_ARG;
    *out = ctx->P;
    return MPS_OK;
}

int mps_upload_state(mps_ctx* ctx, const int* type, const double* pos, const double* vel, const double* press,
                     const double* press_avg) {
    CHECK_CTX(ctx);
    if (!type || !pos || !vel || !press || !press_avg) return MPS_E_ARG;
    MPS_HIP(hipSetDevice(ctx->dev));
    const int n = ctx->N;
    KArgs a = ctx->args();
    hipLaunchKernelGGL(kInitState, dim3(gridFor(n, 256)), dim3(256), 0, ctx->stream, a);
    MPS_HIP(hipMemcpyAsync(ctx->I[ITYPE], type, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    std::vector<double> tmp(n);
    for (int c = 0; c < 3; c++) {
        for (int i = 0; i < n; i++) tmp[i] = pos[i * 3 + c];
        MPS_HIP(hipMemcpyAsync(ctx->D[DX + c], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        MPS_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < n; i++) tmp[i] = vel[i * 3 + c];
        MPS_HIP(hipMemcpyAsync(ctx->D[DU + c], tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        MPS_HIP(hipStreamSynchronize(ctx->stream));
    }
    MPS_HIP(hipMemcpyAsync(ctx->D[DPRESS], press, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    MPS_HIP(hipMemcpyAsync(ctx->D[DPAVG], press_avg, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    {   // phase, density and viscosity of every particle (granular type 1 becomes fluid of phase 2)
        const int rc = launchInitPhase(ctx);
        if (rc) return rc;
    }
    MPS_HIP(hipMemsetAsync(ctx->is, 0, sizeof(int) * NIS, ctx->stream));
    MPS_HIP(hipMemsetAsync(ctx->ds, 0, sizeof(double) * NDS, ctx->stream));
    MPS_HIP(hipMemcpyAsync(ctx->is + I_NLIVE, &n, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    ctx->nLiveHost = n;
    ctx->listValid = false;
    ctx->numIter = 0;
    ctx->timeCurrent = 0.0;
    ctx->asmMode = -1;
    ctx->r0Fresh = false;
    return MPS_OK;
}

int mps_write_vtu(mps_ctx* ctx, const char* path, int flags) {
    CHECK_CTX(ctx);
    if (!path) return MPS_E_ARG;
    const int rc = vtuSubmit(ctx, path, flags);
    return rc != MPS_OK ? rc : vtuWait(ctx);
}

int mps_write_vtu_async(mps_ctx* ctx, const char* path, int flags) {
    CHECK_CTX(ctx);
    if (!path) return MPS_E_ARG;
    return vtuSubmit(ctx, path, flags);
}

int mps_output_wait(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    return vtuWait(ctx);
}

int mps_download(mps_ctx* ctx, const char* field, void* out, size_t count) {
    CHECK_CTX(ctx);
    if (!field || !out) return MPS_E_ARG;
    const FieldDesc* fd = findField(field);
    if (!fd) { ctx->err = std::string("unknown field ") + field; return MPS_E_ARG; }
    MPS_HIP(hipSetDevice(ctx->dev));
    const int n = ctx->N;
    const int comps = fd->kind == 1 ? 3 : 1;
    if (count < (size_t)n * comps) { ctx->err = "output buffer too small"; return MPS_E_ARG; }
    const int g = gridFor(n, 256);
    if (fd->kind == 2) {
        int* tmp = nullptr;
        MPS_HIP(hipMalloc(&tmp, sizeof(int) * n));
        hipLaunchKernelGGL(kToCanonI, dim3(g), dim3(256), 0, ctx->stream, n, ctx->I[ICANON], ctx->I[fd->f0], tmp);
        hipError_t e = hipMemcpyAsync(out, tmp, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream);
        hipStreamSynchronize(ctx->stream);
        hipFree(tmp);
        if (e != hipSuccess) { ctx->err = hipGetErrorString(e); return MPS_E_HIP; }
    } else {
        double* tmp = nullptr;
        MPS_HIP(hipMalloc(&tmp, sizeof(double) * n * comps));
        const double* s0 = ctx->D[fd->f0];
        const double* s1 = comps == 3 ? ctx->D[fd->f0 + 1] : s0;
        const double* s2 = comps == 3 ? ctx->D[fd->f0 + 2] : s0;
        hipLaunchKernelGGL(kToCanonD, dim3(g), dim3(256), 0, ctx->stream, n, ctx->I[ICANON], s0, s1, s2, tmp, comps);
        hipError_t e = hipMemcpyAsync(out, tmp, sizeof(double) * n * comps, hipMemcpyDeviceToHost, ctx->stream);
        hipStreamSynchronize(ctx->stream);
        hipFree(tmp);
        if (e != hipSuccess) { ctx->err = hipGetErrorString(e); return MPS_E_HIP; }
    }
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

int mps_download_noalloc(mps_ctx* ctx, const char* field, void* out, size_t count) {
    CHECK_CTX(ctx);
    if (!field || !out) return MPS_E_ARG;
    const FieldDesc* fd = findField(field);
    if (!fd) { ctx->err = std::string("unknown field ") + field; return MPS_E_ARG; }
    MPS_HIP(hipSetDevice(ctx->dev));
    const int n = ctx->N;
    const int comps = fd->kind == 1 ? 3 : 1;
    if (count < (size_t)n * comps) { ctx->err = "output buffer too small"; return MPS_E_ARG; }
    const size_t bytes = (fd->kind == 2 ? sizeof(int) : sizeof(double)) * (size_t)n * comps;
    if (bytes > kScratchBytes) {
        char msg[200];
        snprintf(msg, sizeof msg, "mps_download_noalloc: field %s needs %zu bytes of staging, more than the %zu bytes "
                 "the library holds; use mps_download", field, bytes, kScratchBytes);
        ctx->err = msg;
        return MPS_E_CAPACITY;
    }
    void* st = nullptr;
    int rc = scratchAddress(ctx, &st);
    if (rc) return rc;
    // canon is a permutation of 0..n-1, so the gather writes every staged entry
    const int g = gridFor(n, 256);
    if (fd->kind == 2) {
        hipLaunchKernelGGL(kToCanonI, dim3(g), dim3(256), 0, ctx->stream, n, ctx->I[ICANON], ctx->I[fd->f0], (int*)st);
    } else {
        const double* s0 = ctx->D[fd->f0];
        const double* s1 = comps == 3 ? ctx->D[fd->f0 + 1] : s0;
        const double* s2 = comps == 3 ? ctx->D[fd->f0 + 2] : s0;
        hipLaunchKernelGGL(kToCanonD, dim3(g), dim3(256), 0, ctx->stream, n, ctx->I[ICANON], s0, s1, s2, (double*)st,
                           comps);
    }
    MPS_HIP(hipGetLastError());
    MPS_HIP(hipMemcpyAsync(out, st, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    return MPS_OK;
}

int mps_upload_field(mps_ctx* ctx, const char* field, const void* in, size_t count) {
    CHECK_CTX(ctx);
    if (!field || !in) return MPS_E_ARG;
    const FieldDesc* fd = findField(field);
    if (!fd) { ctx->err = std::string("unknown field ") + field; return MPS_E_ARG; }
    MPS_HIP(hipSetDevice(ctx->dev));
    const int n = ctx->N;
    const int comps = fd->kind == 1 ? 3 : 1;
    if (count < (size_t)n * comps) { ctx->err = "input buffer too small"; return MPS_E_ARG; }
    const int g = gridFor(n, 256);
    if (fd->kind == 2) {
        int* tmp = nullptr;
        MPS_HIP(hipMalloc(&tmp, sizeof(int) * n));
        MPS_HIP(hipMemcpyAsync(tmp, in, sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(kFromCanonI, dim3(g), dim3(256), 0, ctx->stream, n, ctx->I[ICANON], tmp, ctx->I[fd->f0]);
        hipStreamSynchronize(ctx->stream);
        hipFree(tmp);
    } else {
        double* tmp = nullptr;
        MPS_HIP(hipMalloc(&tmp, sizeof(double) * n * comps));
        MPS_HIP(hipMemcpyAsync(tmp, in, sizeof(double) * n * comps, hipMemcpyHostToDevice, ctx->stream));
        double* d0 = ctx->D[fd->f0];
        double* d1 = comps == 3 ? ctx->D[fd->f0 + 1] : d0;
        double* d2 = comps == 3 ? ctx->D[fd->f0 + 2] : d0;
        hipLaunchKernelGGL(kFromCanonD, dim3(g), dim3(256), 0, ctx->stream, n, ctx->I[ICANON], tmp, d0, d1, d2, comps);
        hipStreamSynchronize(ctx->stream);
        hipFree(tmp);
    }
    // positions or types changed: the neighbour list and its packed copy are stale
    if ((fd->kind != 2 && fd->f0 == DX) || (fd->kind == 2 && fd->f0 == ITYPE)) ctx->listValid = false;
    ctx->p4Flags = false;   // the boundary flags packed by update_bc may be stale
    ctx->r0Fresh = false;   // and so may the residual the assembly left
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

int mps_get_scalar(mps_ctx* ctx, const char* name, double* out) {
    CHECK_CTX(ctx);
    if (!name || !out) return MPS_E_ARG;
    int rc = syncStatus(ctx);
    if (rc) return rc;
    const std::string s(name);
    if (s == "vel_max") *out = ctx->velMaxHost;
    else if (s == "cfl") *out = ctx->P.time_step * ctx->velMaxHost / ctx->P.part_dist;
    else if (s == "time") *out = ctx->timeCurrent;
    else if (s == "solver_iter") *out = ctx->solverIter;
    else if (s == "solver_err") *out = ctx->solverErr;
    else if (s == "outer_iterations") *out = ctx->outerIter;
    else if (s == "num_particles") *out = ctx->nLiveHost;
    else if (s == "iterations") *out = ctx->numIter;
    else if (s == "ppe_ms") {
        const int rc2 = ppeSolveMs(ctx, out);
        if (rc2) return rc2;
    }
    else if (s == "true_residual") *out = ctx->trueResidual;
    else if (s == "fallback") *out = ctx->fallbackUsed;
    else if (s == "learned_stall_ratio") *out = ctx->learnedStallRatio;
    else if (s == "learned_tol_scale") *out = ctx->learnedTolScale;
    else if (s == "ppe_window") *out = ctx->ppeWindow ? 1.0 : 0.0;
    else if (s == "ppe_window_blocks") {
        const int rc2 = ppeWindowBlocks(ctx, out);
        if (rc2) return rc2;
    }
    else if (s == "solver_failed") *out = ctx->solverFailed;
    else if (s == "cfl_visc") {
        const int rc2 = granularCflVisc(ctx, out);
        if (rc2) return rc2;
    }
    else if (s == "solver_failures") *out = (double)ctx->solverFailures;
    else if (s == "list_max") {
        // the longest row of the current neighbour list
        const int rc2 = reduceCounts(ctx, ctx->nbrCnt, ctx->N, true, out);
        if (rc2) return rc2;
    }
    else if (s == "list_skin") *out = ctx->lastSkin / ctx->P.part_dist;
    else if (s == "list_start") *out = ctx->lastStartList ? 1.0 : 0.0;
    else if (s == "list_skin_cap") *out = ctx->listSkinCap;
    else if (s == "list_skin_factor") *out = ctx->listSkinFactor;
    else if (s == "list_rebuilt") *out = ctx->lastSkin > 0.0 ? ctx->hostInts[I_REBUILD] : 1;
    else if (s == "list_patched") *out = ctx->lastSkin > 0.0 ? ctx->hostInts[I_PATCHCNT] : 0;
    else if (s == "list_capacity") *out = 32 * ctx->maskW;
    else if (s == "capacity") *out = ctx->N;
    else if (s == "nnz") {
        // entries of the last PPE assembly
        const int rc2 = reduceCounts(ctx, ctx->aCnt, ctx->N, false, out);
        if (rc2) return rc2;
    }
    else if (s == "nnz_escaped") *out = ctx->hostInts[I_NESC];
    else if (s == "window_escapes") *out = ctx->hostInts[I_NLCESC];
    else if (s == "list_entries" || s == "list_entries_small") {
        // entries of the last neighbour list (large radius / small radius), summed over the slots
        const int* cnt = s == "list_entries" ? ctx->nbrCnt : ctx->args().nbrCntS;
        std::vector<int> h(ctx->N);
        MPS_HIP(hipMemcpyAsync(h.data(), cnt, sizeof(int) * ctx->N, hipMemcpyDeviceToHost, ctx->stream));
        MPS_HIP(hipStreamSynchronize(ctx->stream));
        double sum = 0.0;
        for (int v : h) sum += v;
        *out = sum;
    }
    else if (s == "spmv_ms_total") *out = ctx->spmvMsTotal;
    else if (s == "spmv_launches") *out = (double)ctx->spmvLaunches;
    else if (s == "spmv_bytes_total") *out = ctx->spmvBytesTotal;
    else { ctx->err = "unknown scalar " + s; return MPS_E_ARG; }
    return MPS_OK;
}

int mps_set_scalar(mps_ctx* ctx, const char* name, double value) {
    CHECK_CTX(ctx);
    ctx->r0Fresh = false;   // a solve setting may change: the next solve recomputes its first residual
    const std::string s(name ? name : "");
    if (s == "vel_max") {
        MPS_HIP(hipMemcpyAsync(ctx->ds + S_VELMAX, &value, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
        MPS_HIP(hipStreamSynchronize(ctx->stream));
        ctx->velMaxHost = value;
    } else if (s == "time") {
        ctx->timeCurrent = value;
    } else if (s == "solver_failures") {
        ctx->solverFailures = (long)value;
    } else if (s == "profile") {
        ctx->profile = value != 0.0;
        ctx->spmvMsTotal = 0.0;
        ctx->spmvLaunches = 0;
        ctx->spmvBytesTotal = 0.0;
    } else if (s == "mesh_bvh") {
        // closest-point search: -1 BVH above 64 faces (default), 0 brute force, 1 BVH
        if (value != -1.0 && value != 0.0 && value != 1.0) {
            ctx->err = "mesh_bvh must be -1, 0 or 1";
            return MPS_E_ARG;
        }
        ctx->bvhMode = (int)value;
    } else if (s == "learned_mode") {
        // 0: off, 1: learned warm start, 2: warm start + learned flexible-CG preconditioner
        if (value != 0.0 && value != 1.0 && value != 2.0) {
            ctx->err = "learned_mode must be 0, 1 or 2";
            return MPS_E_ARG;
        }
        ctx->P.learned_mode = (int)value;
    } else if (s == "learned_stall_ratio") {
        // the learned preconditioner stays in use while each refinement pass
        // cuts the true residual norm to at most this fraction (default 0.5)
        if (!(value > 0.0 && value < 1.0)) {
            ctx->err = "learned_stall_ratio must be in (0, 1)";
            return MPS_E_ARG;
        }
        ctx->learnedStallRatio = value;
    } else if (s == "learned_tol_scale") {
        // true-residual target of a solve with the learned preconditioner,
        // as a multiple of the solver tolerance (default 0.1)
        if (!(value > 0.0 && value <= 1.0)) {
            ctx->err = "learned_tol_scale must be in (0, 1]";
            return MPS_E_ARG;
        }
        ctx->learnedTolScale = value;
    } else if (s == "ppe_window") {
        // the CG solves' SpMVs gather from per-block LDS windows (1,
        // default) or from global memory (0); same results, from the next
        // assembly on
        if (!(value == 0.0 || value == 1.0)) {
            ctx->err = "ppe_window must be 0 or 1";
            return MPS_E_ARG;
        }
        ctx->ppeWindow = value != 0.0;
    } else if (s == "mask_words") {
        // candidate window of a stencil row in 32-slot words (at most the
        // allocated width); rows beyond it are listed by a direct walk
        const int w = (int)value;
        if (w < 1 || w > ctx->maskWAlloc || (double)w != value) {
            ctx->err = "mask_words must be an integer in [1, allocated width]";
            return MPS_E_ARG;
        }
        ctx->maskW = w;
        ctx->listValid = false;
    } else if (s == "list_skin") {
        // neighbour list kept over the step: skin in units of l0, < 0 automatic, 0 off
        ctx->listSkin = value;
    } else if (s == "list_skin_cap") {
        // largest automatic skin, in units of l0
        if (!(value >= 0.0 && value <= 1.0)) {
            ctx->err = "list_skin_cap must be in [0, 1]";
            return MPS_E_ARG;
        }
        ctx->listSkinCap = value;
    } else if (s == "list_skin_factor") {
        // automatic skin: multiple of the largest displacement of one step (at least 2: a pair
        // closes in by up to twice one particle's displacement)
        if (!(value >= 2.0 && value <= 10.0)) {
            ctx->err = "list_skin_factor must be in [2, 10]";
            return MPS_E_ARG;
        }
        ctx->listSkinFactor = value;
    } else if (s == "ppe_warm_start") {
        // initial guess of each pressure solve: 0 previous solution, 1 linear, 2 quadratic extrapolation
        if (value != 0.0 && value != 1.0 && value != 2.0 && value != 3.0) {
            ctx->err = "ppe_warm_start must be 0, 1, 2 or 3";
            return MPS_E_ARG;
        }
        ctx->P.ppe_warm_start = (int)value;
    } else if (s == "precision_mode") {
        // takes effect at the next assembly; a solve needs an assembly in the same mode
        if (value != MPS_PREC_FP64 && value != MPS_PREC_MIXED) {
            ctx->err = "precision_mode must be 0 (fp64) or 1 (mixed)";
            return MPS_E_ARG;
        }
        ctx->P.precision_mode = (int)value;
    } else {
        ctx->err = "unknown scalar " + s;
        return MPS_E_ARG;
    }
    return MPS_OK;
}

int mps_sync(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    return syncStatus(ctx);
}

}  // extern "C"

namespace mpsgpu {
// Median-split BVH over the triangles (longest axis of the centroid bounds,
// at most 4 triangles per leaf), built once on the host for a static mesh.
static int buildBvhInto(Ctx* ctx, const double* V, const int* F, int nf, BvhNode** outBvh, int** outTri,
                        int* outNodes) {
    struct Box { double lo[3], hi[3]; };
    std::vector<Box> tb(nf);
    std::vector<double> cen(3 * (size_t)nf);
    for (int f = 0; f < nf; f++) {
        for (int k = 0; k < 3; k++) { tb[f].lo[k] = 1e300; tb[f].hi[k] = -1e300; }
        for (int c = 0; c < 3; c++) {
            const double* v = V + 3 * F[3 * f + c];
            for (int k = 0; k < 3; k++) {
                tb[f].lo[k] = std::min(tb[f].lo[k], v[k]);
                tb[f].hi[k] = std::max(tb[f].hi[k], v[k]);
            }
        }
        for (int k = 0; k < 3; k++) cen[3 * f + k] = 0.5 * (tb[f].lo[k] + tb[f].hi[k]);
    }
    std::vector<int> idx(nf);
    for (int f = 0; f < nf; f++) idx[f] = f;
    std::vector<BvhNode> nodes;
    nodes.reserve(2 * (size_t)nf / 2 + 8);
    struct Item { int node, begin, end; };
    std::vector<Item> work;
    nodes.push_back(BvhNode{});
    work.push_back({0, 0, nf});
    while (!work.empty()) {
        const Item it = work.back();
        work.pop_back();
        BvhNode nd;
        for (int k = 0; k < 3; k++) { nd.lo[k] = 1e300; nd.hi[k] = -1e300; }
        double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
        for (int q = it.begin; q < it.end; q++) {
            const int f = idx[q];
            for (int k = 0; k < 3; k++) {
                nd.lo[k] = std::min(nd.lo[k], tb[f].lo[k]);
                nd.hi[k] = std::max(nd.hi[k], tb[f].hi[k]);
                clo[k] = std::min(clo[k], cen[3 * f + k]);
                chi[k] = std::max(chi[k], cen[3 * f + k]);
            }
        }
        if (it.end - it.begin <= 4) {
            nd.a = it.begin;
            nd.b = -(it.end - it.begin);
            nodes[it.node] = nd;
            continue;
        }
        int ax = 0;
        for (int k = 1; k < 3; k++)
            if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
        const int mid = (it.begin + it.end) / 2;
        std::nth_element(idx.begin() + it.begin, idx.begin() + mid, idx.begin() + it.end, [&](int x, int y) {
            return cen[3 * x + ax] < cen[3 * y + ax] || (cen[3 * x + ax] == cen[3 * y + ax] && x < y);
        });
        nd.a = (int)nodes.size();
        nodes.push_back(BvhNode{});
        nd.b = (int)nodes.size();
        nodes.push_back(BvhNode{});
        nodes[it.node] = nd;
        work.push_back({nd.a, it.begin, mid});
        work.push_back({nd.b, mid, it.end});
    }
    // the tree's shape depends on nf only, so a rebuild of a moved mesh reuses the buffers
    if (*outBvh == nullptr || *outNodes != (int)nodes.size()) {
        if (*outBvh) (void)hipFree(*outBvh);
        if (*outTri) (void)hipFree(*outTri);
        *outBvh = nullptr;
        *outTri = nullptr;
        MPS_HIP(hipMalloc(outBvh, sizeof(BvhNode) * nodes.size()));
        MPS_HIP(hipMalloc(outTri, sizeof(int) * (size_t)nf));
    }
    MPS_HIP(hipMemcpyAsync(*outBvh, nodes.data(), sizeof(BvhNode) * nodes.size(), hipMemcpyHostToDevice, ctx->stream));
    MPS_HIP(hipMemcpyAsync(*outTri, idx.data(), sizeof(int) * (size_t)nf, hipMemcpyHostToDevice, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    *outNodes = (int)nodes.size();
    return MPS_OK;
}
static int buildBvh(Ctx* ctx, const double* V, const int* F, int nf) {
    return buildBvhInto(ctx, V, F, nf, &ctx->bvh, &ctx->bvhTri, &ctx->bvhNodes);
}

// unit face normals (v1 - v0) x (v2 - v0), zero for a degenerate face (the
// normals of mps_mesh_read_stl and of igl::per_face_normals)
static void faceNormals(const std::vector<double>& V, const std::vector<int>& F, std::vector<double>& N) {
    const int nf = (int)(F.size() / 3);
    N.resize(3 * (size_t)nf);
    for (int t = 0; t < nf; t++) {
        const double* a = &V[(size_t)F[t * 3] * 3];
        const double* b = &V[(size_t)F[t * 3 + 1] * 3];
        const double* c = &V[(size_t)F[t * 3 + 2] * 3];
        const double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
        const double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double r = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; k++) N[(size_t)t * 3 + k] = (r == 0) ? 0.0 : n[k] / r;
    }
}

static int uploadSlot(Ctx* ctx, MeshSlot& m, bool rebuildBvh) {
    if (!m.V) {
        MPS_HIP(hipMalloc(&m.V, sizeof(double) * 3 * (size_t)m.nv));
        MPS_HIP(hipMalloc(&m.F, sizeof(int) * 3 * (size_t)m.nf));
        MPS_HIP(hipMalloc(&m.N, sizeof(double) * 3 * (size_t)m.nf));
        MPS_HIP(hipMemcpyAsync(m.F, m.hF.data(), sizeof(int) * 3 * (size_t)m.nf, hipMemcpyHostToDevice, ctx->stream));
    }
    MPS_HIP(hipMemcpyAsync(m.V, m.hV.data(), sizeof(double) * 3 * (size_t)m.nv, hipMemcpyHostToDevice, ctx->stream));
    MPS_HIP(hipMemcpyAsync(m.N, m.hN.data(), sizeof(double) * 3 * (size_t)m.nf, hipMemcpyHostToDevice, ctx->stream));
    // the BVH when the search will use it (launchClosestPoint's rule)
    if (rebuildBvh && (ctx->bvhMode == 1 || (ctx->bvhMode < 0 && m.nf > 64)))
        return buildBvhInto(ctx, m.hV.data(), m.hF.data(), m.nf, &m.bvh, &m.bvhTri, &m.bvhNodes);
    MPS_HIP(hipStreamSynchronize(ctx->stream));   // the host copies may change after the call
    return MPS_OK;
}

static void freeSlot(MeshSlot& m) {
    void* bufs[] = {m.V, m.F, m.N, m.bvh, m.bvhTri};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    m = MeshSlot();
}

// PolygonMesh::updateForcedPolygonMesh (src/PolygonMesh.cpp:173-374) with the
// law the reference has switched on (rotation about the y axis through
// (Xcg, Zcg), steady amplitude, its ramp off) or its translation branch.
// Every node moves by its velocity at the start of the step times dt (the
// reference's explicit update); the angle itself is not integrated.
static int forcedMeshUpdate(Ctx* ctx, double dt, double time) {
    MeshSlot& m = ctx->extraMesh[1];
    if (!m.on) {
        ctx->err = "no forced mesh installed";
        return MPS_E_ARG;
    }
    mps_params& P = ctx->P;
    const bool rotation = P.forced_motion == 1;
    double omegaAmp = 0.0, dx = 0.0, dy = 0.0, dz = 0.0;
    if (rotation) {
        const double kPi = 3.14159265358979323846;
        const double Tmotion = P.forced_period;
        const double R_ksi = P.forced_theta_max, R_ksi_dot = 0.0;
        omegaAmp = R_ksi_dot * sin(2.0 * kPi * time / Tmotion) + R_ksi * 2.0 * kPi / Tmotion * cos(2.0 * kPi * time / Tmotion);
        omegaAmp = omegaAmp * kPi / 180.0;   // degrees/s -> rad/s
    } else {
        const double vel = P.forced_velocity;
        P.uniform_vel_wall[0] = 0.0;
        P.uniform_vel_wall[1] = 0.0;
        P.uniform_vel_wall[2] = vel;
        dz = vel * dt;
    }
    const double Xcg = P.forced_center[0], Zcg = P.forced_center[1];
    for (int q = 0; q < m.nv; q++) {
        double* v = &m.hV[(size_t)q * 3];
        if (rotation) {
            const double xi = v[0] - Xcg;
            const double zi = v[2] - Zcg;
            const double vx = omegaAmp * zi;
            const double vz = -omegaAmp * xi;
            dx = vx * dt;
            dz = vz * dt;
        }
        v[0] += dx;
        v[1] += dy;
        v[2] += dz;
    }
    faceNormals(m.hV, m.hF, m.hN);
    return uploadSlot(ctx, m, ctx->bvhMode != 0);
}
}  // namespace mpsgpu

extern "C" {

int mps_set_mesh(mps_ctx* ctx, int nv, const double* V, int nf, const int* F, const double* N) {
    CHECK_CTX(ctx);
    if (nv <= 0 || nf <= 0 || !V || !F || !N) return MPS_E_ARG;
    for (int q = 0; q < 3 * nf; q++)
        if (F[q] < 0 || F[q] >= nv) { ctx->err = "face index out of range"; return MPS_E_ARG; }
    MPS_HIP(hipSetDevice(ctx->dev));
    if (ctx->mV) hipFree(ctx->mV);
    if (ctx->mF) hipFree(ctx->mF);
    if (ctx->mN) hipFree(ctx->mN);
    MPS_HIP(hipMalloc(&ctx->mV, sizeof(double) * 3 * nv));
    MPS_HIP(hipMalloc(&ctx->mF, sizeof(int) * 3 * nf));
    MPS_HIP(hipMalloc(&ctx->mN, sizeof(double) * 3 * nf));
    MPS_HIP(hipMemcpy(ctx->mV, V, sizeof(double) * 3 * nv, hipMemcpyHostToDevice));
    MPS_HIP(hipMemcpy(ctx->mF, F, sizeof(int) * 3 * nf, hipMemcpyHostToDevice));
    MPS_HIP(hipMemcpy(ctx->mN, N, sizeof(double) * 3 * nf, hipMemcpyHostToDevice));
    ctx->nv = nv;
    ctx->nf = nf;
    ctx->hasMesh = true;
    ctx->rigidV.assign(V, V + 3 * (size_t)nv);
    ctx->rigidF.assign(F, F + 3 * (size_t)nf);
    ctx->rigidN.assign(N, N + 3 * (size_t)nf);
    return buildBvh(ctx, V, F, nf);
}

int mps_set_mesh_type(mps_ctx* ctx, int type, int nv, const double* V, int nf, const int* F, const double* N) {
    CHECK_CTX(ctx);
    if (type == MPS_MESH_FIXED) return mps_set_mesh(ctx, nv, V, nf, F, N);
    if (type != MPS_MESH_DEFORMABLE && type != MPS_MESH_FORCED) { ctx->err = "mesh type must be 0, 1 or 2"; return MPS_E_ARG; }
    if (nv <= 0 || nf <= 0 || !V || !F || !N) return MPS_E_ARG;
    for (int q = 0; q < 3 * nf; q++)
        if (F[q] < 0 || F[q] >= nv) { ctx->err = "face index out of range"; return MPS_E_ARG; }
    MPS_HIP(hipSetDevice(ctx->dev));
    MeshSlot& m = ctx->extraMesh[type - 1];
    freeSlot(m);
    m.nv = nv;
    m.nf = nf;
    m.hV.assign(V, V + 3 * (size_t)nv);
    m.hF.assign(F, F + 3 * (size_t)nf);
    m.hN.assign(N, N + 3 * (size_t)nf);
    m.on = true;
    return uploadSlot(ctx, m, ctx->bvhMode != 0);
}

int mps_get_mesh(mps_ctx* ctx, int type, int* nv, double* V, int* nf, int* F, double* N) {
    CHECK_CTX(ctx);
    const std::vector<double>* hv;
    const std::vector<int>* hf;
    const std::vector<double>* hn;
    if (type == MPS_MESH_FIXED) {
        if (!ctx->hasMesh) { ctx->err = "no rigid mesh installed"; return MPS_E_ARG; }
        hv = &ctx->rigidV; hf = &ctx->rigidF; hn = &ctx->rigidN;
    } else if (type == MPS_MESH_DEFORMABLE || type == MPS_MESH_FORCED) {
        const MeshSlot& m = ctx->extraMesh[type - 1];
        if (!m.on) { ctx->err = "mesh type not installed"; return MPS_E_ARG; }
        hv = &m.hV; hf = &m.hF; hn = &m.hN;
    } else {
        ctx->err = "mesh type must be 0, 1 or 2";
        return MPS_E_ARG;
    }
    if (nv) *nv = (int)(hv->size() / 3);
    if (nf) *nf = (int)(hf->size() / 3);
    if (V) memcpy(V, hv->data(), sizeof(double) * hv->size());
    if (F) memcpy(F, hf->data(), sizeof(int) * hf->size());
    if (N) memcpy(N, hn->data(), sizeof(double) * hn->size());
    return MPS_OK;
}

int mps_forced_mesh_update(mps_ctx* ctx, double dt, double time) {
    CHECK_CTX(ctx);
    MPS_HIP(hipSetDevice(ctx->dev));
    return forcedMeshUpdate(ctx, dt, time);
}

int mps_mesh_closest_point_type(mps_ctx* ctx, int type) {
    CHECK_CTX(ctx);
    MPS_HIP(hipSetDevice(ctx->dev));
    int rc = launchClosestPoint(ctx, type);
    if (rc) return rc;
    return syncStatus(ctx);
}

int mps_get_wall_tables(mps_ctx* ctx, double* xpnd, double* xneigh, double* zij, int* nneigh) {
    CHECK_CTX(ctx);
    for (int q = 0; q < kWallTable; q++) {
        if (xpnd) xpnd[q] = ctx->hxPND[q];
        if (xneigh) xneigh[q] = ctx->hxNeigh[q];
        if (zij) zij[q] = ctx->hZij[q];
        if (nneigh) nneigh[q] = ctx->hnNeigh[q];
    }
    return MPS_OK;
}

#define OPFN(fn, name)                          \
    int fn(mps_ctx* ctx) {                      \
        CHECK_CTX(ctx);                         \
        hipSetDevice(ctx->dev);                 \
        int rc = launchOp(ctx, name);           \
        if (rc) return rc;                      \
        return syncStatus(ctx);                 \
    }

#define GRANULAR(fn, name)                      \
    int fn(mps_ctx* ctx) {                      \
        CHECK_CTX(ctx);                         \
        hipSetDevice(ctx->dev);                 \
        int rc = launchGranular(ctx, name);     \
        if (rc) return rc;                      \
        return syncStatus(ctx);                 \
    }
GRANULAR(mps_volume_fraction, "volume_fraction")
GRANULAR(mps_viscosity_interaction, "viscosity_interaction")
GRANULAR(mps_wall_viscosity_interaction, "wall_viscosity_interaction")
#undef GRANULAR

int mps_bucket_update(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    int rc = launchBucketSort(ctx);
    if (rc) return rc;
    return syncStatus(ctx);
}
OPFN(mps_viscosity, "viscosity")
OPFN(mps_add_gravity, "add_gravity")
OPFN(mps_prediction, "prediction")
int mps_check_out_domain(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    int rc = launchCompaction(ctx);
    if (rc) return rc;
    return syncStatus(ctx);
}
int mps_build_neighbor_list(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    int rc = launchBuildList(ctx);
    if (rc) return rc;
    return syncStatus(ctx);
}
OPFN(mps_collision, "collision")
OPFN(mps_mesh_closest_point, "closest_point")
OPFN(mps_mesh_update_near, "update_near_mesh")
OPFN(mps_wall_npcd, "wall_npcd")
OPFN(mps_correction_matrix, "correction_matrix")
OPFN(mps_initial_pnd, "initial_pnd")
OPFN(mps_pnd_neigh_npcd, "pnd_neigh_npcd")
OPFN(mps_pnd_diffusive, "pnd_diffusive")
OPFN(mps_mean_pnd, "mean_pnd")
OPFN(mps_mean_neigh_fluid_pnd, "mean_neigh_fluid_pnd")
OPFN(mps_update_bc, "update_bc")
OPFN(mps_extrapolate_wall_dummy, "extrapolate_wall_dummy")
OPFN(mps_press_gradient, "press_gradient")
OPFN(mps_wall_press_gradient, "wall_press_gradient")
OPFN(mps_prediction_press_gradient, "prediction_press_gradient")
OPFN(mps_normal_particles, "normal_particles")
OPFN(mps_wall_normal_particles, "wall_normal_particles")
OPFN(mps_prediction_wall_press_gradient, "prediction_wall_press_gradient")
OPFN(mps_wall_viscosity, "wall_viscosity")
OPFN(mps_correction, "correction")
OPFN(mps_shifting, "shifting")
OPFN(mps_wall_dummy_velocity, "wall_dummy_velocity")

int mps_calc_press(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    int rc = calcPress(ctx);
    if (rc) return rc;
    return syncStatus(ctx);
}
int mps_vel_divergence(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    int rc = velDivergence(ctx);
    if (rc) return rc;
    return syncStatus(ctx);
}
int mps_ppe_assemble(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    int rc = ppeAssemble(ctx);
    if (rc) return rc;
    return syncStatus(ctx);
}
int mps_ppe_solve(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    int rc = ppeSolve(ctx);
    if (rc) return rc;
    return syncStatus(ctx);
}

int mps_ppe_warm_start(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    int rc = ppeWarmStart(ctx);
    if (rc) return rc;
    return syncStatus(ctx);
}

int mps_get_buckets(mps_ctx* ctx, int* first, int* last, int* next) {
    CHECK_CTX(ctx);
    MPS_HIP(hipSetDevice(ctx->dev));
    const int n = ctx->N, nB = ctx->nB;
    const mps_params& P = ctx->P;
    std::vector<int> start(nB + 2), canon(n);
    MPS_HIP(hipMemcpyAsync(start.data(), ctx->cellStart, sizeof(int) * (nB + 2), hipMemcpyDeviceToHost, ctx->stream));
    MPS_HIP(hipMemcpyAsync(canon.data(), ctx->I[ICANON], sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<int> f(nB, -1), l(nB, -1), nx(n, -1);
    // bucket b holds slots start[b] .. start[b + 1] - 1 in ascending reference index
    for (int b = 0; b < nB; b++) {
        if (start[b + 1] <= start[b]) continue;
        f[b] = canon[start[b]];
        l[b] = canon[start[b + 1] - 1];
        for (int k = start[b]; k + 1 < start[b + 1]; k++) nx[canon[k]] = canon[k + 1];
    }
    // MpsBucket::copyDataBetweenBuckets (src/MpsBucket.cpp:356-417): the ghost
    // layers of the periodic direction list the opposite interior layer
    if (P.domain_type_bc == 1) {
        const int d = P.periodic_dir[0] ? 0 : P.periodic_dir[1] ? 1 : 2;
        const int nd[3] = {P.num_buckets_x, P.num_buckets_y, (int)P.dim == 3 ? P.num_buckets_z : 1};
        const int stride = d == 0 ? 1 : d == 1 ? P.num_buckets_x : P.num_buckets_xy;
        const int off = stride * (nd[d] - 2);
        for (int b = 0; b < nB; b++) {
            const int c = (b / stride) % nd[d];
            if (c == 0) { f[b] = f[b + off]; l[b] = l[b + off]; }
            else if (c == nd[d] - 1) { f[b] = f[b - off]; l[b] = l[b - off]; }
        }
    }
    if (first) std::copy(f.begin(), f.end(), first);
    if (last) std::copy(l.begin(), l.end(), last);
    if (next) std::copy(nx.begin(), nx.end(), next);
    return nB;
}

int mps_ppe_matrix(mps_ctx* ctx, int* row_ptr, int* cols, double* vals, int* nnz_out) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    const int n = ctx->N, K = ctx->K, KE = ellWidth(K);
    const size_t ne = (size_t)KE * ellRows(n);   // entries of the ELL arrays
    std::vector<int> cnt(n), canon(n), esc((size_t)kPpeEsc * n), col((size_t)K * n);
    std::vector<unsigned short> code(ne);
    std::vector<double> val(ne), diag(n);
    MPS_HIP(hipMemcpy(cnt.data(), ctx->aCnt, sizeof(int) * n, hipMemcpyDeviceToHost));
    MPS_HIP(hipMemcpy(canon.data(), ctx->I[ICANON], sizeof(int) * n, hipMemcpyDeviceToHost));
    MPS_HIP(hipMemcpy(code.data(), ctx->aCol, sizeof(unsigned short) * ne, hipMemcpyDeviceToHost));
    MPS_HIP(hipMemcpy(esc.data(), ctx->aEsc, sizeof(int) * (size_t)kPpeEsc * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < n; k++) {   // decode the column codes
        int j = 0, e = 0;
        for (int c = 0; c < cnt[k]; c++) {
            const unsigned d = code[ellIdx4(c, k, KE)];
            j = d ? j + (int)d : esc[(size_t)(e++) * n + k];
            col[(size_t)c * n + k] = j;
        }
    }
    MPS_HIP(hipMemcpy(val.data(), ctx->aVal, sizeof(double) * ne, hipMemcpyDeviceToHost));
    MPS_HIP(hipMemcpy(diag.data(), ctx->aDiag, sizeof(double) * n, hipMemcpyDeviceToHost));
    const int nl = ctx->nLiveHost;
    std::vector<int> slot(n, -1);
    for (int k = 0; k < n; k++)
        if (canon[k] >= 0 && canon[k] < n) slot[canon[k]] = k;
    int nnz = 0;
    for (int r = 0; r < nl; r++) nnz += cnt[slot[r]] + 1;
    if (nnz_out) *nnz_out = nnz;
    if (!row_ptr && !cols && !vals) return MPS_OK;
    int p = 0;
    for (int r = 0; r < nl; r++) {
        const int k = slot[r];
        if (row_ptr) row_ptr[r] = p;
        for (int c = 0; c < cnt[k]; c++) {
            if (cols) cols[p] = canon[col[(size_t)c * n + k]];
            if (vals) vals[p] = val[ellIdx64(c, k, KE)];
            p++;
        }
        if (cols) cols[p] = r;
        if (vals) vals[p] = diag[k];
        p++;
    }
    if (row_ptr) row_ptr[nl] = p;
    return MPS_OK;
}

// ---------------------------------------------------------- learned operator
int mps_gnn_param_count(void) { return gnnParamCount(); }

int mps_gnn_init(mps_ctx* ctx, unsigned long long seed) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    int rc = gnnEnsure(ctx);
    if (rc) return rc;
    return gnnSetParams(ctx, nullptr, seed);
}

int mps_gnn_set_params(mps_ctx* ctx, const double* theta, int n) {
    CHECK_CTX(ctx);
    if (!theta || n != gnnParamCount()) { ctx->err = "parameter count mismatch"; return MPS_E_ARG; }
    hipSetDevice(ctx->dev);
    int rc = gnnEnsure(ctx);
    if (rc) return rc;
    return gnnSetParams(ctx, theta, 0);
}

int mps_gnn_get_params(mps_ctx* ctx, double* theta, int n) {
    CHECK_CTX(ctx);
    if (!theta || n != gnnParamCount()) { ctx->err = "parameter count mismatch"; return MPS_E_ARG; }
    hipSetDevice(ctx->dev);
    int rc = gnnEnsure(ctx);
    if (rc) return rc;
    return gnnGetParams(ctx, theta);
}

int mps_gnn_forward(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    return gnnForward(ctx);
}

int mps_gnn_graph(mps_ctx* ctx, int* indptr, int* cols, double* e, double* w, double* feat, double* diag, double* cs,
                  double* dsi, int* nedges) {
    CHECK_CTX(ctx);
    if (!ctx->gnn) { ctx->err = "call mps_gnn_forward first"; return MPS_E_ARG; }
    hipSetDevice(ctx->dev);
    return gnnDownloadGraph(ctx, indptr, cols, e, w, feat, diag, cs, dsi, nedges);
}

int mps_gnn_backward(mps_ctx* ctx, const double* r, const double* gz, double* grad, int n) {
    CHECK_CTX(ctx);
    if (!ctx->gnn || !r || !gz || !grad || n != gnnParamCount()) { ctx->err = "bad arguments"; return MPS_E_ARG; }
    hipSetDevice(ctx->dev);
    int rc = gnnUploadVectors(ctx, r, gz);
    if (rc) return rc;
    if ((rc = gnnBackward(ctx))) return rc;
    std::vector<float> g(n);
    MPS_HIP(hipMemcpyAsync(g.data(), gnnGradPtr(ctx), sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < n; k++) grad[k] = g[k];
    return MPS_OK;
}

int mps_gnn_train_step(mps_ctx* ctx, double lr, unsigned long long seed, double* loss) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    return gnnTrainStep(ctx, lr, seed, loss);
}

int mps_gnn_save(mps_ctx* ctx, const char* path) {
    CHECK_CTX(ctx);
    if (!path) return MPS_E_ARG;
    std::vector<double> th(gnnParamCount());
    int rc = mps_gnn_get_params(ctx, th.data(), (int)th.size());
    if (rc) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) { ctx->err = std::string("cannot write ") + path; return MPS_E_IO; }
    const char magic[8] = {'M', 'P', 'S', 'G', 'N', 'N', '1', 0};
    const int np = (int)th.size();
    std::vector<float> tf(th.begin(), th.end());
    const bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(&np, sizeof(int), 1, f) == 1 &&
                    fwrite(tf.data(), sizeof(float), tf.size(), f) == tf.size();
    fclose(f);
    if (!ok) { ctx->err = std::string("short write to ") + path; return MPS_E_IO; }
    return MPS_OK;
}

int mps_gnn_load(mps_ctx* ctx, const char* path) {
    CHECK_CTX(ctx);
    if (!path) return MPS_E_ARG;
    FILE* f = fopen(path, "rb");
    if (!f) { ctx->err = std::string("cannot open ") + path; return MPS_E_IO; }
    char magic[8];
    int np = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "MPSGNN1", 8) == 0 && fread(&np, sizeof(int), 1, f) == 1 &&
              np == gnnParamCount();
    std::vector<float> tf(ok ? np : 0);
    if (ok) ok = fread(tf.data(), sizeof(float), tf.size(), f) == tf.size();
    fclose(f);
    if (!ok) { ctx->err = std::string("not a parameter file of this network: ") + path; return MPS_E_IO; }
    std::vector<double> th(tf.begin(), tf.end());
    return mps_gnn_set_params(ctx, th.data(), np);
}

// ------------------------------------------------------------ checkpoint
// File layout (little-endian): CkptHeader, the parameter block
// (paramsSize bytes), then every carried per-particle field in reference
// order (capacity entries each: the double fields of kCkptD, the int fields
// of kCkptI, and with the granular phase kNNCarried and IPTYPE), then the
// node and face-normal arrays of the deformable and forced meshes.
namespace {
constexpr char kCkptMagic[8] = {'M', 'P', 'S', 'C', 'K', 'P', 'T', '1'};
constexpr uint32_t kCkptVersion = 1;
struct CkptHeader {
    char magic[8];
    uint32_t version, paramsSize;
    int32_t capacity, nLive, numIter, ppeSolves;
    double time, velMax;
    int64_t solverFailures;
    int32_t nD, nI;                  // per-particle double / int fields that follow
    int32_t meshNv[2], meshNf[2];    // deformable, forced mesh (0: none)
};
// persistent int fields besides the reference index, which the order implies
constexpr int kCkptI[] = {ITYPE, IBC, INNEIGH};

std::vector<int> ckptDFields(const mps_params& P) {
    std::vector<int> f;
    for (int q = 0; q < NPD; q++) f.push_back(q);
    if (P.fluid_type == 1)
        for (int q = 0; q < kNNCarriedCount; q++) f.push_back(kNNCarried[q]);
    return f;
}
std::vector<int> ckptIFields(const mps_params& P) {
    std::vector<int> f(std::begin(kCkptI), std::end(kCkptI));
    if (P.fluid_type == 1) f.push_back(IPTYPE);
    return f;
}
// parameters that fix the context's allocations and bucket grid
bool ckptCompatible(const mps_params& a, const mps_params& b, std::string& why) {
    char msg[160];
#define SAME(f)                                                                        \
    if (a.f != b.f) {                                                                  \
        snprintf(msg, sizeof msg, "checkpoint parameter " #f " differs from the context's"); \
        why = msg;                                                                     \
        return false;                                                                  \
    }
    SAME(dim) SAME(num_buckets_x) SAME(num_buckets_y) SAME(num_buckets_z) SAME(fluid_type) SAME(wall_type)
    SAME(domain_type_bc) SAME(part_dist) SAME(re_s) SAME(re_l) SAME(bucket_side)
#undef SAME
    return true;
}
}  // namespace

int mps_checkpoint_write(mps_ctx* ctx, const char* path) {
    CHECK_CTX(ctx);
    if (!path) return MPS_E_ARG;
    int rc = syncStatus(ctx);
    if (rc) return rc;
    MPS_HIP(hipSetDevice(ctx->dev));
    const int n = ctx->N;
    CkptHeader h{};
    memcpy(h.magic, kCkptMagic, 8);
    h.version = kCkptVersion;
    h.paramsSize = sizeof(mps_params);
    h.capacity = n;
    h.nLive = ctx->nLiveHost;
    h.numIter = ctx->numIter;
    h.ppeSolves = ctx->ppeSolves;
    h.time = ctx->timeCurrent;
    MPS_HIP(hipMemcpy(&h.velMax, ctx->ds + S_VELMAX, sizeof(double), hipMemcpyDeviceToHost));
    h.solverFailures = ctx->solverFailures;
    const std::vector<int> df = ckptDFields(ctx->P), itf = ckptIFields(ctx->P);
    h.nD = (int)df.size();
    h.nI = (int)itf.size();
    for (int t = 0; t < 2; t++) {
        const MeshSlot& m = ctx->extraMesh[t];
        h.meshNv[t] = m.on ? m.nv : 0;
        h.meshNf[t] = m.on ? m.nf : 0;
    }
    FILE* fp = fopen(path, "wb");
    if (!fp) { ctx->err = std::string("cannot write ") + path; return MPS_E_IO; }
    bool ok = fwrite(&h, sizeof h, 1, fp) == 1 && fwrite(&ctx->P, sizeof(mps_params), 1, fp) == 1;
    double* dtmp = nullptr;
    int* itmp = nullptr;
    std::vector<double> hd(n);
    std::vector<int> hi(n);
    hipError_t e = hipMalloc(&dtmp, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc(&itmp, sizeof(int) * n);
    const int g = gridFor(n, 256);
    for (size_t q = 0; ok && e == hipSuccess && q < df.size(); q++) {
        const double* src = ctx->D[df[q]];
        hipLaunchKernelGGL(kToCanonD, dim3(g), dim3(256), 0, ctx->stream, n, ctx->I[ICANON], src, src, src, dtmp, 1);
        e = hipMemcpyAsync(hd.data(), dtmp, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        ok = ok && fwrite(hd.data(), sizeof(double), n, fp) == (size_t)n;
    }
    for (size_t q = 0; ok && e == hipSuccess && q < itf.size(); q++) {
        hipLaunchKernelGGL(kToCanonI, dim3(g), dim3(256), 0, ctx->stream, n, ctx->I[ICANON], ctx->I[itf[q]], itmp);
        e = hipMemcpyAsync(hi.data(), itmp, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        ok = ok && fwrite(hi.data(), sizeof(int), n, fp) == (size_t)n;
    }
    for (int t = 0; ok && t < 2; t++) {
        const MeshSlot& m = ctx->extraMesh[t];
        if (!m.on) continue;
        ok = fwrite(m.hV.data(), sizeof(double), m.hV.size(), fp) == m.hV.size() &&
             fwrite(m.hN.data(), sizeof(double), m.hN.size(), fp) == m.hN.size();
    }
    if (dtmp) hipFree(dtmp);
    if (itmp) hipFree(itmp);
    ok = (fclose(fp) == 0) && ok;
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); return MPS_E_HIP; }
    if (!ok) { ctx->err = std::string("write failed: ") + path; return MPS_E_IO; }
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

int mps_checkpoint_read(mps_ctx* ctx, const char* path) {
    CHECK_CTX(ctx);
    if (!path) return MPS_E_ARG;
    MPS_HIP(hipSetDevice(ctx->dev));
    FILE* fp = fopen(path, "rb");
    if (!fp) { ctx->err = std::string("cannot open ") + path; return MPS_E_IO; }
    CkptHeader h{};
    mps_params P{};
    auto fail = [&](int code, const std::string& msg) {
        fclose(fp);
        ctx->err = msg;
        return code;
    };
    if (fread(&h, sizeof h, 1, fp) != 1 || memcmp(h.magic, kCkptMagic, 8) != 0)
        return fail(MPS_E_IO, std::string("not a checkpoint file: ") + path);
    if (h.version != kCkptVersion || h.paramsSize != sizeof(mps_params))
        return fail(MPS_E_IO, "checkpoint written by another version of the library");
    if (fread(&P, sizeof P, 1, fp) != 1) return fail(MPS_E_IO, "truncated checkpoint");
    const int n = ctx->N;
    if (h.capacity != n) return fail(MPS_E_ARG, "checkpoint capacity differs from the context's particle count");
    if (h.nLive < 0 || h.nLive > n) return fail(MPS_E_IO, "corrupt checkpoint (live count)");
    std::string why;
    if (!ckptCompatible(P, ctx->P, why)) return fail(MPS_E_ARG, why);
    const std::vector<int> df = ckptDFields(P), itf = ckptIFields(P);
    if (h.nD != (int)df.size() || h.nI != (int)itf.size()) return fail(MPS_E_IO, "corrupt checkpoint (field count)");
    for (int t = 0; t < 2; t++) {
        const MeshSlot& m = ctx->extraMesh[t];
        if ((h.meshNv[t] > 0) != m.on || (m.on && (h.meshNv[t] != m.nv || h.meshNf[t] != m.nf)))
            return fail(MPS_E_ARG, "checkpoint meshes differ from the context's: install the same deformable / forced "
                                   "meshes before reading");
    }
    // read everything on the host first: a short file leaves the context as it was
    std::vector<std::vector<double>> hd(df.size(), std::vector<double>(n));
    std::vector<std::vector<int>> hi(itf.size(), std::vector<int>(n));
    for (auto& v : hd)
        if (fread(v.data(), sizeof(double), n, fp) != (size_t)n) return fail(MPS_E_IO, "truncated checkpoint");
    for (auto& v : hi)
        if (fread(v.data(), sizeof(int), n, fp) != (size_t)n) return fail(MPS_E_IO, "truncated checkpoint");
    std::vector<double> mv[2], mn[2];
    for (int t = 0; t < 2; t++) {
        if (h.meshNv[t] == 0) continue;
        mv[t].resize(3 * (size_t)h.meshNv[t]);
        mn[t].resize(3 * (size_t)h.meshNf[t]);
        if (fread(mv[t].data(), sizeof(double), mv[t].size(), fp) != mv[t].size() ||
            fread(mn[t].data(), sizeof(double), mn[t].size(), fp) != mn[t].size())
            return fail(MPS_E_IO, "truncated checkpoint");
    }
    fclose(fp);
    int rc = vtuWait(ctx);   // snapshots in flight read the old state
    if (rc) return rc;
    ctx->P = P;
    // slot k holds the particle of reference index k; the per-step fields
    // start as in a new context
    KArgs a = ctx->args();
    hipLaunchKernelGGL(kInitState, dim3(gridFor(n, 256)), dim3(256), 0, ctx->stream, a);
    for (size_t q = 0; q < df.size(); q++)
        MPS_HIP(hipMemcpyAsync(ctx->D[df[q]], hd[q].data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    for (size_t q = 0; q < itf.size(); q++)
        MPS_HIP(hipMemcpyAsync(ctx->I[itf[q]], hi[q].data(), sizeof(int) * n, hipMemcpyHostToDevice, ctx->stream));
    MPS_HIP(hipMemsetAsync(ctx->is, 0, sizeof(int) * NIS, ctx->stream));
    MPS_HIP(hipMemsetAsync(ctx->ds, 0, sizeof(double) * NDS, ctx->stream));
    MPS_HIP(hipMemcpyAsync(ctx->is + I_NLIVE, &h.nLive, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    MPS_HIP(hipMemcpyAsync(ctx->ds + S_VELMAX, &h.velMax, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    MPS_HIP(hipStreamSynchronize(ctx->stream));
    for (int t = 0; t < 2; t++) {
        MeshSlot& m = ctx->extraMesh[t];
        if (!m.on) continue;
        m.hV = mv[t];
        m.hN = mn[t];
        rc = uploadSlot(ctx, m, ctx->bvhMode != 0);
        if (rc) return rc;
    }
    ctx->nLiveHost = h.nLive;
    ctx->velMaxHost = h.velMax;
    ctx->timeCurrent = h.time;
    ctx->numIter = h.numIter;
    ctx->ppeSolves = h.ppeSolves;
    ctx->solverFailures = h.solverFailures;
    ctx->solverFailed = 0;
    ctx->listValid = false;
    ctx->p4Flags = false;
    ctx->asmMode = -1;
    ctx->r0Fresh = false;
    MPS_HIP(hipGetLastError());
    return MPS_OK;
}

int mps_init(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    return runInit(ctx);
}

int mps_step(mps_ctx* ctx) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    return runStep(ctx);
}

int mps_steps(mps_ctx* ctx, int k) {
    CHECK_CTX(ctx);
    hipSetDevice(ctx->dev);
    int soft = MPS_OK;
    std::string msg;
    for (int t = 0; t < k; t++) {
        int rc = runStep(ctx);
        if (softSolverFailure(ctx, rc)) {
            if (soft == MPS_OK) msg = "step " + std::to_string(ctx->numIter) + ": " + ctx->err;
            soft = rc;
        } else if (rc) {
            return rc;
        }
    }
    if (soft != MPS_OK) ctx->err = msg;
    return soft;
}

}  // extern "C"
