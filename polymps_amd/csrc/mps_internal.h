// Internal declarations of the MI355X MPS library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "mps_gpu.h"

namespace mpsgpu {

void setGlobalError(const std::string& e);
void paramsDefaults(mps_params* p);
int paramsFromJson(const char* path, mps_params* p, std::string& grid, std::string& mesh, std::string& err);
int paramsDerive(mps_params* p, int n, const double* pos, std::string& err);

// ---------------------------------------------------------------- fields
// Per-particle double fields.  The first NPD are persistent across steps and
// are gathered into bucket order by every sort; the rest are recomputed
// inside a step before they are read.
enum DField {
    // persistent
    // (pndki, div_u, diffusive and grad_conc are written for a subset of the
    // particles in some modes, so they are carried with their particle)
    DX, DY, DZ, DU, DV, DW, DMX, DMY, DMZ, DPRESS, DPNDI, DPPEX, DPAVG, DFWX, DFWY, DFWZ, DPNDSKI,
    DPNDKI, DDIVU, DDIFF, DGCX, DGCY, DGCZ, DPPEX1, DPPEX2, DPPEX3,   // the three solutions before ppe_x
    NPD,
    // per step
    DAX = NPD, DAY, DAZ, DASX, DASY, DASZ, DNPX, DNPY, DNPZ, DNPCD2, DPNDSMALL, DCONC,
    DAUX1, DAUX2, DPNDWALL, DRIW2, DDEVN, DWPX, DWPY, DWPZ, DPNX, DPNY, DPNZ,
    DCM0, DCM1, DCM2, DCM3, DCM4, DCM5, DCM6, DCM7, DCM8, DDVX, DDVY, DDVZ, DPPEB, DNRMX, DNRMY, DNRMZ,
    // granular phase (fluid_type 1): density and viscosity per particle (a
    // Newtonian fluid uses DNS_FL1 and KNM_VS1 DNS_FL1 for every particle),
    // volume fraction, second strain-rate invariant, yield viscosity,
    // inertial number, normal stress, smoothed pressure, grain volume
    // fraction, strain-rate tensor
    DRHO, DMEU, DCV, DII, DMEUY, DINERTIA, DPRHEO, DPSMOOTH, DVF, DS11, DS12, DS13, DS22, DS23, DS33,
    ND
};
// per-step fields that a granular-phase run carries with its particles
// through the sort (the correction matrix is read at the start of the next
// step by the strain rate)
constexpr int kNNCarried[] = {DRHO, DMEU, DCM0, DCM1, DCM2, DCM3, DCM4, DCM5, DCM6, DCM7, DCM8};
constexpr int kNNCarriedCount = sizeof(kNNCarried) / sizeof(kNNCarried[0]);
// Per-particle int fields; the first NPI are persistent.
enum IField { ITYPE, ICANON, IBC, INNEIGH, NPI, INNWALL = NPI, INEARWALL, INEARMESH, IPTYPE, NI };

// device scalar slots (double)
enum DScalar { S_VELMAX, S_ABSNEW, S_ABSOLD, S_RHSN2, S_THR, S_RN2, S_ALPHA, S_BETA, S_PAP, S_BN2, S_TRUE_RN2,
               S_RELTOL, S_RHO, S_RHOOLD, S_BW, S_R0N2,
               S_NNZ,   // PPE entries of the last assembly, counted on the device when profiling
               NDS };
// device scalar slots (int)
enum IScalar { I_NLIVE, I_OUTCNT, I_OVERFLOW, I_CG_ITER, I_CG_DONE, I_COMPACT_OVF, I_NESC,
               I_BCG_RESTART, I_BCG_RESTARTS, I_REBUILD, I_PATCHCNT, I_FIXCNT, I_NLCESC, NIS };

constexpr int kWallTable = 50;
// escaped (non-delta) columns per PPE row: the first entry and stencil-row
// jumps of 65536 slots or more, at most 1 + 8
constexpr int kPpeEsc = 12;
// Window index escapes per PPE row (one per stencil row the row's entries
// enter, usually nine, and per step of 256 or more within one); a block with
// a row that needs more gathers from global memory
constexpr int kLcEsc = 16;
constexpr int kPpeCodePad = 16;   // ELL entries per row past K (one batch the row kernels read ahead)
constexpr int kRedBlocks = 1024;   // partial-sum slots of a reduction

// Bounding-volume hierarchy over the wall triangles: a leaf holds the
// triangles bvhTri[a .. a - b - 1] (b < 0); an inner node has children a, b.
struct BvhNode {
    double lo[3], hi[3];
    int a, b;
};

// Everything a kernel may touch, passed by value.
struct KArgs {
    mps_params P;
    int N;                 // slots (capacity)
    int nB;                // buckets
    double* D[ND];
    int* I[NI];
    const int* cellStart;  // nB + 2 (last entry = N)
    const unsigned* nbrMask;  // candidate bitmasks [rows][maskW][N], see kBuildList
    int maskW;                // 32-bit words per stencil row
    const int* nbr;           // neighbour list, column-major [K][N] (decoded from the bitmasks)
    const int* nbrCnt;
    const int* nbrS;          // entries within the small radius (re_s, collision), column-major [KS][N]
    const int* nbrCntS;
    double4* P4;              // (x, y, z, w) per slot, written by the list build; valid with the list.
                              // w = type, or after update_bc (p4Flags) type & 1023 | inner << 10 | far-from-wall << 11
    int p4Flags;
    const double4* V4;        // (u, v, w, 0) per slot, packed for the list-based viscosity
    int K;                    // PPE row capacity (ELL width)
    int KS;                   // small-radius list capacity
    const double* mV;      // mesh
    const int* mF;
    const double* mN;
    int nf;
    const BvhNode* bvh;    // nullptr: brute-force closest point
    const int* bvhTri;
    const double* wtab;    // xPND[50] xNeigh[50] Zij[50] nNeigh[50] (as double)
    double* ds;            // device scalars
    int* is;
    double* red;           // reduction partials: up to three per block, then the vel-max slot
    double volForce;       // pow(l0, dim) evaluated on the host, as the reference's CPU build does
};

struct Gnn;
struct VtuQueue;

// A wall mesh besides the rigid one (deformable or forced): device arrays,
// its BVH and the host copy the forced motion updates.
struct MeshSlot {
    bool on = false;
    int nv = 0, nf = 0;
    double* V = nullptr;
    int* F = nullptr;
    double* N = nullptr;
    BvhNode* bvh = nullptr;
    int* bvhTri = nullptr;
    int bvhNodes = 0;
    std::vector<double> hV, hN;
    std::vector<int> hF;
};


struct Ctx {
    mps_params P;
    int dev = 0;
    hipStream_t stream = nullptr;
    int N = 0;            // slots = initial particle count
    int nLiveHost = 0;
    int nB = 0;
    double* D[ND] = {};
    double* Dalt[NPD] = {};
    int* I[NI] = {};
    int* Ialt[NPI] = {};
    // granular phase: second buffers of the carried per-step fields and of
    // PTYPE, and the bed-level grid (allocated with fluid_type 1 only)
    double* DaltNN[kNNCarriedCount] = {};
    int* IaltPtype = nullptr;
    unsigned long long* bedKeys = nullptr;   // [2][cells]: ordered keys of the bed and water levels
    int* bedIdx = nullptr;                   // [cells]: reference index of the bed-level particle
    int bedKx = 0, bedKy = 0;
    double cflVisc = 0.0;
    // buckets
    int* cellStart = nullptr;   // nB + 2
    int* cellCount = nullptr;   // nB + 1
    int* key = nullptr;         // N
    int* rank = nullptr;        // N
    int* order = nullptr;       // N
    int* order2 = nullptr;      // N
    int* slotOf = nullptr;      // N (canonical -> slot)
    int* scanTmp = nullptr;
    int* outList = nullptr;     // compaction candidates
    int* patchList = nullptr;   // N: particles whose kept list is rebuilt alone (allocated on first use)
    int* fixList = nullptr;     // N: particles whose candidate rows overflow the bitmask windows
    // neighbour list
    unsigned* nbrMask = nullptr;
    int maskW = 0;
    int maskWAlloc = 0;         // words per stencil row the masks were allocated with
    int* nbr = nullptr;
    int* nbrCnt = nullptr;
    unsigned* nbrMaskS = nullptr;   // small-radius bitmasks, list and counts
    int* nbrS = nullptr;
    int* nbrCntS = nullptr;
    int KS = 0;
    double4* P4 = nullptr;
    bool p4Flags = false;       // P4.w carries the inner / far-from-wall flags of update_bc
    double4* V4 = nullptr;      // packed velocities (viscosity with the list; allocated on first use)
    int K = 0;
    bool listValid = false;
    // list kept over the time step: skin in units of l0 (< 0 automatic from
    // the last velocity maximum, 0 off) and the skin of the last step (m)
    double listSkin = -1.0;
    double lastSkin = 0.0;
    bool lastStartList = false;   // the last step built a list at its start positions (no kept list)
    // largest automatic skin (l0); beyond it the step builds its list after
    // the prediction and the viscosity walks the buckets
    double listSkinCap = 0.08;
    // automatic skin: this many times the largest displacement of one step
    double listSkinFactor = 3.0;
    // mesh
    double* mV = nullptr;
    int* mF = nullptr;
    double* mN = nullptr;
    int nv = 0, nf = 0;
    BvhNode* bvh = nullptr;
    int* bvhTri = nullptr;
    int bvhNodes = 0;
    int bvhMode = -1;            // -1: BVH for meshes of more than 64 faces, 0: brute force, 1: BVH
    double* wtab = nullptr;
    std::vector<double> hxPND, hxNeigh, hZij;
    std::vector<int> hnNeigh;
    bool hasMesh = false;
    std::vector<double> rigidV, rigidN;   // host copy of the rigid mesh
    std::vector<int> rigidF;
    MeshSlot extraMesh[2];                // meshType 1 (deformable) and 2 (forced)
    // PPE
    double* aVal = nullptr;     // ellIdx64 (64-row tiles of ellWidth(K) entries per row)
    unsigned short* aCol = nullptr;   // ellIdx4: column codes, slot delta from the previous entry, 0 = escape
    int* aEsc = nullptr;              // [kPpeEsc][N] absolute slots of the escaped entries
    int* aCnt = nullptr;        // N
    double* aDiag = nullptr;    // N
    float* aVal32 = nullptr;    // ellIdx4 (mixed precision)
    unsigned char* aLc8 = nullptr;    // ELL, groups of 8: one-byte window index codes (step or 0 = escape)
    unsigned short* aLcEsc = nullptr; // ELL, kLcEsc per row: the escapes' absolute window indices
    int* aWin = nullptr;              // per 256-row block: the LDS window table of the mixed SpMV
    bool winValid = false;            // the last assembly wrote windows
    bool ppeWindow = true;            // scalar ppe_window: build and use the windows
    float* aDiagInv32 = nullptr;
    float* aDiag32 = nullptr;
    double* cgR = nullptr;
    double* cgP = nullptr;
    double* cgZ = nullptr;
    double* cgQ = nullptr;
    double* cgInvD = nullptr;
    double* cgX = nullptr;      // refinement correction
    double* bcgR0 = nullptr;    // BiCGSTAB shadow residual, y, s, t (allocated on first use)
    double* bcgY = nullptr;
    double* bcgS = nullptr;
    double* bcgT = nullptr;
    float* fR = nullptr;
    float* fP = nullptr;
    float* fZ = nullptr;
    float* fQ = nullptr;
    double* fX = nullptr;   // the inner solve's correction, summed in FP64
    // scalars
    double* ds = nullptr;
    int* is = nullptr;
    double* red = nullptr;
    double* hostScalars = nullptr;   // pinned
    int* hostInts = nullptr;         // pinned
    // host mirrors / stats
    double velMaxHost = 0.0;
    double timeCurrent = 0.0;
    int numIter = 0;
    int solverIter = 0;
    int outerIter = 0;
    Gnn* gnn = nullptr;          // learned PPE operator (created on first use)
    int redCap = 0;              // partial-sum slots in red (pairs)
    int ppeSolves = 0;           // pressure solves so far (history of the extrapolated warm start)
    int lastCgIters = 0;         // iterations of the last FP64 CG solve / first mixed inner pass
    int lastInnerIters = 0;
    int lastPass2Iters = 8;          // iterations of the last solve's second refinement pass
    bool r0Fresh = false;        // cgR and asmRed hold b - A x of the current guess and its partials (assembly)
    bool r0FreshMixed = false;   // ... in the mixed solve's partial order (else the FP64 CG's, with 1/diag)
    double* asmRed = nullptr;    // block partials of the assembly's first residual (no other kernel writes them)
    unsigned long long* red64 = nullptr;   // result of reduceCounts (allocated on first use)
    // solver kind of the last PPE assembly (MPS_SOLVER_* / precision), -1 when
    // there is none for the current slot order; a solve must match it
    int asmMode = -1;
    // outcome of the last pressure solve: 0 converged, 1 stopped at an
    // iteration limit above the tolerance, 2 non-finite residual
    int solverFailed = 0;
    long solverFailures = 0;     // solves that did not converge since the count was last reset
    int lastDone = 0;            // the solver's final device status (5: a non-finite residual or product)
    double solverErr = 0.0;
    double trueResidual = 0.0;
    double ppeMs = 0.0;
    bool ppeMsPending = false;   // ppeMs still to be read from the solve's events (ev0, ev1)
    int fallbackUsed = 0;
    double learnedStallRatio = 0.5;   // learned preconditioner: largest residual ratio of a refinement pass
    double learnedTolScale = 0.1;     // learned preconditioner: true-residual target as a multiple of the tolerance
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // optional per-launch timing of the SpMV kernel (HIP events on the
    // context stream); enabled with mps_set_scalar(ctx, "profile", 1)
    bool profile = false;
    std::vector<hipEvent_t> evPool;
    double spmvMsTotal = 0.0;
    long spmvLaunches = 0;
    double spmvBytesTotal = 0.0;
    std::string err;
    // asynchronous VTU output (output.hip)
    VtuQueue* vtu = nullptr;

    KArgs args() const;
};

#define MPS_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess) {                                                         \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(_e);               \
            return MPS_E_HIP;                                                           \
        }                                                                               \
    } while (0)

inline int gridFor(int n, int block) { return (n + block - 1) / block; }

// launchers (defined in the .hip translation units)
int launchBucketSort(Ctx* ctx);
int launchBuildList(Ctx* ctx, double skin = 0.0, const int* gate = nullptr, bool small = true);
int launchListKeepStart(Ctx* ctx, double skin);
int launchListKeepRefresh(Ctx* ctx, double skin);
int launchOp(Ctx* ctx, const char* op);
int ppeAssemble(Ctx* ctx);
int ppeSolve(Ctx* ctx);
int ppeWarmStart(Ctx* ctx);
int launchCompaction(Ctx* ctx);
// sum (max = false) or maximum of the n counts cnt[] (synchronises the stream)
int ppeWindowBlocks(Ctx* ctx, double* out);
int reduceCounts(Ctx* ctx, const int* cnt, int n, bool max, double* out);
int launchCountToScalar(Ctx* ctx, const int* cnt, int n, int slot);
int ppeSolveMs(Ctx* ctx, double* out);
int launchClosestPoint(Ctx* ctx, int type);   // one mesh type (0 rigid, 1 deformable, 2 forced)
// granular phase (granular.hip)
void granularGridSize(const mps_params& P, int* kx, int* ky);
int launchInitPhase(Ctx* ctx);
int launchViscosityNN(Ctx* ctx);
int launchGranular(Ctx* ctx, const char* op);
int granularCflVisc(Ctx* ctx, double* out);
// learned PPE operator (gnn.hip)
int gnnEnsure(Ctx* ctx);
int gnnSetParams(Ctx* ctx, const double* theta, unsigned long long seed);
int gnnGetParams(Ctx* ctx, double* theta);
int gnnParamCount();
int gnnForward(Ctx* ctx);
int gnnBackward(Ctx* ctx);
int gnnTrainStep(Ctx* ctx, double lr, unsigned long long seed, double* loss);
int gnnWarmStart(Ctx* ctx);
int gnnApply32(Ctx* ctx, const float* r, float* z);
int gnnDownloadGraph(Ctx* ctx, int* indptr, int* cols, double* e, double* w, double* feat, double* diag,
                     double* cs, double* dsi, int* nedges);
int gnnUploadVectors(Ctx* ctx, const double* r, const double* gz);
float* gnnGradPtr(Ctx* ctx);
void gnnFree(Ctx* ctx);
// VTU output (output.hip): snapshot in stream order, file written by a
// background thread; vtuWait joins every pending writer.
int vtuSubmit(Ctx* ctx, const char* path, int flags);
int vtuWait(Ctx* ctx);
void vtuFree(Ctx* ctx);

}  // namespace mpsgpu
