// polymps_gpu.hpp -- the PolyMPS operator classes over the MI355X C ABI.
//
// A program written against the reference's classes (MpsParticleSystem,
// MpsParticle, MpsBucket, MpsPndNeigh, MpsPressure, MpsViscosity,
// MpsParticleCollision, MpsShifting, MpsBoundaryCondition, MpsVectorMatrix,
// MpsParticleVelPos, MpsInputOutput, PolygonMesh) compiles against this
// header instead and runs every operator on the GPU.  Method names and
// signatures are those of include/Mps*.h and include/PolygonMesh.h of the
// reference; each body forwards to one call of include/mps_gpu.h.
//
// Particle state lives on the device, owned by the mps_ctx that
// MpsBucket::allocateMemory creates (it is the first call of the reference's
// start-up that knows both the parameters and the particle count).  The
// array members of MpsParticle that the reference's main() passes to the
// PolygonMesh calls are kept only so that those calls compile; the mesh
// operators work on the device state of the context in use.
//
// Errors: the reference prints and throws std::exception or exits; here a
// failed call throws MpsGpuError (derived from std::exception) carrying the
// library's message.
#ifndef POLYMPS_GPU_HPP_
#define POLYMPS_GPU_HPP_

#include <sys/stat.h>
#include <sys/time.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "mps_gpu.h"

// Enumerations of include/MpsParticleSystem.h:20-119 (same names and values,
// so that code comparing against them compiles unchanged).
enum boundaryWallType { PARTICLE = 0, POLYGON = 1 };
enum partType { FLUID = 0, WALL = 1, DUMMY_WALL = 2 };
enum calcPressType { EXPLICIT = 0, WEAKLY = 1, IMPLICIT_PND = 2, IMPLICIT_PND_DIVU = 3 };
enum solvPressType { CG = 0, BICGSTAB = 1 };
enum calcPNDType { SUM_WIJ = 0, MEAN_SUM_WIJ = 1, DIFFUSIVE = 2 };
enum calcBCType { PND_NEIGH = 0, PND_NPCD = 1, PND_ARC = 2 };
enum slipBC { NO_SLIP = 0, FREE_SLIP = 1 };
enum repForceType { HARADA = 0, MITSUME = 1, LENNARD_JONES = 2, MONAGHAN_KAJTAR = 3 };
enum meshType { FIXED = 0, DEFORMABLE = 1, FORCED = 2 };
enum viscType { NEWTONIAN = 0, NON_NEWTONIAN = 1 };
enum colType { PC = 0, DPC = 1 };
enum domainBC { NONE_BC = 0, PERIODIC = 1 };

class MpsGpuError : public std::exception {
public:
    explicit MpsGpuError(std::string m) : msg_(std::move(m)) {}
    const char* what() const noexcept override { return msg_.c_str(); }

private:
    std::string msg_;
};

namespace polymps_gpu {

// The context that the mesh operators act on (the reference's PolygonMesh
// methods receive raw arrays, not the particle object).
inline mps_ctx*& currentContext() {
    static mps_ctx* ctx = nullptr;
    return ctx;
}

inline void check(int rc, const mps_ctx* ctx, const char* what) {
    if (rc == MPS_OK) return;
    std::string m = std::string(what) + " failed (" + std::to_string(rc) + "): " + mps_last_error(ctx);
    throw MpsGpuError(m);
}

inline double wallClock() {
    struct timeval t;
    gettimeofday(&t, nullptr);
    return (double)t.tv_sec + (double)t.tv_usec * 1e-6;
}

}  // namespace polymps_gpu

// include/MpsParticleSystem.h:125-316: the parameters shared by all
// particles.  The full parameter block is `prm`; the members main() reads
// are mirrored under the reference's names and refreshed by syncFromParams().
class MpsParticleSystem {
public:
    mps_params prm;
    int wallType = 1, numOfMeshs = 0, mpsType = 0, freeSurfType = 0, fluidType = 0, collisionType = 0;
    int pndType = 0, slipCondition = 0, shiftingType = 0, weightType = 0, fluid = 0, iterOutput = 80;
    int numOfIterations = 0, fileNumber = 0, gradientCorrection = 0;
    bool forcedOn = false, femOn = false;
    double relaxPress = 1.0, timeCurrent = 0.0, timeStep = 0.0, timeSimulation = 0.0, dim = 3.0, partDist = 0.0;
    double reS = 0.0, reL = 0.0, reS2 = 0.0, reL2 = 0.0, velMax = 0.0, CFLcurrent = 0.0, CFLvisc = 0.0;
    double timerStart = 0.0, timerEnd = 0.0;
    double uniformVelWall[3] = {0.0, 0.0, 0.0};

    MpsParticleSystem() { mps_params_defaults(&prm); syncFromParams(); }

    void syncFromParams() {
        wallType = prm.wall_type;
        numOfMeshs = (prm.wall_type == POLYGON) ? 1 + (prm.fem_on ? 1 : 0) + (prm.forced_on ? 1 : 0) : 0;
        mpsType = prm.mps_type;
        freeSurfType = prm.free_surf_type;
        fluidType = prm.fluid_type;
        collisionType = prm.collision_type;
        pndType = prm.pnd_type;
        slipCondition = prm.slip_condition;
        shiftingType = prm.shifting_type;
        weightType = prm.weight_type;
        fluid = prm.fluid;
        iterOutput = prm.iter_output;
        gradientCorrection = prm.gradient_correction;
        forcedOn = prm.forced_on != 0;
        femOn = prm.fem_on != 0;
        relaxPress = prm.relax_press;
        timeStep = prm.time_step;
        timeSimulation = prm.time_simulation;
        dim = prm.dim;
        partDist = prm.part_dist;
        reS = prm.re_s;
        reL = prm.re_l;
        reS2 = prm.re_s2;
        reL2 = prm.re_l2;
    }
};

// include/MpsParticle.h:20-221.  Owns the device context once created.
class MpsParticle {
public:
    int numParticles = 0;
    mps_ctx* ctx = nullptr;
    // Host copy of the .grid state until the context exists.
    std::vector<int> hostType;
    std::vector<double> hostPos, hostVel, hostPress, hostPressAvg;
    // Present so that the reference's PolygonMesh calls compile; the device
    // state is what the mesh operators use.
    int* particleType = nullptr;
    double* pos = nullptr;
    double* particleAtWallPos = nullptr;
    double* mirrorParticlePos = nullptr;
    double* distParticleWall2 = nullptr;
    int* elementID = nullptr;
    int* nearMeshType = nullptr;
    double* polygonNormal = nullptr;
    double* pndWallContribution = nullptr;
    double* numNeighWallContribution = nullptr;
    bool* particleNearWall = nullptr;

    MpsParticle() = default;
    MpsParticle(const MpsParticle&) = delete;
    MpsParticle& operator=(const MpsParticle&) = delete;
    ~MpsParticle() {
        if (ctx) {
            if (polymps_gpu::currentContext() == ctx) polymps_gpu::currentContext() = nullptr;
            mps_destroy(ctx);
        }
    }

    // MpsParticle::setParameters (src/MpsParticle.cpp:136-227); the derived
    // constants are computed with the bucket grid in MpsBucket::allocateMemory.
    void setParameters(MpsParticleSystem* PSystem) { PSystem->syncFromParams(); }

    // MpsParticle::addGravity (src/MpsParticle.cpp:111-130)
    void addGravity(MpsParticleSystem* PSystem, MpsParticle* Particles) {
        (void)PSystem;
        polymps_gpu::check(mps_add_gravity(Particles->ctx), Particles->ctx, "addGravity");
    }
    // MpsParticle::checkParticleOutDomain (src/MpsParticle.cpp:234-402)
    void checkParticleOutDomain(MpsParticleSystem* PSystem) {
        (void)PSystem;
        polymps_gpu::check(mps_check_out_domain(ctx), ctx, "checkParticleOutDomain");
    }

    // Download helpers (reference order)
    std::vector<double> field(const char* name, int comps) const {
        std::vector<double> v((size_t)capacity() * comps);
        polymps_gpu::check(mps_download(ctx, name, v.data(), v.size()), ctx, name);
        return v;
    }
    int capacity() const {
        double c = 0;
        mps_get_scalar(ctx, "capacity", &c);
        return (int)c;
    }
    void refreshCount() {
        double n = 0;
        polymps_gpu::check(mps_get_scalar(ctx, "num_particles", &n), ctx, "num_particles");
        numParticles = (int)n;
    }
};

// include/MpsBucket.h
class MpsBucket {
public:
    // MpsBucket::allocateMemory (src/MpsBucket.cpp:20-67) with calcDomainLimits:
    // derives the bucket grid, creates the device context and uploads the state.
    void allocateMemory(MpsParticleSystem* PSystem, MpsParticle* Particles) {
        const int n = Particles->numParticles;
        polymps_gpu::check(mps_params_derive(&PSystem->prm, n, Particles->hostPos.data(), Particles->hostType.data()),
                           nullptr, "allocateMemory (parameters)");
        PSystem->syncFromParams();
        polymps_gpu::check(mps_create(&Particles->ctx, &PSystem->prm, n, 0), nullptr, "allocateMemory (device)");
        polymps_gpu::check(mps_upload_state(Particles->ctx, Particles->hostType.data(), Particles->hostPos.data(),
                                            Particles->hostVel.data(), Particles->hostPress.data(),
                                            Particles->hostPressAvg.data()),
                           Particles->ctx, "allocateMemory (upload)");
        polymps_gpu::currentContext() = Particles->ctx;
    }
    // MpsBucket::updateParticlesID (src/MpsBucket.cpp:299-353)
    void updateParticlesID(MpsParticleSystem* PSystem, MpsParticle* Particles) {
        (void)PSystem;
        polymps_gpu::check(mps_bucket_update(Particles->ctx), Particles->ctx, "updateParticlesID");
    }
};

// include/MpsBoundaryCondition.h
class MpsBoundaryCondition {
public:
    // MpsBoundaryCondition::setBucketBC (src/MpsBoundaryCondition.cpp:20-63):
    // the periodic ghost layers follow from the bucket grid that
    // mps_params_derive computes, and the device walks resolve them; nothing
    // is stored per bucket.
    void setBucketBC(MpsParticleSystem* PSystem, MpsParticle* Particles, MpsBucket* Buckets) {
        (void)PSystem; (void)Particles; (void)Buckets;
    }
    // MpsBoundaryCondition::updateParticleBC (src/MpsBoundaryCondition.cpp:66-400)
    void updateParticleBC(MpsParticleSystem* PSystem, MpsParticle* Particles, MpsBucket* Buckets) {
        (void)PSystem; (void)Buckets;
        polymps_gpu::check(mps_update_bc(Particles->ctx), Particles->ctx, "updateParticleBC");
    }
};

// include/MpsPndNeigh.h.  The diffusive-PND family and the mean-PND pair are
// single device operators; their other members are no-ops.
class MpsPndNeigh {
public:
    void setInitialPndNumberOfNeighNPCD(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_initial_pnd(P->ctx), P->ctx, "setInitialPndNumberOfNeighNPCD");
    }
    void calcWallNPCD(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_wall_npcd(P->ctx), P->ctx, "calcWallNPCD");
    }
    void calcPndnNeighNPCD(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_pnd_neigh_npcd(P->ctx), P->ctx, "calcPndnNeighNPCD");
    }
    // calcPndDiffusiveTerm + wall terms + updatePnd + meanPndParticlesWallDummySurface
    void calcPndDiffusiveTerm(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_pnd_diffusive(P->ctx), P->ctx, "calcPndDiffusiveTerm");
    }
    void calcWallSlipPndDiffusiveTerm(MpsParticleSystem*, MpsParticle*, MpsBucket*) {}
    void calcWallNoSlipPndDiffusiveTerm(MpsParticleSystem*, MpsParticle*, MpsBucket*) {}
    void updatePnd(MpsParticle*) {}
    void meanPndParticlesWallDummySurface(MpsParticleSystem*, MpsParticle*, MpsBucket*) {}
    // meanWallPnd + meanPnd
    void meanWallPnd(MpsParticleSystem*, MpsParticle*, MpsBucket*) {}
    void meanPnd(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_mean_pnd(P->ctx), P->ctx, "meanPnd");
    }
    void meanNeighFluidPnd(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_mean_neigh_fluid_pnd(P->ctx), P->ctx, "meanNeighFluidPnd");
    }
};

// include/MpsPressure.h
class MpsPressure {
public:
    void allocateMemory(MpsParticle*) {}
    // MpsPressure::calcPress (src/MpsPressure.cpp:32-58): EOS pressure or PPE
    // assembly + Jacobi-PCG + setZeroOnNegativePressure.
    // A solve that stops above its tolerance prints the reference's message
    // and the run goes on, as there (src/MpsPressure.cpp:440-442); a
    // non-finite residual throws.
    void calcPress(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        const int rc = mps_calc_press(P->ctx);
        double failed = 0;
        if (rc == MPS_E_SOLVER && mps_get_scalar(P->ctx, "solver_failed", &failed) == MPS_OK && failed == 1.0)
            fprintf(stderr, "Error: Failed solving.\n");
        else
            polymps_gpu::check(rc, P->ctx, "calcPress");
        double v = 0;
        mps_get_scalar(P->ctx, "solver_iter", &v);
        solverIter_ = (int)v;
        mps_get_scalar(P->ctx, "solver_err", &solverError_);
    }
    void extrapolatePressParticlesWallDummy(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_extrapolate_wall_dummy(P->ctx), P->ctx, "extrapolatePressParticlesWallDummy");
    }
    void predictionPressGradient(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_prediction_press_gradient(P->ctx), P->ctx, "predictionPressGradient");
    }
    void predictionWallPressGradient(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_prediction_wall_press_gradient(P->ctx), P->ctx, "predictionWallPressGradient");
    }
    void calcPressGradient(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_press_gradient(P->ctx), P->ctx, "calcPressGradient");
    }
    void calcWallPressGradient(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_wall_press_gradient(P->ctx), P->ctx, "calcWallPressGradient");
    }
    // include/MpsPressure.h:242-253
    int getSolverIter() const { return solverIter_; }
    double getSolverError() const { return solverError_; }

private:
    int solverIter_ = 0;
    double solverError_ = 0.0;
};

// include/MpsViscosity.h
class MpsViscosity {
public:
    void calcViscosity(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_viscosity(P->ctx), P->ctx, "calcViscosity");
    }
    void calcWallSlipViscosity(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_wall_viscosity(P->ctx), P->ctx, "calcWallSlipViscosity");
    }
    void calcWallNoSlipViscosity(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_wall_viscosity(P->ctx), P->ctx, "calcWallNoSlipViscosity");
    }
    // the granular phase (src/MpsViscosity.cpp:104-1526)
    void calcVolumeFraction(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_volume_fraction(P->ctx), P->ctx, "calcVolumeFraction");
    }
    // the 3-D normal stress ramps with PSystem->timeCurrent (src/MpsViscosity.cpp:596-597)
    void calcViscosityInteractionVal(MpsParticleSystem* S, MpsParticle* P, MpsBucket*) {
        mps_set_scalar(P->ctx, "time", S->timeCurrent);
        polymps_gpu::check(mps_viscosity_interaction(P->ctx), P->ctx, "calcViscosityInteractionVal");
    }
    void calcWallSlipViscosityInteractionVal(MpsParticleSystem* S, MpsParticle* P, MpsBucket*) {
        mps_set_scalar(P->ctx, "time", S->timeCurrent);
        polymps_gpu::check(mps_wall_viscosity_interaction(P->ctx), P->ctx, "calcWallSlipViscosityInteractionVal");
    }
    void calcWallNoSlipViscosityInteractionVal(MpsParticleSystem* S, MpsParticle* P, MpsBucket*) {
        mps_set_scalar(P->ctx, "time", S->timeCurrent);
        polymps_gpu::check(mps_wall_viscosity_interaction(P->ctx), P->ctx, "calcWallNoSlipViscosityInteractionVal");
    }
};

// include/MpsParticleCollision.h.  The per-step neighbour list is built here
// (positions are final for the step after the collision of the PC model and
// before the DPC model moves none of the list's consumers' inputs).
class MpsParticleCollision {
public:
    void checkParticleCollisions(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_build_neighbor_list(P->ctx), P->ctx, "neighbour list");
        polymps_gpu::check(mps_collision(P->ctx), P->ctx, "checkParticleCollisions");
    }
    void checkDynamicParticleCollisions(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_collision(P->ctx), P->ctx, "checkDynamicParticleCollisions");
        polymps_gpu::check(mps_build_neighbor_list(P->ctx), P->ctx, "neighbour list");
    }
};

// include/MpsShifting.h
class MpsShifting {
public:
    void calcShifting(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_shifting(P->ctx), P->ctx, "calcShifting");
    }
    // calcConcAndConcGradient + calcWallConcAndConcGradient
    void calcConcAndConcGradient(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_shifting(P->ctx), P->ctx, "calcConcAndConcGradient");
    }
    void calcWallConcAndConcGradient(MpsParticleSystem*, MpsParticle*, MpsBucket*) {}
    void calcNormalParticles(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_normal_particles(P->ctx), P->ctx, "calcNormalParticles");
    }
    void calcWallNormalParticles(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_wall_normal_particles(P->ctx), P->ctx, "calcWallNormalParticles");
    }
};

// include/MpsVectorMatrix.h
class MpsVectorMatrix {
public:
    void correctionMatrix(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_correction_matrix(P->ctx), P->ctx, "correctionMatrix");
    }
};

// include/MpsParticleVelPos.h
class MpsParticleVelPos {
public:
    void predictionVelocityPosition(MpsParticleSystem*, MpsParticle* P) {
        polymps_gpu::check(mps_prediction(P->ctx), P->ctx, "predictionVelocityPosition");
    }
    // correctionVelocityPosition also refreshes velMax / CFL on the host copy
    void correctionVelocityPosition(MpsParticleSystem* S, MpsParticle* P) {
        polymps_gpu::check(mps_correction(P->ctx), P->ctx, "correctionVelocityPosition");
        (void)S;
    }
    void updateVelocityParticlesWallDummy(MpsParticleSystem*, MpsParticle* P, MpsBucket*) {
        polymps_gpu::check(mps_wall_dummy_velocity(P->ctx), P->ctx, "updateVelocityParticlesWallDummy");
    }
};

// include/PolygonMesh.h.  One rigid mesh (me == 0); the vertex/face counts
// answer the NV.rows() / NF.rows() queries of the reference's main().
class PolygonMesh {
public:
    struct Rows {
        int n = 0;
        const std::vector<double>* v = nullptr;
        int rows() const { return n; }
        double operator()(int r, int c) const { return v ? (*v)[(size_t)r * 3 + c] : 0.0; }
    };
    Rows NV, NF;

    void initPolygonMesh(const int nP) { (void)nP; }
    // which of the context's meshes this object is (meshType: FIXED,
    // DEFORMABLE, FORCED); main.cpp indexes its mesh array by that type
    void setMeshType(int t) { type_ = t; }
    int meshTypeOf() const { return type_; }
    // PolygonMesh::readPolygonMeshFile (src/PolygonMesh.cpp:42-108)
    void readPolygonMeshFile(const std::string& path) {
        int nv = 0, nf = 0;
        polymps_gpu::check(mps_mesh_read_stl(path.c_str(), &nv, nullptr, &nf, nullptr, nullptr), nullptr,
                           ("readPolygonMeshFile " + path).c_str());
        V_.resize(3 * (size_t)nv);
        F_.resize(3 * (size_t)nf);
        N_.resize(3 * (size_t)nf);
        polymps_gpu::check(mps_mesh_read_stl(path.c_str(), &nv, V_.data(), &nf, F_.data(), N_.data()), nullptr,
                           ("readPolygonMeshFile " + path).c_str());
        NV.n = nv;
        NV.v = &V_;
        NF.n = nf;
    }
    // PolygonMesh::initWijnNeigh (src/PolygonMesh.cpp:377-527): the tables are
    // built from the context's parameters when the mesh is installed.
    void initWijnNeigh(int dim, int wijType, double lo, double reL, double reS) {
        (void)dim; (void)wijType; (void)lo; (void)reL; (void)reS;
        mps_ctx* ctx = polymps_gpu::currentContext();
        polymps_gpu::check(mps_set_mesh_type(ctx, type_, NV.n, V_.data(), NF.n, F_.data(), N_.data()), ctx,
                           "initWijnNeigh");
    }
    // PolygonMesh::closestPointPNDBoundaryAABB (src/PolygonMesh.cpp:585-715)
    void closestPointPNDBoundaryAABB(double reS2, double reL2, int nP, int wijType, int* Typ, int fld, int msh_id,
                                     int sta_id, int fem_id, int frw_id, const double* Pos, double* wallPos,
                                     double* mirrorPos, double* riw2, int* elementID, int* meshID, double* NormalWall) {
        (void)reS2; (void)reL2; (void)nP; (void)wijType; (void)Typ; (void)fld; (void)sta_id; (void)fem_id;
        (void)frw_id; (void)Pos; (void)wallPos; (void)mirrorPos; (void)riw2; (void)elementID; (void)meshID;
        (void)NormalWall; (void)msh_id;
        mps_ctx* ctx = polymps_gpu::currentContext();
        polymps_gpu::check(mps_mesh_closest_point_type(ctx, type_), ctx, "closestPointPNDBoundaryAABB");
    }
    // PolygonMesh::updateParticlesNearPolygonMesh (src/PolygonMesh.cpp:718-795)
    void updateParticlesNearPolygonMesh(double reS2, double reL2, int nP, int wijType, int* Typ, int fld,
                                        const double* riw2, double* niw, double* numNeighw,
                                        std::vector<int>& partNearMesh, bool* Nw) {
        (void)reS2; (void)reL2; (void)nP; (void)wijType; (void)Typ; (void)fld; (void)riw2; (void)niw;
        (void)numNeighw; (void)partNearMesh; (void)Nw;
        mps_ctx* ctx = polymps_gpu::currentContext();
        polymps_gpu::check(mps_mesh_update_near(ctx), ctx, "updateParticlesNearPolygonMesh");
    }
    // PolygonMesh::writePolygonMeshFile (src/PolygonMesh.cpp:110-139): <path>/meshNN_XXXXX.stl
    void writePolygonMeshFile(const int mesh_ID, const std::string& path, const int iF) {
        mps_ctx* ctx = polymps_gpu::currentContext();
        int nv = 0, nf = 0;
        polymps_gpu::check(mps_get_mesh(ctx, type_, &nv, nullptr, &nf, nullptr, nullptr), ctx, "writePolygonMeshFile");
        std::vector<double> V(3 * (size_t)nv);
        std::vector<int> F(3 * (size_t)nf);
        polymps_gpu::check(mps_get_mesh(ctx, type_, &nv, V.data(), &nf, F.data(), nullptr), ctx, "writePolygonMeshFile");
        char name[64];
        snprintf(name, sizeof name, "/mesh%02d_%05d.stl", mesh_ID, iF);
        polymps_gpu::check(mps_mesh_write_stl((path + name).c_str(), nv, V.data(), nf, F.data()), ctx,
                           "writePolygonMeshFile");
    }
    // PolygonMesh::updateForcedPolygonMesh (src/PolygonMesh.cpp:173-374): the
    // nodes move in the context; the caller's node arrays and wall velocity
    // receive the new values.
    void updateForcedPolygonMesh(double* nodeX, double* nodeY, double* nodeZ, double* velVWall, const double dt,
                                 const double time) {
        mps_ctx* ctx = polymps_gpu::currentContext();
        polymps_gpu::check(mps_forced_mesh_update(ctx, dt, time), ctx, "updateForcedPolygonMesh");
        std::vector<double> V(3 * (size_t)NV.n);
        polymps_gpu::check(mps_get_mesh(ctx, type_, nullptr, V.data(), nullptr, nullptr, nullptr), ctx,
                           "updateForcedPolygonMesh");
        for (int q = 0; q < NV.n; q++) {
            if (nodeX) nodeX[q] = V[3 * (size_t)q];
            if (nodeY) nodeY[q] = V[3 * (size_t)q + 1];
            if (nodeZ) nodeZ[q] = V[3 * (size_t)q + 2];
        }
        V_ = V;
        mps_params prm;
        polymps_gpu::check(mps_get_params(ctx, &prm), ctx, "updateForcedPolygonMesh");
        if (velVWall)
            for (int k = 0; k < 3; k++) velVWall[k] = prm.uniform_vel_wall[k];
    }

private:
    std::vector<double> V_, N_;
    std::vector<int> F_;
    int type_ = FIXED;
};

// include/MpsInputOutput.h: input (JSON, .grid) and output (VTU, PVD).
class MpsInputOutput {
public:
    std::string gridFilename = "oops", meshRigidFilename = "oops", meshDeformableFilename = "oops",
                meshForcedFilename = "oops", vtuOutputFoldername = "oops";
    int vtuFlags = 0;
    // When set, readInputFile reads this JSON file instead of prompting.
    std::string jsonPath;

    // MpsInputOutput::readInputFile (src/MpsInputOutput.cpp:58-641): asks for
    // the input name on stdin (input/<name>.json) unless jsonPath is set,
    // then reads the .grid file.
    void readInputFile(MpsParticleSystem* PSystem, MpsParticle* Particles) {
        PSystem->timerStart = polymps_gpu::wallClock();
        std::string path = jsonPath;
        while (path.empty()) {
            char name[1000];
            printf("Enter the name of the MPS input file:\n");
            if (scanf("%999s", name) != 1) throw MpsGpuError("no input file name");
            printf("\n");
            std::string p = std::string("input/") + name + ".json";
            FILE* f = fopen(p.c_str(), "r");
            if (!f) { printf("Error reading the input file %s. Try again.\n", p.c_str()); continue; }
            fclose(f);
            path = p;
        }
        printf("Reading JSON File... ");
        char grid[4096], mesh[4096], folder[4096];
        polymps_gpu::check(mps_params_from_json(path.c_str(), &PSystem->prm, grid, sizeof grid, mesh, sizeof mesh),
                           nullptr, "readInputFile");
        polymps_gpu::check(mps_output_from_json(path.c_str(), folder, sizeof folder, &vtuFlags), nullptr,
                           "readInputFile (output)");
        gridFilename = grid;
        meshRigidFilename = mesh;
        vtuOutputFoldername = folder;
        char deformable[4096], forced[4096];
        polymps_gpu::check(mps_mesh_paths_from_json(path.c_str(), deformable, sizeof deformable, forced, sizeof forced),
                           nullptr, "readInputFile (mesh paths)");
        meshDeformableFilename = deformable;
        meshForcedFilename = forced;
        PSystem->syncFromParams();
        printf("OK\n");
        printf("Reading GRID File... ");
        readMpsParticleFile(PSystem, Particles);
        printf("OK\n");
    }
    // MpsInputOutput::readMpsParticleFile (src/MpsInputOutput.cpp:645-857)
    void readMpsParticleFile(MpsParticleSystem* PSystem, MpsParticle* Particles) {
        (void)PSystem;
        int n = 0;
        polymps_gpu::check(mps_read_grid(gridFilename.c_str(), &n, nullptr, nullptr, nullptr, nullptr, nullptr),
                           nullptr, ("readMpsParticleFile " + gridFilename).c_str());
        Particles->numParticles = n;
        Particles->hostType.resize(n);
        Particles->hostPos.resize(3 * (size_t)n);
        Particles->hostVel.resize(3 * (size_t)n);
        Particles->hostPress.resize(n);
        Particles->hostPressAvg.resize(n);
        polymps_gpu::check(mps_read_grid(gridFilename.c_str(), &n, Particles->hostType.data(), Particles->hostPos.data(),
                                         Particles->hostVel.data(), Particles->hostPress.data(),
                                         Particles->hostPressAvg.data()),
                           nullptr, ("readMpsParticleFile " + gridFilename).c_str());
    }
    // MpsInputOutput::displayInfo (src/MpsInputOutput.cpp:37-56)
    void displayInfo(MpsParticleSystem* PSystem, MpsParticle* Particles, MpsPressure* PPress, const int intervalIter) {
        if (PSystem->numOfIterations % intervalIter != 0) return;
        refresh(PSystem, Particles);
        PSystem->timerEnd = polymps_gpu::wallClock();
        const int seconds = (int)(PSystem->timerEnd - PSystem->timerStart);
        const int minutes = seconds / 60, hours = minutes / 60;
        printf("Iteration: %5dth Time: %lfsec Num. Particles: %d Max Velocity: %lfm/s Courant: %lf",
               PSystem->numOfIterations, PSystem->timeCurrent, Particles->numParticles, PSystem->velMax,
               PSystem->CFLcurrent);
        if (PSystem->fluidType == NON_NEWTONIAN) {
            mps_get_scalar(Particles->ctx, "cfl_visc", &PSystem->CFLvisc);
            printf(" CFLvisc: %lf", PSystem->CFLvisc);
        }
        if (PSystem->mpsType == IMPLICIT_PND || PSystem->mpsType == IMPLICIT_PND_DIVU)
            printf(" Solver iterations: %3d Estimated error: %.2e", PPress->getSolverIter(), PPress->getSolverError());
        printf(" RunTime: %02dh%02dm%02dsec\n", hours, minutes % 60, seconds % 60);
    }
    // MpsInputOutput::writeOutputFiles (src/MpsInputOutput.cpp:861-886):
    // VTU, ASCII or (output_VTU.type = 1) binary, of every particle or
    // (output_VTU.only_freeSurface) of the free-surface and wall particles
    // (writeVtuAsciiFreeSurface, src/MpsInputOutput.cpp:1582-1875).  The file is
    // written in the background from a snapshot of the current state, so the
    // next steps run while it is formatted; finishOutput waits for it.
    void writeOutputFiles(MpsParticleSystem* PSystem, MpsParticle* Particles) {
        char name[4352];
        snprintf(name, sizeof name, "%s/output%05d.vtu", vtuOutputFoldername.c_str(), PSystem->fileNumber);
        polymps_gpu::check(mps_write_vtu_async(Particles->ctx, name, vtuFlags),
                           Particles->ctx, "writeOutputFiles");
    }
    // Waits for every VTU file still being written (reports write errors).
    void finishOutput(MpsParticle* Particles) {
        if (Particles->ctx) polymps_gpu::check(mps_output_wait(Particles->ctx), Particles->ctx, "writeOutputFiles");
    }
    // MpsInputOutput::writePvd (src/MpsInputOutput.cpp:1878-1935)
    void writePvd(MpsParticleSystem* PSystem, MpsParticle* Particles) {
        (void)Particles;
        if (mkdir(vtuOutputFoldername.c_str(), 0777) == -1)
            printf("Unable to create OUTPUT directory or it has already been created!\n");
        else
            printf("Directory created.\n");
        std::string base = vtuOutputFoldername;
        const size_t slash = base.find('/');
        if (slash != std::string::npos) base.erase(0, slash + 1);
        FILE* fp = fopen((vtuOutputFoldername + ".pvd").c_str(), "w");
        if (!fp) throw MpsGpuError("cannot write " + vtuOutputFoldername + ".pvd");
        const int nIter = (int)std::ceil(PSystem->timeSimulation / PSystem->timeStep);
        fprintf(fp, "<VTKFile type='Collection' version='0.1' byte_order='LittleEndian'>\n");
        fprintf(fp, "  <Collection>\n");
        int j = 0;
        for (int i = 0; i < nIter; i++) {
            if (i % PSystem->iterOutput != 0) continue;
            fprintf(fp, "    <DataSet timestep='%.6f' group='A' part='0' file='%s/output%05d.vtu'/>\n",
                    PSystem->timeStep * i, base.c_str(), j);
            j++;
        }
        fprintf(fp, "  </Collection>\n");
        fprintf(fp, "</VTKFile>\n");
        fclose(fp);
    }
    void deleteDirectoryFiles() {}
    void writeBuckets(MpsParticleSystem*, MpsParticle*) {}
    void stringToChar(char** out_char) {
        *out_char = new char[vtuOutputFoldername.size() + 1];
        strcpy(*out_char, vtuOutputFoldername.c_str());
    }

    // velMax, CFL, particle count and time from the device
    static void refresh(MpsParticleSystem* S, MpsParticle* P) {
        P->refreshCount();
        mps_get_scalar(P->ctx, "vel_max", &S->velMax);
        mps_get_scalar(P->ctx, "cfl", &S->CFLcurrent);
    }
};

#endif  // POLYMPS_GPU_HPP_
