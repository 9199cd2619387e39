/*
 * mps_gpu.h -- C ABI of the MI355X (gfx950) MPS time-step library.
 *
 * Plain pointers and sizes only.  Every call returns 0 on success or a
 * negative MPS_E* code; mps_last_error() gives the message.  A context owns
 * all device buffers and one HIP stream; it is single-thread affine.
 *
 * Particle data crossing the boundary is in the reference's order and
 * layout: AoS double[3N] for vectors, double[N] / int[N] for scalars, ids as
 * in the JSON file (fluid 0, wall 2, dummy 3, ghost -1; BC inner 0,
 * surface 1, other -1).  Internally particles live in bucket-sorted SoA
 * order; that order is never visible to the caller.
 *
 * Each operator entry point replaces one method of the reference; the
 * citation names the method it stands for.  Operators run in stream order
 * and leave their results on the device.
 */
#ifndef MPS_GPU_H_
#define MPS_GPU_H_

#include <stddef.h>
#include "mps_params.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MPS_OK 0
#define MPS_E_ARG (-1)      /* invalid argument                 */
#define MPS_E_HIP (-2)      /* HIP runtime error                */
#define MPS_E_UNSUPPORTED (-3) /* configuration not supported   */
#define MPS_E_IO (-4)       /* file could not be read / written */
#define MPS_E_CAPACITY (-5) /* a fixed-size device table overflowed */
#define MPS_E_SOLVER (-6)   /* the PPE solve did not converge: it stopped above its
                             * tolerance (FP64 CG / BiCGSTAB: Eigen's rule, the
                             * recursive relative residual; mixed: the FP64 true
                             * residual) or met a non-finite residual.  The
                             * pressure is written either way, as the reference
                             * does after printing "Failed solving"
                             * (src/MpsPressure.cpp:440-442); scalar
                             * "solver_failed" tells the two cases apart. */

typedef struct mps_ctx mps_ctx;

/* ------------------------------------------------------------ parameters */
/* sizeof(mps_params), for bindings that mirror the struct */
size_t mps_params_size(void);
/* JSON defaults of MpsInputOutput::readInputFile (src/MpsInputOutput.cpp:136-500) */
int mps_params_defaults(mps_params* p);
/* MpsInputOutput::readInputFile (src/MpsInputOutput.cpp:58-641): parse a
 * JSON file (comments allowed, missing keys take the defaults).  grid_path
 * and mesh_path receive pathNames.particle_grid_file / mesh_rigid_file. */
int mps_params_from_json(const char* json_path, mps_params* p, char* grid_path, size_t grid_cap,
                         char* mesh_path, size_t mesh_cap);
/* MpsParticle::setParameters (src/MpsParticle.cpp:136-227) plus
 * MpsBucket::allocateMemory / calcDomainLimits (src/MpsBucket.cpp:20-295).
 * pos/type are needed only for particle walls (domain from particles). */
int mps_params_derive(mps_params* p, int n, const double* pos, const int* type);

/* --------------------------------------------------------------- context */
/* Create a context for n particles on HIP device `device`.  p must be
 * derived.  Replaces the allocation in MpsInputOutput::readMpsParticleFile
 * (src/MpsInputOutput.cpp:667-750) and MpsBucket::allocateMemory. */
int mps_create(mps_ctx** out, const mps_params* p, int n, int device);
int mps_destroy(mps_ctx* ctx);
const char* mps_last_error(const mps_ctx* ctx);
/* Copy of the context's parameter block */
int mps_get_params(const mps_ctx* ctx, mps_params* out);

/* Upload the .grid state (src/MpsInputOutput.cpp:757-776): type[N],
 * pos[3N], vel[3N], press[N], press_avg[N]; the remaining fields take the
 * initial values of readMpsParticleFile (:812-838). */
int mps_upload_state(mps_ctx* ctx, const int* type, const double* pos, const double* vel,
                     const double* press, const double* press_avg);
/* Download one field in reference order.  Double fields: "pos" "vel" "acc"
 * "acc_star" "npcd" "mirror" "wall_pos" "polygon_normal" "force_wall"
 * "cm1" "cm2" "cm3" "grad_conc" (3 per particle); "press" "press_avg" "pndi"
 * "pndki" "pndski" "pnd_small" "npcd2" "conc" "div_u" "diffusive" "pnd_wall"
 * "riw2" "ppe_x" "ppe_b" "dev_dot_n" (1 per particle).  Int fields: "type"
 * "bc" "num_neigh" "num_neigh_wall" "near_wall" "near_mesh_type".  `count`
 * is the element capacity of out. */
int mps_download(mps_ctx* ctx, const char* field, void* out, size_t count);
/* mps_download without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory (tight serving
 * loops, code that forbids hipMalloc).  The field is staged in a 64 MiB area
 * the library holds per GPU until the process ends; a field larger than that
 * (a vector field beyond 2,796,202 particles, a scalar beyond 8,388,608)
 * returns MPS_E_CAPACITY.  All contexts on one GPU share the area: calls that
 * overlap on one GPU (other threads, other contexts) race. */
int mps_download_noalloc(mps_ctx* ctx, const char* field, void* out, size_t count);
/* Overwrite one field (reference order), same names as mps_download. */
int mps_upload_field(mps_ctx* ctx, const char* field, const void* in, size_t count);
/* Scalars: "vel_max" "cfl" "cfl_visc" (granular phase, src/MpsViscosity.cpp:691) "time" "solver_iter" "solver_err" "num_particles"
 * "iterations" "ppe_ms" "true_residual" "fallback" "outer_iterations"
 * "solver_failed" (last pressure solve: 0 converged, 1 stopped above the
 * tolerance, 2 non-finite residual) "solver_failures" (failed solves since
 * the count was last set; settable)
 * "list_max" "list_capacity" "capacity" "nnz" "list_start" (1 when the last
 * step built a list at its start positions) "spmv_ms_total"
 * "spmv_launches" "spmv_bytes_total" (with "profile" on: the time, count and
 * algorithmic bytes of the timed SpMV launches, every eighth launch of a
 * batch) "list_skin" (skin of the last step's
 * kept neighbour list, in l0; 0: built after the prediction) "list_rebuilt"
 * (1 when the last step rebuilt its list after the prediction)
 * "list_entries" "list_entries_small" (entries of the current neighbour
 * lists at the large and the small radius, summed over the particles)
 * "list_patched" (particles whose list the last step rebuilt alone).
 * Settable: "vel_max" "time" "list_skin" (< 0 automatic, 0 off, > 0 a fixed
 * skin in l0; every choice gives the same results) "list_skin_cap" (largest
 * automatic skin in l0, default 0.08; a step that needs more builds its list
 * after the prediction) "list_skin_factor" (automatic skin as a multiple of
 * the largest one-step displacement, in [2, 10], default 3) "mask_words" (candidate
 * window per stencil row in 32-slot words, at most the allocated width;
 * particles with a longer row are listed by a direct walk, same results)
 * "profile" (nonzero: time every eighth SpMV launch with HIP events, counters reset)
 * "precision_mode" (MPS_PREC_*, used from the next PPE assembly)
 * "learned_mode" (0 off, 1 learned warm start, 2 warm start + learned
 * flexible-CG preconditioner) "learned_stall_ratio" (in (0, 1), default
 * 0.5: the learned preconditioner is dropped for Jacobi, "fallback" 1, when a
 * refinement pass leaves more than this fraction of the true residual)
 * "learned_tol_scale" (in (0, 1], default 0.1: with the learned
 * preconditioner the true residual is driven to this multiple of the
 * tolerance, since the flexible CG's error at a given residual is larger)
 * "ppe_window" (1 default: the CG solves' SpMVs read each 256-row block's
 * vector window from LDS; 0: from global memory; same results,
 * used from the next assembly; read-only "ppe_window_blocks": blocks of the
 * last assembly served from a window, "window_escapes": window index
 * escapes of the last assembly, the rows' absolute indices next to their
 * one-byte steps) "ppe_warm_start" (initial guess of each
 * pressure solve: 0 previous solution, 1 / 2 / 3 linear / quadratic / cubic
 * extrapolation of the last solutions) "mesh_bvh" (closest-point search: -1 BVH for
 * meshes above 64 faces, 0 brute force, 1 BVH; both give the same result). */
int mps_get_scalar(mps_ctx* ctx, const char* name, double* out);
int mps_set_scalar(mps_ctx* ctx, const char* name, double value);
int mps_sync(mps_ctx* ctx);

/* ------------------------------------------------------------- mesh wall */
/* PolygonMesh::readPolygonMeshFile (src/PolygonMesh.cpp:42-108): binary or
 * ASCII STL -> duplicate vertices merged (tolerance 2e-15) -> unit face
 * normals.  Writes counts; arrays may be NULL to query sizes. */
int mps_mesh_read_stl(const char* path, int* nv, double* V, int* nf, int* F, double* N);
/* Install the rigid wall mesh (V[3nv], F[3nf], unit normals N[3nf]) and
 * build the wall tables of PolygonMesh::initWijnNeigh (src/PolygonMesh.cpp:377-527). */
int mps_set_mesh(mps_ctx* ctx, int nv, const double* V, int nf, const int* F, const double* N);
/* Mesh by type, as main.cpp's mesh[meshType] (src/main.cpp:471-515):
 * MPS_MESH_FIXED (= mps_set_mesh), MPS_MESH_DEFORMABLE (fem_MESH; static here,
 * the reference's FEM coupling is not part of it), MPS_MESH_FORCED (forced_MESH,
 * moved by mps_forced_mesh_update).  The closest-point search visits the
 * installed meshes in type order and keeps a strictly closer point
 * (src/PolygonMesh.cpp:615-650), recording the type in nearMeshType. */
#define MPS_MESH_FIXED 0
#define MPS_MESH_DEFORMABLE 1
#define MPS_MESH_FORCED 2
int mps_set_mesh_type(mps_ctx* ctx, int type, int nv, const double* V, int nf, const int* F, const double* N);
/* Current geometry of an installed mesh (the forced one moves); arrays may be
 * NULL to query the counts. */
int mps_get_mesh(mps_ctx* ctx, int type, int* nv, double* V, int* nf, int* F, double* N);
/* PolygonMesh::updateForcedPolygonMesh (src/PolygonMesh.cpp:173-374): move the
 * forced mesh's nodes by one step of the forced motion at time `time` (the
 * parameters forced_*), recompute its face normals, and set uniform_vel_wall. */
int mps_forced_mesh_update(mps_ctx* ctx, double dt, double time);
/* PolygonMesh::writePolygonMeshFile (src/PolygonMesh.cpp:110-139): ASCII STL
 * of a mesh (facet normals from the faces). */
int mps_mesh_write_stl(const char* path, int nv, const double* V, int nf, const int* F);
/* Wall tables xPND, xNeigh, Zij, nNeighWall (50 entries each). */
int mps_get_wall_tables(mps_ctx* ctx, double* xpnd, double* xneigh, double* zij, int* nneigh);

/* ------------------------------------------------------------- operators */
/* MpsBucket::updateParticlesID (src/MpsBucket.cpp:299-353): counting sort of
 * particles into buckets, particles reordered on the device. */
int mps_bucket_update(mps_ctx* ctx);
/* The bucket lists of the last mps_bucket_update as the reference keeps them
 * (firstParticleInBucket, lastParticleInBucket, nextParticleInSameBucket,
 * include/MpsParticle.h:139-141, filled by src/MpsBucket.cpp:299-353 and, for
 * a periodic domain, copyDataBetweenBuckets :356-417): reference particle
 * indices, -1 for none; first / last have num_buckets_xyz entries, next the
 * context's capacity.  Arrays may be NULL.  Returns num_buckets_xyz (>= 0) or
 * an error code. */
int mps_get_buckets(mps_ctx* ctx, int* first, int* last, int* next);
/* MpsViscosity::calcViscosity (src/MpsViscosity.cpp:20-102) */
int mps_viscosity(mps_ctx* ctx);
/* MpsParticle::addGravity (src/MpsParticle.cpp:111-130) */
int mps_add_gravity(mps_ctx* ctx);
/* MpsParticleVelPos::predictionVelocityPosition (src/MpsParticleVelPos.cpp:20-50) */
int mps_prediction(mps_ctx* ctx);
/* MpsParticle::checkParticleOutDomain (src/MpsParticle.cpp:234-402) */
int mps_check_out_domain(mps_ctx* ctx);
/* Build the per-step neighbour list at the current positions (no reference
 * counterpart: the reference re-walks the buckets in every operator). */
int mps_build_neighbor_list(mps_ctx* ctx);
/* MpsParticleCollision::checkParticleCollisions / checkDynamicParticleCollisions
 * (src/MpsParticleCollision.cpp:22-119, 124-292) */
int mps_collision(mps_ctx* ctx);
/* PolygonMesh::closestPointPNDBoundaryAABB (src/PolygonMesh.cpp:585-715), for
 * every installed mesh in type order (main.cpp's loop, src/main.cpp:296-300) */
int mps_mesh_closest_point(mps_ctx* ctx);
/* The same for one mesh type (one call of that loop). */
int mps_mesh_closest_point_type(mps_ctx* ctx, int type);
/* PolygonMesh::updateParticlesNearPolygonMesh (src/PolygonMesh.cpp:718-795) */
int mps_mesh_update_near(mps_ctx* ctx);
/* MpsPndNeigh::calcWallNPCD (src/MpsPndNeigh.cpp:105-195) */
int mps_wall_npcd(mps_ctx* ctx);
/* MpsVectorMatrix::correctionMatrix (src/MpsVectorMatrix.cpp:20-93) */
int mps_correction_matrix(mps_ctx* ctx);
/* MpsPndNeigh::setInitialPndNumberOfNeighNPCD (src/MpsPndNeigh.cpp:21-101) */
int mps_initial_pnd(mps_ctx* ctx);
/* MpsPndNeigh::calcPndnNeighNPCD (src/MpsPndNeigh.cpp:199-331) */
int mps_pnd_neigh_npcd(mps_ctx* ctx);
/* MpsPndNeigh::calcPndDiffusiveTerm + calcWall{Slip,NoSlip}PndDiffusiveTerm
 * + updatePnd + meanPndParticlesWallDummySurface (src/MpsPndNeigh.cpp:334-852) */
int mps_pnd_diffusive(mps_ctx* ctx);
/* MpsPndNeigh::meanWallPnd (:923-994) then meanPnd (:855-920) */
int mps_mean_pnd(mps_ctx* ctx);
/* MpsPndNeigh::meanNeighFluidPnd (:997-1050) */
int mps_mean_neigh_fluid_pnd(mps_ctx* ctx);
/* MpsBoundaryCondition::updateParticleBC (src/MpsBoundaryCondition.cpp:66-400) */
int mps_update_bc(mps_ctx* ctx);
/* MpsPressure::calcPress (src/MpsPressure.cpp:32-58): EOS pressure, or PPE
 * assembly (solvePressurePoissonPnd[DivU] :169-421) + Jacobi-PCG
 * (replacing solveConjugateGradient :424-445) + setZeroOnNegativePressure. */
int mps_calc_press(mps_ctx* ctx);
/* MpsPressure::calcVelDivergence (src/MpsPressure.cpp:479-551) and, with
 * polygon walls, calcWall{Slip,NoSlip}VelDivergence (:554-788): the velocity
 * divergence of the PND + DivU source term (part of mps_calc_press when
 * mps_type is 3) */
int mps_vel_divergence(mps_ctx* ctx);
/* PPE assembly only (the first half of mps_calc_press) */
int mps_ppe_assemble(mps_ctx* ctx);
/* PPE solve only, from the current ppe_x guess (second half of mps_calc_press).
 * MPS_E_ARG when no assembly matches the current slot order (a bucket update,
 * compaction or state upload came after it) or the assembly was made for
 * another solver_type / precision_mode; MPS_E_SOLVER as above. */
int mps_ppe_solve(mps_ctx* ctx);
/* The initial guess mps_calc_press gives the solve (no reference
 * counterpart): ppe_x becomes the extrapolation of order ppe_warm_start from
 * ppe_x and the two solutions before it, which are shifted back by one;
 * nothing happens for ppe_warm_start 0. */
int mps_ppe_warm_start(mps_ctx* ctx);
/* Assembled iteration matrix in reference order: row_ptr[N+1], cols/vals[nnz]
 * (column = reference particle index).  Arrays may be NULL to query nnz. */
int mps_ppe_matrix(mps_ctx* ctx, int* row_ptr, int* cols, double* vals, int* nnz);
/* MpsPressure::extrapolatePressParticlesWallDummy (src/MpsPressure.cpp:792-846) */
int mps_extrapolate_wall_dummy(mps_ctx* ctx);
/* MpsShifting::calcNormalParticles (src/MpsShifting.cpp:106-175) and
 * calcWallNormalParticles (:319-430): the free-surface normals of the ARC
 * rule (free_surface_threshold.type 2) */
int mps_normal_particles(mps_ctx* ctx);
int mps_wall_normal_particles(mps_ctx* ctx);
/* MpsPressure::predictionPressGradient (src/MpsPressure.cpp:932-1028) and
 * predictionWallPressGradient (:1031-1214): acc += (1 - relax_fact) grad p / rho
 * before the predictor when relax_fact < 1 */
int mps_prediction_press_gradient(mps_ctx* ctx);
int mps_prediction_wall_press_gradient(mps_ctx* ctx);
/* MpsPressure::calcPressGradient (src/MpsPressure.cpp:1217-1338) */
int mps_press_gradient(mps_ctx* ctx);
/* MpsPressure::calcWallPressGradient (src/MpsPressure.cpp:1341-1550) */
int mps_wall_press_gradient(mps_ctx* ctx);
/* Granular (non-Newtonian) phase, fluid_type 1; the particles of type 1 in
 * the state uploaded by mps_upload_state are its particles (fluid of phase
 * 2, src/MpsInputOutput.cpp:839-855).
 * MpsViscosity::calcVolumeFraction (src/MpsViscosity.cpp:104-191) */
int mps_volume_fraction(mps_ctx* ctx);
/* MpsViscosity::calcViscosityInteractionVal (src/MpsViscosity.cpp:194-700):
 * bed levels, strain rate, mu(I) rheology (MEU, RHO, ...).  The reference
 * fills its bed-level grid from an unsynchronised parallel loop; the result
 * here is that of the loop in index order. */
int mps_viscosity_interaction(mps_ctx* ctx);
/* MpsViscosity::calcWall{Slip,NoSlip}ViscosityInteractionVal
 * (src/MpsViscosity.cpp:703-1526), polygon walls in 3-D */
int mps_wall_viscosity_interaction(mps_ctx* ctx);
/* MpsViscosity::calcWall{Slip,NoSlip}Viscosity (src/MpsViscosity.cpp:1527-1910) */
int mps_wall_viscosity(mps_ctx* ctx);
/* MpsParticleVelPos::correctionVelocityPosition (src/MpsParticleVelPos.cpp:53-97) */
int mps_correction(mps_ctx* ctx);
/* MpsShifting::calcShifting or calcConcAndConcGradient +
 * calcWallConcAndConcGradient (src/MpsShifting.cpp:23-104, 432-692), with
 * read-old / write-new semantics */
int mps_shifting(mps_ctx* ctx);
/* MpsParticleVelPos::updateVelocityParticlesWallDummy (src/MpsParticleVelPos.cpp:100-158) */
int mps_wall_dummy_velocity(mps_ctx* ctx);

/* ---------------------------------------------------------------- driver */
/* The initial-step pipeline of main() (src/main.cpp:119-180, initMesh
 * :471-556): buckets, out-of-domain check, wall distance, initial PND, mean
 * PND, BC, first pressure. */
int mps_init(mps_ctx* ctx);
/* One time step of mainLoopOfSimulation (src/main.cpp:252-467).  A pressure
 * solve that stops above its tolerance lets the step finish, as the
 * reference's loop goes on after "Failed solving", and the step then returns
 * MPS_E_SOLVER; a non-finite residual returns at once. */
int mps_step(mps_ctx* ctx);
/* k steps without host synchronisation in between; solves that stop above
 * their tolerance do not stop the k steps, MPS_E_SOLVER is returned after
 * the last one (message: the first failing step); a non-finite residual
 * returns at once. */
int mps_steps(mps_ctx* ctx, int k);

/* ------------------------------------------------------------------ I/O */
/* .grid reader (src/MpsInputOutput.cpp:645-857); arrays may be NULL to query n */
int mps_read_grid(const char* path, int* n, int* type, double* pos, double* vel, double* press,
                  double* press_avg);
/* .grid writer in the same format (the state a .grid file can hold) */
int mps_write_grid(const char* path, int n, const int* type, const double* pos, const double* vel,
                   const double* press, const double* press_avg);
/* Checkpoint of the whole simulation state (no reference counterpart: the
 * reference restarts from a .grid file, src/MpsInputOutput.cpp:645-857,
 * which leaves out the pressure solver's warm-start vector pressurePPE,
 * include/MpsParticle.h:205, and the other fields carried between steps).
 * The file holds every per-particle field carried from one step to the
 * next, in reference order (positions, velocities, pressures, the PPE
 * solution and the extrapolation history, PND fields, wall forces, the
 * granular density / viscosity / correction matrix / phase), the live count,
 * time, step count, velocity maximum, pressure-solve count, the parameters
 * and the nodes of the deformable and forced meshes.  mps_checkpoint_read
 * restores it into a context created with the same parameters, capacity
 * and meshes (the rigid mesh, and the deformable / forced meshes with the
 * same node and face counts, installed first); the run then continues bit
 * for bit as the one that wrote it.  A learned PPE operator's weights are
 * not included (mps_gnn_save / mps_gnn_load).  Errors: MPS_E_IO for an
 * unreadable, truncated or foreign file, MPS_E_ARG for a context whose
 * capacity, bucket grid, phases or meshes differ; the context is unchanged
 * by a failed read. */
int mps_checkpoint_write(mps_ctx* ctx, const char* path);
int mps_checkpoint_read(mps_ctx* ctx, const char* path);
/* Output options (flags.output_VTU of the JSON file) */
#define MPS_VTU_PND 1
#define MPS_VTU_NEIGH 2
#define MPS_VTU_DEVIATION 4
#define MPS_VTU_CONCENTRATION 8
#define MPS_VTU_AUXILIAR 16
#define MPS_VTU_ONLY_FREE_SURFACE 32  /* free-surface and wall particles only (writeVtuAsciiFreeSurface) */
#define MPS_VTU_BINARY 64             /* binary VTU (appended raw data) instead of ASCII */
#define MPS_VTU_NON_NEWTONIAN 128     /* the granular phase's arrays (output_VTU.non_newtonian) */
/* The output part of MpsInputOutput::readInputFile (src/MpsInputOutput.cpp:140-151,
 * 173-179): pathNames.vtu_output_folder and the MPS_VTU_* flags. */
int mps_output_from_json(const char* json_path, char* vtu_folder, size_t folder_cap, int* vtu_flags);
/* pathNames.mesh_deformable_file and mesh_forced_file (src/MpsInputOutput.cpp:177-178). */
int mps_mesh_paths_from_json(const char* json_path, char* deformable_path, size_t deformable_cap, char* forced_path,
                             size_t forced_cap);
/* MpsInputOutput::writeVtuAscii (src/MpsInputOutput.cpp:1313-1579) for the
 * context's current state; flags: MPS_VTU_PND .. MPS_VTU_AUXILIAR;
 * MPS_VTU_ONLY_FREE_SURFACE for writeVtuAsciiFreeSurface (:1582-1875: the
 * particles with the free-surface BC or the wall type, no ParticleType array,
 * polygonNormal of every particle as the reference prints it); and
 * MPS_VTU_BINARY for a binary file with the same arrays (appended raw data,
 * UInt64 block headers; the reference's binary writer, :921-1108, is marked
 * as not working in include/MpsInputOutput.h:85). */
int mps_write_vtu(mps_ctx* ctx, const char* path, int flags);
/* The same file, written in the background: the state is snapshotted in
 * stream order (later operator calls may run at once), copied to pinned host
 * memory on a second stream and formatted by a writer thread.  Two snapshots
 * can be in flight; a third call waits for the oldest.  Errors of background
 * writes are returned by mps_output_wait (and by mps_write_vtu, which waits
 * for every pending write).  mps_destroy waits as well.  No reference
 * counterpart: the reference's writeOutputFiles (src/MpsInputOutput.cpp:861-886)
 * blocks the step loop while it prints. */
int mps_write_vtu_async(mps_ctx* ctx, const char* path, int flags);
int mps_output_wait(mps_ctx* ctx);

/* ------------------------------------------------- learned PPE operator */
/* No counterpart in the reference.  A message-passing network over the reS
 * neighbour graph that outputs the sparse operator
 *   (M r)_i = s_i r_i / a_ii + sum_j c_ij r_j / sqrt(a_ii a_jj),
 * used as warm start and flexible-CG preconditioner when
 * params.learned_mode is 1 or 2 (Jacobi when its output weights are 0). */
int mps_gnn_param_count(void);
/* seeded initial parameters (the Jacobi operator); resets the optimiser */
int mps_gnn_init(mps_ctx* ctx, unsigned long long seed);
int mps_gnn_set_params(mps_ctx* ctx, const double* theta, int n);
int mps_gnn_get_params(mps_ctx* ctx, double* theta, int n);
int mps_gnn_load(mps_ctx* ctx, const char* path);
int mps_gnn_save(mps_ctx* ctx, const char* path);
/* graph, features and operator for the current state and assembled matrix */
int mps_gnn_forward(mps_ctx* ctx);
/* the last forward pass in reference order: CSR indptr[nlive+1], cols, edge
 * features e[4 E], weights w[E], node features feat[4 nlive], diagonal,
 * operator cs[E] (c_ij / sqrt(a_ii a_jj)), dsi[nlive] (s_i / a_ii); arrays
 * may be NULL to query E */
int mps_gnn_graph(mps_ctx* ctx, int* indptr, int* cols, double* e, double* w, double* feat, double* diag,
                  double* cs, double* dsi, int* nedges);
/* gradient of sum_i gz_i (M r)_i with respect to the parameters */
int mps_gnn_backward(mps_ctx* ctx, const double* r, const double* gz, double* grad, int n);
/* one Adam step on the current assembled matrix (lr = 0: loss only) */
int mps_gnn_train_step(mps_ctx* ctx, double lr, unsigned long long seed, double* loss);

#ifdef __cplusplus
}
#endif
#endif /* MPS_GPU_H_ */
