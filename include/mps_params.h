/*
 * mps_params.h -- plain-C parameter block of one MPS particle system.
 *
 * One struct carries every scalar the hot path reads: the JSON inputs of the
 * reference's MpsParticleSystem (include/MpsParticleSystem.h:133-237), the
 * constants MpsParticle::setParameters derives from them
 * (src/MpsParticle.cpp:136-227) and the bucket grid of
 * MpsBucket::allocateMemory / calcDomainLimits (src/MpsBucket.cpp:20-295).
 *
 * The layout is part of the C ABI (ints and doubles only, natural alignment),
 * so C, C++ and ctypes callers share it.  mps_params_size() in mps_gpu.h
 * returns sizeof(mps_params) so a binding can check its mirror.
 */
#ifndef MPS_PARAMS_H_
#define MPS_PARAMS_H_

#ifdef __cplusplus
extern "C" {
#endif

/* enum values follow include/MpsParticleSystem.h:20-119 */
enum { MPS_WALL_PARTICLE = 0, MPS_WALL_POLYGON = 1 };
enum { MPS_EXPLICIT = 0, MPS_WEAKLY = 1, MPS_IMPLICIT_PND = 2, MPS_IMPLICIT_PND_DIVU = 3 };
enum { MPS_SOLVER_CG = 0, MPS_SOLVER_BICGSTAB = 1 };
enum { MPS_PND_SUM_WIJ = 0, MPS_PND_MEAN_SUM_WIJ = 1, MPS_PND_DIFFUSIVE = 2 };
enum { MPS_FS_PND_NEIGH = 0, MPS_FS_PND_NPCD = 1, MPS_FS_PND_ARC = 2 };
enum { MPS_NO_SLIP = 0, MPS_FREE_SLIP = 1 };
enum { MPS_REP_HARADA = 0, MPS_REP_MITSUME = 1, MPS_REP_LENNARD_JONES = 2, MPS_REP_MONAGHAN_KAJTAR = 3 };
enum { MPS_COLLISION_PC = 0, MPS_COLLISION_DPC = 1 };
/* precision of the pressure Poisson solve (no reference counterpart) */
enum { MPS_PREC_FP64 = 0, MPS_PREC_MIXED = 1 };
/* shifting update semantics (the reference's in-place OpenMP loops race,
 * src/MpsShifting.cpp:77-79,533-535,681-683): 0 = read old / write new */
enum { MPS_SHIFT_JACOBI = 0, MPS_SHIFT_SEQUENTIAL = 1 };

typedef struct mps_params {
    /* ---- JSON "flags" ---- */
    int wall_type;              /* wall_type                          */
    int fem_on;                 /* fem_MESH: a second, static mesh (meshType 1) */
    int forced_on;              /* forced_MESH: a mesh moved by the forced motion (meshType 2) */
    /* ---- JSON "domain" ---- */
    double domain_min[3];       /* min.x/y/z (extended by 3 l0 in derive) */
    double domain_max[3];
    int domain_type_bc;         /* boundary.type (0 none, 1 periodic) */
    int limit_type_bc;          /* boundary.limit                     */
    int num_bc;                 /* 1 when "boundary" present          */
    int periodic_dir[3];
    /* ---- JSON "physical" ---- */
    double density_fluid;       /* fluid_density                      */
    double density_wall;        /* wall_density                       */
    double knm_vs1;             /* kinematic_visc                     */
    int fluid_type;             /* fluid_type (0 = Newtonian only)    */
    int pad0;
    double gravity[3];
    double knm_vs2;             /* rheological.kinematic_visc_phase_2 */
    double dns_fl1;             /* rheological.fluid_density_phase_1  */
    double dns_fl2;             /* rheological.fluid_density_phase_2  */
    /* rheological parameters of the non-Newtonian (granular) phase,
       fluid_type 1 (src/MpsInputOutput.cpp:266-314) */
    double dns_sdt;             /* rheological.sediment_density (DNS_SDT) */
    double power_n;             /* rheological.power_law_index (N)        */
    double meu0;                /* rheological.consistency_index (MEU0)   */
    double phi_1, phi_wal, phi_bed, phi_2;   /* rheological.phi.lower / wall / bed / second */
    double cohes;               /* rheological.cohes_coeff                */
    double dg;                  /* rheological.grain_size                 */
    double i0;                  /* rheological.I0                         */
    double mm;                  /* rheological.mm                         */
    double cd;                  /* rheological.drag_coeff                 */
    double vf_min, vf_max;      /* rheological.volume_fraction.min / max  */
    int fraction_method;        /* rheological.fraction_method: 1 neighbour count, 2 kernel weighted */
    int stress_calc_method;     /* rheological.stress_calc_method (read, not used by the reference) */
    /* ---- JSON "numerical" ---- */
    double dim;                 /* dimension                          */
    double part_dist;           /* particle_dist  (l0)                */
    double time_step;           /* time_step                          */
    double time_simulation;     /* final_time                         */
    int iter_output;            /* iter_output                        */
    int weight_type;            /* weight_type                        */
    double cfl_number;          /* CFL_number                         */
    int slip_condition;         /* slip_condition                     */
    int gradient_type;          /* gradient.type                      */
    double re_s;                /* effective_radius.small (x l0 after derive) */
    double re_l;                /* effective_radius.large (x l0 after derive) */
    int gradient_correction;    /* gradient.correction                */
    int mps_type;               /* mps_type                           */
    double relax_press;         /* gradient.relax_fact                */
    double sound_speed;         /* explicit_mps.equation_state.speed_sound */
    double gamma;               /* explicit_mps.equation_state.gamma  */
    int solver_type;            /* semi_implicit_mps.solver_type (0 CG, 1 BiCGSTAB) */
    int shifting_type;          /* particle_shifting.type             */
    double alpha_compressibility; /* semi_implicit_mps.weak_compressibility.alpha */
    double relax_pnd;           /* semi_implicit_mps.source_term.relax_pnd */
    double dri;                 /* particle_shifting.DRI              */
    double coef_a;              /* particle_shifting.coef_A           */
    double mach_number;         /* particle_shifting.mach_number      */
    int pnd_type;               /* pnd.type                           */
    int repulsive_force_type;   /* wall_repulsive_force.type          */
    double diffusive_coef;      /* pnd.diffusive_coeff                */
    double re_repulsive_force;  /* wall_repulsive_force.re (x l0 after derive) */
    double expect_max_velocity; /* wall_repulsive_force.maxVel        */
    double rep_coef_mitsume;
    double rep_coef_lennard_jones;
    double rep_coef_monaghan_kajtar;
    double eps_re;              /* wall_repulsive_force.eps_re        */
    int free_surf_type;         /* free_surface_threshold.type        */
    int collision_type;         /* particle_collision.type            */
    double pnd_threshold;
    double neigh_threshold;
    double npcd_threshold;
    double theta_threshold;
    double norm_threshold;
    double collision_ratio;
    double dist_limit_ratio;
    double lambda_collision;
    int ghost, fluid, wall, dummy_wall;   /* particle_type ids     */
    int surface, inner, other;            /* boundary_type ids     */
    int pad1;

    /* ---- solver / execution extensions (no reference key) ---- */
    double solver_tol;          /* relative residual target; <=0 -> DBL_EPSILON (Eigen default) */
    int solver_max_iter;        /* <=0 -> 2 N (Eigen default)         */
    int precision_mode;         /* MPS_PREC_*                         */
    int shift_semantics;        /* MPS_SHIFT_*                        */
    int ppe_transpose;          /* 1: iterate with A^T as Eigen's Lower|Upper CG does on a
                                   column-major matrix; 0: iterate with A          */
    double mixed_inner_tol;     /* inner FP32 PCG tolerance of the refinement loop */
    int learned_mode;           /* 0 off, 1 warm start, 2 warm start + preconditioner */
    int ppe_warm_start;         /* initial guess of each pressure solve: 0 the previous solution
                                   (the reference), 1 / 2 / 3 linear / quadratic / cubic
                                   extrapolation of the last 2 / 3 / 4 solutions */
    /* forced mesh motion (the reference hard-codes it in updateForcedPolygonMesh,
       src/PolygonMesh.cpp:173-374; the defaults are its active law): 1 rotation
       about the y axis through (center_x, center_z), angle theta_max sin(2 pi t / period)
       degrees; 0 translation along z at forced_velocity */
    int forced_motion;
    int pad2;
    double forced_center[2];    /* Xcg, Zcg                           */
    double forced_theta_max;    /* degrees                            */
    double forced_period;       /* s                                  */
    double forced_velocity;     /* m/s (translation)                  */
    double uniform_vel_wall[3]; /* uniformVelWall: velocity of the forced mesh's wall points */

    /* ---- derived by mps_params_derive (setParameters + bucket grid) ---- */
    double eps_zero;            /* 10 * DBL_EPSILON                   */
    double re_s2, re_l2;
    double pnd_small_zero, pnd_large_zero, pnd_gradient_zero;
    double lambda_zero, num_neigh_zero;
    double coeff_viscosity, coeff_visc_multiphase;
    double coeff_press_emps, coeff_press_grad, coeff_press_wcmps;
    double coeff_shifting1, coeff_shifting2;
    double coeff_ppe, coeff_ppe_source;
    double dns_fluid, dns_wall, inv_dns_fluid, inv_dns_wall;
    double inv_part_dist;
    double dist_collision_limit, dist_collision_limit2;
    double restitution_collision;
    double beta_pnd, beta_neigh, delta2;
    double theta_arc, h_threshold2, dst_threshold2, norm_threshold2;
    double bucket_side, inv_bucket_side;
    double periodic_length[3];  /* periodicLength (0 where not periodic) */
    double phys_min[3], phys_max[3];   /* physDomMin/Max: the periodic wrap limits */
    int num_buckets_x, num_buckets_y, num_buckets_z;
    int num_buckets_xy, num_buckets_xyz;
    int num_of_meshs;
    int derived;                /* set to 1 by mps_params_derive      */
    int pad3;
} mps_params;

#ifdef __cplusplus
}
#endif
#endif /* MPS_PARAMS_H_ */
