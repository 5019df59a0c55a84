/*
 * wf_abi.h — C-ABI of libwfhip, the gfx950 HIP implementation of the
 * WeldFormGPU per-step SPH particle pipeline (neighbour search, pair loops,
 * stress/plasticity update, integration + velocity BCs, thermal conduction,
 * rigid triangle-mesh penalty contact, internal energy).
 *
 * The entry points replace the reference's Domain_d device call surface
 * (src/cuda/Domain_d.cuh:101-426) one for one; each declaration cites the
 * reference kernel or method it stands in for.  The ABI is plain C: opaque
 * handle, plain pointers and sizes, int status codes, no C++ types and no
 * exceptions across it.
 *
 * Conventions
 *  - All per-particle arrays crossing the ABI are in the ORIGINAL particle
 *    order (the order the caller generated them in).  The library reorders
 *    particles internally along a Morton curve of the neighbour-search cells
 *    and keeps the permutation; wf_get_field undoes it.
 *  - Vectors are interleaved xyz (double[3*n]) like the reference's double3
 *    arrays; symmetric and antisymmetric tensors are 6-component flat
 *    [xx,yy,zz,xy,yz,xz] (reference ToFlatSymPtr, src/cuda/tensor3.cu:95-100).
 *  - All calls enqueue on the handle's own HIP stream and return; wf_sync
 *    waits.  Field transfers and queries are synchronous.
 *  - A handle is not thread-safe (like the reference's Domain_d).
 *  - Return value 0 = WF_OK, negative = WF_E_*; wf_last_error() gives the
 *    message of the last failure on the calling thread.
 */
#ifndef WF_ABI_H
#define WF_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WF_ABI_VERSION 7

/* ---- status codes ------------------------------------------------------ */
enum {
  WF_OK = 0,
  WF_E_ARG = -1,         /* bad argument (null, size mismatch, unknown enum) */
  WF_E_HIP = -2,         /* HIP runtime error                                */
  WF_E_STATE = -3,       /* call not valid in the current state              */
  WF_E_NBCAP = -4,       /* a particle has more than WF_MAX_NEIGHBORS        */
  WF_E_NONFINITE = -5,   /* NaN/Inf detected in a checked field              */
  WF_E_NOMEM = -6,       /* device allocation failed                         */
  WF_E_UNSUPPORTED = -7  /* option not built / not supported                 */
};

/* cap on neighbours per particle: cuNSearchKernels.cuh:6 (CUDA_MAX_NEIGHBORS) */
#define WF_MAX_NEIGHBORS 70
/* list capacity under WF_REBUILD_PLASTIC_ZONE (the skin adds the next
 * lattice shell: 80 neighbours in the bulk of a cubic lattice) */
#define WF_MAX_NEIGHBORS_SKIN 96

/* ---- material models: src/cuda/Material.cuh:6-8 ------------------------- */
enum { WF_MAT_BILINEAR = 0, WF_MAT_HOLLOMON = 1, WF_MAT_JOHNSON_COOK = 2 };

/* ---- arithmetic of the pair loops -------------------------------------- */
enum {
  WF_FP64 = 0,  /* everything in FP64 (the reference's arithmetic)          */
  WF_MIXED = 1  /* FP32 pair loops with compensated sums and cell-relative
                   coordinates; FP64 integration, return map and reductions */
};

/* ---- plastic return map ------------------------------------------------- */
enum {
  WF_RETURN_REFERENCE = 0, /* scale back with old sigma_y, explicit re-eval,
                              dep=(s_trial-s_y)/(3G): Mechanical.cu:504-535 */
  WF_RETURN_NEWTON = 1     /* FP64 Newton radial return on
                              s_trial - 3G*dg - s_y(ep+dg) = 0 (extension)  */
};

/* ---- reference quirks reproduced by default (bit flags) ---------------- */
#define WF_Q_ROT_XZ_SIGN   0x1 /* W.xz negated: InteractionAlt.cu:362           */
#define WF_Q_TENSOR_PRODUCT 0x2 /* tensor3 operator* takes A*B in the first column
                                  and A*B^T elsewhere: tensor3.cu:241-255       */
#define WF_Q_STALE_NS_POS  0x4 /* rebuild from the last output snapshot x_h:
                                  SolverFraser.cu:171-175                       */
#define WF_Q_ALL (WF_Q_ROT_XZ_SIGN | WF_Q_TENSOR_PRODUCT | WF_Q_STALE_NS_POS)

/* ---- thermal boundary codes (per particle BC_T) ------------------------- */
enum {
  WF_BCT_NONE = 0,
  WF_BCT_CONVECTION = 1, /* CalcConvHeatKernel, Thermal.cu:36-64             */
  WF_BCT_FIXED = 2       /* Dirichlet: temperature held (extension)         */
};

/* ---- contact penalty law ------------------------------------------------ */
enum {
  WF_CONTACT_REFERENCE = 0, /* F = 0.4 (2 m_i/dt^2) (h_i - d) n, friction
                               stiffness 2 m_i/dt^2: Contact.cu:247,266  */
  WF_CONTACT_STIFFNESS = 1  /* F = kappa E dx (h_i - d) n (a stiffness
                               kappa E/dx over a particle face dx^2),
                               friction stiffness kappa E dx (extension) */
};

/* ---- material record: src/cuda/Material.cuh:24-93; the JSON "const" array
 *      order is WeldFormGPU.cu:753-792 ------------------------------------- */
typedef struct wf_material_params {
  int model;       /* WF_MAT_*                                              */
  int _pad0;
  double E, nu;    /* youngsModulus, poissonsRatio                          */
  double sy0;      /* yieldStress0                                          */
  double Et;       /* bilinear tangent modulus, const[0]                    */
  double K_h, m_h; /* Hollomon K, m                                         */
  double A, B, n, C, eps0, m, Tm, Tt; /* Johnson-Cook const[0..7]            */
} wf_material_params;

/* Neighbour-rebuild policy.  WF_REBUILD_EVERY_N is the reference's: every
 * ns_interval steps once max eps_p > ns_min_plastic (SolverFraser.cu:166,
 * 260-271, 395-397).  WF_REBUILD_PLASTIC_ZONE (an extension) builds the
 * lists with radius 2h + ns_skin from the current positions and rebuilds,
 * once yielding, only when a particle with eps_p > ep_hot has moved more
 * than ns_skin / 2 since the last build; pairs beyond 2h contribute nothing
 * to the pair sums, and the surface pass counts only pairs within 2h. */
enum {
  WF_REBUILD_EVERY_N = 0,
  WF_REBUILD_PLASTIC_ZONE = 1,
  WF_REBUILD_PLASTIC_ZONE_LOCAL = 2 /* the same trigger; a rebuild re-lists
    in place only the tiles within two search cells of a particle with
    eps_p > ep_hot (no reorder), and a full rebuild follows every
    ns_global_every re-lists or when a tile outgrows its staging region */
};

typedef struct wf_config {
  int abi_version;       /* WF_ABI_VERSION                                  */
  int n_particles;       /* particle_count                                  */
  int solid_count;       /* solid_part_count (== n_particles, no contact)   */
  int precision;         /* WF_FP64 / WF_MIXED                              */
  double dt;             /* deltat                                          */
  double alpha, beta;    /* artifViscAlpha / artifViscBeta                  */
  double h_glob;         /* search radius is 2*h_glob (SolverFraser.cu:72)  */
  wf_material_params mat;
  int return_map;        /* WF_RETURN_*                                     */
  int quirks;            /* WF_Q_* bit set                                  */
  int mechanics;         /* 1: run the mechanical pipeline in wf_step       */
  int thermal;           /* 1: conduction + leapfrog (ThermalCalcs)         */
  double taylor_quinney; /* fraction of plastic work to heat (0 = off)      */
  double T_inf, h_conv;  /* convection BC parameters (Thermal.cu:36-64)    */
  int ns_interval;       /* rebuild every k steps once yielding (5)         */
  int surrogate;         /* 1: MLP initial guess for the Newton return map  */
  double ns_min_plastic; /* yielding threshold on max eps_p (1e-6)          */
  int newton_max_iter;   /* Newton iteration cap (default 50)               */
  int mlp_train_interval;/* online training cadence in steps (200)          */
  double newton_tol;     /* |residual| <= newton_tol * s_trial              */
  /* rigid-mesh contact (SolverFraser.cu:233-257, WeldFormGPU.cu:495-636)   */
  int contact;           /* 1: surface + contact forces + mesh motion       */
  int contact_law;       /* WF_CONTACT_*                                    */
  double contact_kappa;  /* WF_CONTACT_STIFFNESS factor (0.1)               */
  double contact_dx;     /* WF_CONTACT_STIFFNESS lattice spacing dx          */
  double fric_sta;       /* Contact[0] fricCoeffStatic (0 = no friction)    */
  double fric_dyn;       /* Contact[0] fricCoeffDynamic                      */
  int id_free_surf;      /* ID given to free-surface particles (1,
                            SolverFraser.cu:120)                            */
  int rebuild_policy;    /* WF_REBUILD_*                                    */
  double fric_heat;      /* fraction of the sliding-friction power that heats
                            the particle (0 = off; an extension)            */
  double ns_skin;        /* WF_REBUILD_PLASTIC_ZONE: list radius 2h + skin  */
  double ep_hot;         /* WF_REBUILD_PLASTIC_ZONE: eps_p of the particles
                            whose displacement triggers a rebuild (0.05)    */
  /* surrogate training (extension): each online round runs mlp_train_iters
   * Adam steps at rate mlp_lr.  Rounds are due every mlp_train_interval steps
   * while yielding; with mlp_interval_max > 0 the interval doubles, up to
   * mlp_interval_max, after a round whose starting loss (the network before
   * the round, on the round's samples) is at most mlp_refit_ratio times the
   * previous round's final loss, and returns to mlp_train_interval after a
   * round where it is not. */
  int mlp_train_iters;   /* Adam steps per round (250)                       */
  int mlp_interval_max;  /* cap of the adaptive interval (0 = fixed cadence) */
  double mlp_lr;         /* Adam learning rate (2e-3)                        */
  double mlp_refit_ratio;/* starting/previous loss ratio that still fits (2) */
  int ns_global_every;   /* WF_REBUILD_PLASTIC_ZONE_LOCAL: a full rebuild
                            (reorder) after this many re-lists (0: 32)      */
  int axisymm;           /* 1: the axisymmetric-slice acceleration clamp
                            after the acceleration pass (dom_bid_type
                            AxiSymm_3D, SolverFraser.cu:229-231; set by a
                            cylindrical-grid slice, WeldFormGPU.cu:366-375) */
  int fuse_integrate;    /* 1 (default): in contact-free, non-axisymmetric
                            mixed-precision steps with every particle solid,
                            the acceleration pass integrates each particle
                            (BC + position/velocity update) as soon as its
                            acceleration is summed; 0: a separate pass     */
  int ns_compact_max;    /* a tile whose adjacent cells hold at most this
                            many candidates gets its staged set compacted to
                            the particles its lists reference; larger sets
                            are staged whole (0: 4096, the list builder's
                            LDS flag array; larger values count as 4096)   */
} wf_config;

/* Fill *cfg with the reference's defaults: alpha=1, beta=0
 * (Domain_d.cu:191-192), ns_interval=5, ns_min_plastic=1e-6
 * (SolverFraser.cu:17,91), all quirks on, FP64, reference return map,
 * contact off with the reference penalty law, id_free_surf=1. */
void wf_config_default(wf_config* cfg);

/* ---- fields ------------------------------------------------------------- */
enum {
  /* double[3*n] */
  WF_F_X = 0, WF_F_V, WF_F_A, WF_F_U, WF_F_VA, WF_F_X_SNAPSHOT,
  WF_F_NORMAL,    /* surface normal of CalculateSurface (Geometry.cu:72)   */
  WF_F_CONTFORCE, /* contact force of the last contact call (Contact.cu:247) */
  /* double[n] */
  WF_F_RHO = 16, WF_F_RHO0, WF_F_DRHO, WF_F_M, WF_F_H, WF_F_P, WF_F_CS, WF_F_G,
  WF_F_SIGMA_Y, WF_F_PL_STRAIN, WF_F_SIGMA_EQ, WF_F_EFF_STRAIN_RATE,
  WF_F_INT_ENERGY, WF_F_T, WF_F_TA, WF_F_TB, WF_F_DTDT, WF_F_K_T, WF_F_CP_T,
  WF_F_Q_PLASTIC,
  WF_F_Q_FRICTION, /* sliding-friction heat power into the particle [W]     */
  /* double[6*n], flat symmetric order xx yy zz xy yz xz; the rotation rate
     holds the upper triangle of the antisymmetric W (InteractionAlt.cu:358-365),
     so its diagonal is always zero: values written there are stored as zero */
  WF_F_SIGMA = 48, WF_F_SHEARSTRESS, WF_F_STRAIN, WF_F_STRRATE, WF_F_ROTRATE,
  /* int32[n] */
  WF_F_ID = 64, WF_F_BC_T,
  /* int32[n], read-only: neighbour count, contact-neighbour count          */
  WF_F_NB_COUNT, WF_F_CONT_NB_COUNT,
  WF_F_COUNT_
};

/* Components per particle and element byte size of a field; 0 if unknown. */
int wf_field_shape(int field, int* components, int* elem_bytes);

/* ---- lifecycle: SetDimension (Domain_d.cu:19-234) ----------------------- */
int wf_create(const wf_config* cfg, struct wf_domain** out);
int wf_destroy(struct wf_domain* d);
const char* wf_last_error(void);
int wf_sync(struct wf_domain* d);

/* Upload / download a whole field (original order).  bytes must equal
 * n_particles * components * elem_bytes.  Replaces the Set* setters
 * (Domain_d.cu:307-477) and the cudaMemcpy's of WeldFormGPU.cu:694-841. */
int wf_set_field(struct wf_domain* d, int field, const void* host, size_t bytes);
int wf_get_field(struct wf_domain* d, int field, void* host, size_t bytes);
/* Fill a double field with a constant (SetDensity, Set_h, SetShearModulus,
 * SetSigmay ... Domain_d.cu:307-477). */
int wf_fill_field(struct wf_domain* d, int field, double value);

/* Copy x into the neighbour-search snapshot x_h, as the reference does at
 * output times (SolverFraser.cu:323-325). */
int wf_snapshot_positions(struct wf_domain* d);

/* Velocity boundary conditions applied by wf_step (UserAcc,
 * WeldFormGPU.cu:78-140; boundaryCondition, Boundary_Condition.cuh:5-14).
 * At most WF_MAX_BCS; applied in order, so the last matching zone wins. */
#define WF_MAX_BCS 16
int wf_set_bcs(struct wf_domain* d, int count, const int* zone_ids,
               const double* values /* 3*count */);

/* ---- neighbour search: cuNSearch::NeighborhoodSearch::find_neighbors
 *      (src/cuNSearch/src/cuNSearch.cu:111-137).  Radius 2*h_glob, FP32
 *      criterion 0 < |xj-xi|^2 < r^2 (cuNSearchKernels.cu:163-164).
 *      Positions come from the snapshot (WF_Q_STALE_NS_POS) or from x.
 *      Under WF_REBUILD_PLASTIC_ZONE: radius 2*h_glob + ns_skin, positions
 *      from x.  A list holds up to WF_MAX_NEIGHBORS_SKIN neighbours under
 *      either policy; particles with more than the policy's cap (70, the
 *      reference's CUDA_MAX_NEIGHBORS; WF_MAX_NEIGHBORS_SKIN for skin lists)
 *      are counted in nb_over_cap.  The reference stops a list at its 70th
 *      hit in an order that its atomic cell insertion leaves undefined
 *      (cuNSearchKernels.cu:169-173), so no fixed rule reproduces its set;
 *      here every neighbour up to the storage limit is kept. ---- */
int wf_nsearch_build(struct wf_domain* d);
/* Total number of stored neighbour entries and particles over the cap. */
int wf_nsearch_stats(struct wf_domain* d, int64_t* total, int* over_cap);
/* CSR of the current lists in ORIGINAL indices, each row sorted ascending.
 * counts[n], offsets[n+1], list[list_cap]. */
int wf_get_neighbors(struct wf_domain* d, int* counts, int* offsets, int* list,
                     int64_t list_cap);

/* ---- the per-step entry points (reference kernels) ---------------------- */
/* CalcDensIncKernel, InteractionAlt.cu:153-246                             */
int wf_calc_dens_inc(struct wf_domain* d);
/* UpdateDensityKernel, Mechanical.cu:194-203                               */
int wf_update_density(struct wf_domain* d, double dt);
/* CalcRateTensorsKernel, InteractionAlt.cu:251-383                         */
int wf_calc_rate_tensors(struct wf_domain* d);
/* PressureKernelExt, Mechanical.cu:289-298 (EOS case 0, Functions.cu:211)  */
int wf_pressure(struct wf_domain* d);
/* StressStrainKickDriftKernel -> StressStrainOne, Mechanical.cu:458-565,
 * plus the per-particle internal-energy increment of CalcIntEnergy
 * (Energy.cu:5-19, per-particle indexing corrected).                       */
int wf_stress_strain(struct wf_domain* d);
/* CalcAccelKernel, InteractionAlt.cu:5-150 (KernelType 0, XSPH 0)          */
int wf_calc_accel(struct wf_domain* d);
/* ApplyAxiSymmBCKernel, Domain_d.cu:665-695: a particle with in-plane
 * acceleration |a_xy| > 1e-3, x > h/2 and a_x > 1e-2 gets a_x = a_y = 0 (the
 * projection the reference computes there is overwritten by those zeros;
 * its printf output is not reproduced).                                    */
int wf_apply_axisymm_bc(struct wf_domain* d);
/* ApplyBCVelKernel, Mechanical.cu:572-589                                  */
int wf_apply_bc_vel(struct wf_domain* d, int zone_id, double vx, double vy,
                    double vz);
/* UpdatePosFraserKernel, Mechanical.cu:179-191                             */
int wf_update_pos_fraser(struct wf_domain* d, double dt);
/* UpdateVelKernel, Mechanical.cu:142-152                                   */
int wf_update_vel(struct wf_domain* d, double dt);
/* ThermalSolveKernel, Thermal.cu:101-142 (the GPU form of
 * Domain::CalcTempInc, Thermal_ONLY4REF_NOTUSE.cpp:10-78), using the
 * mechanical neighbour lists; then CalcConvHeatKernel (Thermal.cu:36-64)
 * and the plastic-work heat source.                                         */
int wf_calc_temp_inc(struct wf_domain* d);
/* Domain_d::ThermalCalcs, Thermal.cu:68-99: wf_calc_temp_inc followed by
 * TempCalcLeapfrogFirst / TempCalcLeapfrog (Thermal.cu:16-31).              */
int wf_thermal_calcs(struct wf_domain* d, double dt);
/* FP64 deterministic sum of the per-particle internal energy
 * (Energy.cu:17 int_energy_sum, without its race).                         */
int wf_int_energy(struct wf_domain* d, double* sum);

/* ---- fused step: Domain_d::MechFraserSolve loop body, SolverFraser.cu:164-399:
 *      rebuild policy, dens-inc + density update, rate tensors, pressure +
 *      stress, accel, BC, position/velocity update, BC, thermal, energy. */
int wf_step(struct wf_domain* d, int n_steps);

/* ---- run statistics ------------------------------------------------------ */
typedef struct wf_stats {
  int64_t step;            /* steps taken by wf_step                        */
  double time;             /* simulated time                                */
  int is_yielding;         /* max eps_p has exceeded ns_min_plastic         */
  int ns_builds;           /* neighbour rebuilds                            */
  int64_t newton_iters;    /* total Newton iterations (all particles)       */
  int64_t newton_calls;    /* particles that ran the Newton return map      */
  int64_t newton_iters_wave; /* sum over wavefronts of their max iterations */
  int mlp_trainings;       /* online training rounds                        */
  int nb_over_cap;         /* particles whose candidates exceeded the cap   */
  int cont_over_cap;       /* contact candidates beyond WF_MAX_CONT_NB      */
  double mlp_loss;         /* training loss of the last surrogate round     */
  int64_t mlp_samples;     /* logged samples available at that round        */
  int mlp_interval;        /* current training interval in steps            */
  int _pad0;
  double mlp_start_loss;   /* loss of the network at the start of that round */
  int ns_local_builds;     /* local re-lists (WF_REBUILD_PLASTIC_ZONE_LOCAL) */
  int _pad1;
  int64_t ns_local_tiles;  /* tiles re-listed by them, in total             */
} wf_stats;
int wf_get_stats(struct wf_domain* d, wf_stats* out);

/* ---- per-kernel timing with HIP events on the handle's stream ----------- */
/* Enable (1) / disable (0) event timing; resets the accumulators.           */
int wf_set_timing(struct wf_domain* d, int on);
/* Kernel ids for wf_kernel_time. */
enum {
  WF_K_DENS = 0,    /* density-rate pair loop + density update      */
  WF_K_RATES,       /* strain/rotation-rate pair loop               */
  WF_K_STRESS,      /* pressure + stress/plasticity + energy, with the
                       Newton Johnson-Cook return map and its
                       surrogate guess solved in place              */
  WF_K_ACCEL,       /* acceleration pair loop                       */
  WF_K_INTEGRATE,   /* BC + position/velocity update + BC           */
  WF_K_THERMAL,     /* conduction pair loop + leapfrog              */
  WF_K_NSEARCH,     /* neighbour rebuild (all its kernels)          */
  WF_K_MLP,         /* surrogate training                           */
  WF_K_RETURN,      /* reserved: the return map runs inside
                       WF_K_STRESS, so this id reports no launches  */
  WF_K_CONTACT,     /* surface, contact neighbours, forces, mesh move */
  WF_K_COND_DENS,   /* conduction pair loop + leapfrog fused with the
                       next step's density pass + update (inside one
                       wf_step call, when no rebuild runs between)  */
  WF_K_ACCEL_INTEGRATE, /* acceleration pair loop with each particle's BC +
                           position/velocity update (contact-free steps)  */
  WF_K_COUNT_
};
/* Accumulated milliseconds and launch count of kernel id k.  Each timed
 * kernel launch is bracketed by its own event pair (events without the
 * system-scope fence; the host does not wait per launch), read here. */
int wf_kernel_time(struct wf_domain* d, int k, double* total_ms, int64_t* launches);
/* A measurement aid with no counterpart in the reference: launch the pair
 * pass k (WF_K_RATES or WF_K_ACCEL, which only read the state their inputs
 * come from) reps times back to back between two events on the handle's
 * stream, and return the elapsed time over reps: the kernel's own duration,
 * as a kernel trace of the same launches reports it. */
int wf_time_pass(struct wf_domain* d, int k, int reps, double* ms_per_launch);

/* ---- return-map surrogate ----------------------------------------------------
 * No counterpart in the reference (an extension of the Johnson-Cook return
 * map, Mechanical.cu:507-535): a 4 -> 32 -> 32 -> 1 ReLU MLP gives the FP64
 * Newton its initial plastic increment.  Features per yielding particle:
 * s_trial/A, 10 eps_p, log10(max(rate,1e-3))/3, homologous temperature;
 * target: dg / dg_max with dg_max = (s_trial - s_y(eps_p))/(3G).  Inference
 * runs on BF16 MFMA with FP32 accumulation inside the return-map kernel;
 * with cfg.surrogate it is trained online every mlp_train_interval steps on
 * the samples the solver logs (Adam, FP32, deterministic).  Newton always
 * finishes to newton_tol, so the surrogate changes iteration counts only.
 * Parameter order: W1[4][32], b1[32], W2[32][32], b2[32], W3[32], b3. */
#define WF_MLP_PARAMS 1249
int wf_mlp_get_params(struct wf_domain* d, float* out, int n);
int wf_mlp_set_params(struct wf_domain* d, const float* in, int n);
/* zero the Adam moments and step counter */
int wf_mlp_reset_adam(struct wf_domain* d);
/* BF16-MFMA inference on n feature rows X[n][4] -> y[n] */
int wf_mlp_infer(struct wf_domain* d, const float* X, int n, float* y);
/* gradient of mean((y - Y)^2) over the rows (FP32, deterministic) */
int wf_mlp_grad(struct wf_domain* d, const float* X, const float* Y, int n, float* grad,
                float* loss);
/* iters full-batch Adam steps (lr; beta1 0.9, beta2 0.999, eps 1e-8) */
int wf_mlp_fit(struct wf_domain* d, const float* X, const float* Y, int n, int iters,
               float lr, float* loss);
/* one online round on the logged samples */
int wf_mlp_train(struct wf_domain* d, int iters, float lr, int64_t* samples, float* loss);

/* wf_mlp_grad on device arrays, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  params (WF_MLP_PARAMS floats), X (n x 4), Y (n), grad
 * (WF_MLP_PARAMS) and loss (1 float) are device pointers on the current
 * device; stream is a hipStream_t (NULL: the null stream).  The call only
 * enqueues; results are ready when the stream has reached it.  The gradient's
 * per-block partial sums (at most 256 blocks) live in a fixed device array of
 * the library (1.22 MiB per GPU, held until the process ends) instead of an
 * allocation; n must lie in 1..WF_MLP_NOALLOC_MAX_SAMPLES (WF_E_ARG
 * otherwise), and calls that overlap on one GPU race on that array.  Results
 * are bit-identical to wf_mlp_grad's on the same inputs. */
#define WF_MLP_NOALLOC_MAX_SAMPLES (1 << 24)
int wf_mlp_grad_noalloc(const float* params, const float* X, const float* Y, int n,
                        float* grad, float* loss, void* stream);

/* ---- rigid triangle meshes and penalty contact ---------------------------
 * The reference stores each rigid mesh as a device TriMesh_d (Mesh.cuh:19-54)
 * and adds one zero-mass "particle" per triangle centroid to the SPH set
 * (Domain::AddTrimeshParticles, Domain.cpp:967-983) so that contact pairs
 * come out of the SPH neighbour search.  Here meshes are kept apart from the
 * particles: nodes, connectivity, centroids, unit normals and plane
 * constants live in their own device arrays, and contact candidates come
 * from a grid over triangle centroids.  Element indices are global over all
 * meshes, in the order the meshes were added. */
#define WF_MAX_MESHES 8
#define WF_MAX_CONT_NB 20 /* MAX_NB_COUNT, Contact.cu:7 */

/* TriMesh_d::AxisPlaneMesh(2, false, p1, p2, dens), Mesh.cu:48-194: a plate
 * of (dens+1)^2 nodes spaced (p2.x-p1.x)/dens at z = p1.z, two triangles per
 * quad with normal -z.  Host only.  Fills nodes (3*(dens+1)^2) and elnode
 * (3*2*dens^2) when the caps allow; returns the element count.             */
int wf_gen_plane_mesh(const double p1[3], const double p2[3], int dens, double* nodes,
                      int node_cap, int* elnode, int elem_cap, int* n_nodes);
/* Nastran small-field triangle mesh: NastranReader::read + ReadFromNastran
 * (NastranReader.cpp:5-153, Mesh_reader.cu:5-60).  GRID cards (id, x, y, z
 * in 8-character fields 2 and 4-6, exponent letter optional: "1.5-3") and
 * CTRIA3 cards (node ids in fields 4-6) -> nodes (3*n_nodes), elnode
 * (0-based, 3 per triangle); flip reverses each triangle's orientation.
 * Host only.  Returns the triangle count or a negative WF_E_* code.        */
int wf_read_nastran(const char* path, int flip, double* nodes, int node_cap, int* elnode,
                    int elem_cap, int* n_nodes);
/* Upload a rigid mesh (WeldFormGPU.cu:495-575): nodes xyz interleaved,
 * elnode 3 per triangle (0-based, into this mesh's nodes).  zone_id is the
 * mesh's ID (contact_surf_id).  Centroids, normals (n0n1 x n0n2, unit) and
 * plane constants are computed on the device.  *mesh gets its index.       */
int wf_add_mesh(struct wf_domain* d, const double* nodes, int n_nodes, const int* elnode,
                int n_elems, int zone_id, int* mesh);
/* Rigid motion of a mesh: node velocity v + w x (x - center) (SetMeshVel,
 * Domain_d.cuh:553-556; TriMesh_d::Move uses w x x, i.e. center 0).  The
 * centre then travels with v (extension: a tool spins about its own moving
 * axis; the reference's fixed origin is the case v = 0 or w = 0).          */
int wf_set_mesh_motion(struct wf_domain* d, int mesh, const double v[3], const double w[3],
                       const double center[3]);
/* Counts of mesh m (nodes, elements, zone id); m = -1: totals over meshes. */
int wf_mesh_info(struct wf_domain* d, int mesh, int* n_nodes, int* n_elems, int* zone_id);
/* Download mesh m: nodes/node_v (3*n_nodes), centroids/normals (3*n_elems),
 * plane constants (n_elems); any pointer may be null.                      */
int wf_get_mesh(struct wf_domain* d, int mesh, double* nodes, double* node_v,
                double* centroids, double* normals, double* pplane);
/* CalculateSurfaceKernel, Geometry.cu:13-88: normal_i = sum_j m_j x_ij *
 * N/(M nb_i) with N = particles + triangles and M the total mass; particles
 * with |normal_i| >= 0.25 h_i and nb_i <= 46 get ID = id_free_surf unless
 * their ID was non-zero at the first call (not_write_surf_ID,
 * WeldFormGPU.cu:845-858).                                                 */
int wf_calc_surface(struct wf_domain* d);
/* CalcContactNbKernel, Contact.cu:15-62: for each particle with ID ==
 * id_free_surf, the triangles whose centroid lies within 2 h_i, ascending
 * element index, at most WF_MAX_CONT_NB.                                   */
int wf_calc_contact_nb(struct wf_domain* d);
/* CalcContactForcesKernel -> CalcContactForcesWang, Contact.cu:114-321:
 * predicted position, plane distance, inside-triangle test, penalty normal
 * force and static/dynamic Coulomb friction added to a; contact force kept
 * in WF_F_CONTFORCE.                                                       */
int wf_calc_contact_forces(struct wf_domain* d);
/* MeshUpdateKernel (Domain_d.cuh:545-551: Move, CalcCentroids, CalcNormals,
 * UpdatePlaneCoeff) for every mesh, then UpdateContactParticlesKernel
 * (Contact.cu:64-112): element velocity = mean of its node velocities.     */
int wf_mesh_update(struct wf_domain* d, double dt);
/* Contact candidate lists in ORIGINAL particle order: counts[n] and
 * list[n * WF_MAX_CONT_NB] (global element indices, -1 past the count).    */
int wf_get_contact_neighbors(struct wf_domain* d, int* counts, int* list);
/* Sum of |contact force| over particles with ID == id_free_surf (the
 * "Contact Force Sum" line, SolverFraser.cu:355-360), FP64, deterministic. */
int wf_contact_force_sum(struct wf_domain* d, double* sum);

/* ---- host-side input builders (no device needed) ---------------------------
 * Lattice generators and zone assignment of the host domain build, in the
 * reference's generation order and floating-point arithmetic, so particle
 * counts and positions match a reference run of the same deck.
 * Each returns the particle count; xyz (3*count doubles) is filled when
 * non-null and cap is large enough; *mass_out gets the per-particle mass. */
/* Domain::AddBoxLength, cubic packing (Domain.cpp:194-329)                  */
int wf_gen_box(const double start[3], double Lx, double Ly, double Lz, double r,
               double density, double* xyz, int cap, double* mass_out);
/* Domain::AddCylinderLength (Domain.cpp:490-574, calcHalfPartCount :473-488) */
int wf_gen_cylinder(const double start[3], double R, double Lz, double r,
                    double density, double* xyz, int cap, double* mass_out);
/* Domain::AddCylUniformLength (Domain.cpp:1071-1193), the cylindrical grid:
 * layers z = r, 3r, ... while z <= Lz (the domain's start point is not
 * used); on each, rings of radius ri = r_i, r_i + 2r, ... < R, with
 * ceil(ang ri / 2r) particles at angles k ang / count (one particle when
 * ri == 0), x = ri cos, y = ri sin.  ID (when non-null) gets tag, or, on a
 * slice (ang < 2 pi), 2 for the first `rows` and 3 for the last `rows`
 * angles of a ring; 4 for a particle within r/10 of the axis in both x and
 * y.  *mass_out = pi R^2 Lz density / count (the whole cylinder's volume,
 * also for a slice).  Full grid: ang = 2 pi, rows = 1, r_i = 0; the slice
 * the reference builds when sliceAngle is not exactly 2 pi: ang = pi/16,
 * rows = 1, r_i = R/2 (WeldFormGPU.cu:366-375).                            */
int wf_gen_cylinder_polar(int tag, double R, double Lz, double r, double density,
                          double ang, int rows, double r_i, double* xyz, int* ID,
                          int cap, double* mass_out);
/* Domain::AssignZone (Domain.cpp:989-1005): ID[a] = id for particles inside
 * the closed box [start,end]; returns how many.                             */
int wf_assign_zone(const double* xyz, int n, const double start[3],
                   const double end[3], int id, int* ID);

#ifdef __cplusplus
}
#endif
#endif /* WF_ABI_H */
