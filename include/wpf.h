/*
 * wpf.h -- C ABI of the MI355X (gfx950) wPFAW3D per-time-step hot path.
 *
 * One context owns one grid resident in HBM.  The host keeps the reference
 * module's view of the state: haloed Fortran-order arrays (index 0..n+1 on
 * every axis, x fastest), exactly as allocated in wPFAW3D.f90:440-449.  The
 * library stores interior cells only, structure-of-arrays, and applies the
 * single-rank ghost rule ghost(i,j,k) = field(clamp(i),clamp(j),clamp(k))
 * (wPFAW3D.f90:1645-1657 after parallel.f90:278-293) inside its kernels.
 *
 * Every entry point returns a status code (WPF_OK == 0).  The reference has
 * no status codes: it stops on a KSPSolve error (wPFAW3D.f90:1005) and never
 * checks convergence; here non-convergence and NaN are reported.
 *
 * Threading: one host thread per context; calls on one context are ordered
 * on that context's single HIP stream.  The caller owns host memory, the
 * library owns device memory.
 */
#ifndef WPF_H
#define WPF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WPF_ABI_VERSION 1

/* ---- status codes ---------------------------------------------------- */
#define WPF_OK          0
#define WPF_ERR_ARG     1   /* bad argument or size mismatch               */
#define WPF_ERR_HIP     2   /* HIP runtime error                            */
#define WPF_ERR_NOCONV  3   /* pressure solve hit maxit before tolerance    */
#define WPF_ERR_NAN     4   /* NaN or Inf in a solver residual              */
#define WPF_ERR_NODEV   5   /* no GPU visible                               */
#define WPF_ERR_STATE   6   /* call order (e.g. solve before permeability)  */
#define WPF_ERR_NOMEM   7   /* device allocation failed                     */
#define WPF_ERR_IO      8   /* file missing or malformed                    */

/* ---- field identifiers (haloed, Fortran order) ------------------------- */
#define WPF_FIELD_PERMEAB_ROCK 0  /* permeab_rock(0:nx+1,0:ny+1,0:nz+1)      */
#define WPF_FIELD_PRESSURE     1  /* pressure                                 */
#define WPF_FIELD_PHASE        2  /* phase                                    */
#define WPF_FIELD_CONCENT      3  /* concent                                  */
#define WPF_FIELD_DPHASEDT     4  /* dphasedt                                 */
#define WPF_FIELD_DCONCENTDT   5  /* dconcentdt                               */
#define WPF_FIELD_PERMEAB      6  /* permeab(0:3,...), download only          */
#define WPF_FIELD_VELOCITY     7  /* velocity(1:3,...), download only         */
/* The two above are rebuilt on the host from the device's phase, k_rock and
 * pressure in the reference's AoS layout (every face 0..n, halos included).
 * The next two are the device's own arrays, (1:3, 0:nx+1, 0:ny+1, 0:nz+1)
 * with the +x/+y/+z faces of each interior cell and 0 on the halo: */
#define WPF_FIELD_FACE_T       8  /* t = permeab(face)/viscosity, 0 on the
                                     global edge (the pressure operator)    */
#define WPF_FIELD_FACE_V       9  /* face velocities of calcVelocity        */

/* The subset of Param_type (wPFAW3D.f90:27-76) that the hot path reads.
 * Global and local sizes coincide: one GPU holds the whole grid. */
typedef struct wpf_params {
    int32_t nx, ny, nz;             /* nx_g, ny_g, nz_g                      */
    int32_t ix_ref, iy_ref, iz_ref; /* reference cell (1-based)              */
    double dtime;                   /* dt of steps 1-2 (cfg "dt")            */
    double tfac;                    /* dt factor (cfg "tfac")                */
    double epsilon;                 /* interface width [m] (cfg "w")         */
    double dx;                      /* grid spacing [m] = dxw*w              */
    double lambda;                  /* phi-c coupling                        */
    double porosity;
    double rhos, rhoa;
    double viscosity;               /* cfg "visco"                           */
    double cwtp;
    double diffusion;               /* cfg "D"                               */
    double b100;
    double ral, rat;                /* dispersivities                        */
    double permeab_hole;            /* cfg "kwh"                             */
    double alpha;                   /* 0.5/b/kc (wPFAW3D.f90:397-399)        */
    double pini;
} wpf_params;

/* Well_type (wPFAW3D.f90:82-90); indices are 1-based global cell indices. */
typedef struct wpf_well {
    int32_t ix_g, iy_g, iz_g, pad;
    double pressure;   /* Dirichlet value; a well is Dirichlet iff != 0      */
    double injection;  /* injection rate; an inlet iff != 0                  */
    double phase;
    double concent;
} wpf_well;

/* ---- pressure solver -------------------------------------------------- */
#define WPF_SOLVER_CG_JACOBI 0  /* FP64 matrix-free CG + Jacobi              */
#define WPF_SOLVER_MGCG      1  /* FP64 CG + geometric-multigrid V-cycle     */
#define WPF_SOLVER_MIXED     2  /* FP32 MG-PCG inside FP64 refinement        */

#define WPF_FLAG_LEARNED_WARM  1  /* learned first stage (needs a net)      */
#define WPF_FLAG_LEARNED_FCG   2  /* WPF_SOLVER_CG: the learned stage inside
                                     flexible CG; other solvers: the first
                                     stage, as WPF_FLAG_LEARNED_WARM        */
#define WPF_FLAG_EXTRAPOLATE   4  /* linear-in-time initial guess            */
#define WPF_FLAG_CAPTURE       8  /* keep (input, correction) for training   */
#define WPF_FLAG_CHEBYSHEV    16  /* degree-2 Chebyshev smoothing (sweeps 2) */
#define WPF_FLAG_OP16         32  /* mixed: the V-cycle's fine faces stored
                                     as FP16 (CG's matvec keeps FP32 faces) */
#define WPF_FLAG_F32_STENCIL  64  /* wpf_step: phase/concentration stencils
                                     in FP32, FP64 fields, update and taumin */
#define WPF_FLAG_PROJECT     128  /* mixed: initial guess projected onto the
                                     last 5 solution changes (A-norm optimal) */
#define WPF_FLAG_SPLIT_TAIL  256  /* multigrid: the last two levels as five
                                     separate launches instead of one (the
                                     same values; kept for comparison)      */

typedef struct wpf_solver_opts {
    int32_t solver;       /* WPF_SOLVER_*                                    */
    int32_t maxit;        /* Krylov iteration cap (reference: 1000)          */
    double  rtol;         /* reference: 1e-10                                */
    double  atol;         /* PETSc default 1e-50                             */
    int32_t mg_levels;    /* multigrid levels including the fine one (5)     */
    int32_t mg_sweeps;    /* damped-Jacobi sweeps before and after (2)       */
    double  mg_omega;     /* Jacobi damping                                  */
    double  mg_coarse_scale; /* weight of the coarse-grid correction         */
    double  inner_rtol;   /* mixed: FP32 inner reduction per refinement step */
    int32_t inner_maxit;  /* mixed: FP32 iterations cap per refinement step  */
    int32_t flags;        /* WPF_FLAG_*                                      */
} wpf_solver_opts;

/* Residual norms are Jacobi-preconditioned 2-norms of the reference's own
 * system (Dirichlet rows included), the norm PETSc's default test uses for
 * a left-preconditioned Krylov method: converged when
 * rnorm <= max(rtol*bnorm, atol). */
typedef struct wpf_solve_stats {
    int32_t iterations;   /* Krylov iterations (FP32 inner total if mixed)   */
    int32_t outer;        /* refinement steps (mixed), 0 otherwise           */
    int32_t converged;    /* 1 when the test above holds on the true residual*/
    int32_t learned;      /* learned first stage: 1 when accepted; flexible
                             CG: the number of learned stage applications  */
    double  rnorm0;       /* initial true residual                           */
    double  rnorm;        /* final true residual                             */
    double  bnorm;        /* right-hand side                                  */
    double  seconds;      /* wall time of the solve (host clock)             */
} wpf_solve_stats;

/* showProgress reductions (wPFAW3D.f90:1695-1810) */
typedef struct wpf_diag {
    double pv;            /* dx^3*poro*0.5*sum(1-phase)   (1:1695)           */
    double vcal;          /* dx^3*(1-poro)*0.5*sum(1+phase)  (1:1776)        */
    double rwhmax;        /* max distance of phase>0 cells from ref cell     */
    double phase_max, phase_min, concent_max, concent_min; /* haloed arrays  */
    double pinj;          /* pressure at the reference cell                  */
} wpf_diag;

typedef struct wpf_ctx wpf_ctx;

/* ---- lifecycle --------------------------------------------------------- */
/* Build a context: allocate the device fields and derive the per-cell type
 * mask from lw(0:nx+1,0:ny+1,0:nz+1) and the well table, the state that
 * loadConfig builds (wPFAW3D.f90:451-461).  Fields start as the reference
 * initialises them (1:463-471) with pressure = pini.  dev < 0 keeps the
 * current device. */
int wpf_create(const wpf_params* p, const int32_t* lw_haloed,
               const wpf_well* wells, int32_t nwell, int32_t dev,
               wpf_ctx** out);
int wpf_destroy(wpf_ctx* ctx);
int wpf_device_count(int32_t* n);
const char* wpf_status_string(int32_t status);
/* Text of the last error on this thread (empty if none). */
const char* wpf_last_error(void);

/* ---- host <-> device ----------------------------------------------------- */
/* Copy a haloed host field in.  Only interior cells are read. */
int wpf_upload(wpf_ctx* ctx, int32_t field, const double* host_haloed);
/* Copy a field out, halos filled as updateHalo leaves them. */
int wpf_download(wpf_ctx* ctx, int32_t field, double* host_haloed);

/* ---- the reference's module procedures ---------------------------------- */
/* calcPermeability, wPFAW3D.f90:561-599: node permeability and harmonic
 * face means; also prepares the matrix-free pressure operator. */
int wpf_calc_permeability(wpf_ctx* ctx);
/* solvePressure, wPFAW3D.f90:603-1160 (assembly 630-936, KSPSolve
 * 989-1005, scatter 1016-1050): matrix-free, same linear system. */
int wpf_solve_pressure(wpf_ctx* ctx, const wpf_solver_opts* opts,
                       wpf_solve_stats* stats);
/* calcVelocity, wPFAW3D.f90:1164-1193. */
int wpf_calc_velocity(wpf_ctx* ctx);
/* calcPhaseDeriv(taumin), wPFAW3D.f90:1197-1475 including the MPI_MIN
 * reduction of taumin at 1:1439-1440. */
int wpf_calc_phase_deriv(wpf_ctx* ctx, double* taumin);
/* No counterpart in the reference: wpf_calc_phase_deriv for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same kernels and the same
 * results; its per-block taumin minima and the taumin key go to a fixed
 * device array of 8 MiB per GPU (held until the process ends) instead of
 * the context's reduction buffers, and the call makes no allocation.  Grids
 * that need more than 2^20 - 1 blocks of 64 x 4 cells
 * (ceil(nx/64)*ceil(ny/4)*nz) are rejected with WPF_ERR_ARG.  All calls on
 * one GPU share the array: calls that overlap on one GPU (several host
 * threads or contexts at once) race. */
int wpf_calc_phase_deriv_noalloc(wpf_ctx* ctx, double* taumin);
/* calcConcentDeriv, wPFAW3D.f90:1554-1624. */
int wpf_calc_concent_deriv(wpf_ctx* ctx);
/* evolvePhase / evolveConcent, wPFAW3D.f90:1628-1640 (forward Euler). */
int wpf_evolve_phase(wpf_ctx* ctx, double dtime);
int wpf_evolve_concent(wpf_ctx* ctx, double dtime);
/* updateHalo, wPFAW3D.f90:1645-1657.  Ghosts are clamped inside every
 * kernel, so this only validates the field id. */
int wpf_update_halo(wpf_ctx* ctx, int32_t field);
/* getPressure, wPFAW3D.f90:2151-2166 (1-based global indices). */
int wpf_get_pressure(wpf_ctx* ctx, int32_t ix_g, int32_t iy_g, int32_t iz_g,
                     double* p);
/* showProgress reductions, wPFAW3D.f90:1695-1810. */
int wpf_diagnostics(wpf_ctx* ctx, wpf_diag* d);

/* ---- fused time step ----------------------------------------------------- */
/* One pass of the loop body of wPFAW3D.f90:2225-2294: the dt rule
 * (2227-2230), calcPermeability, solvePressure, then velocity, phase and
 * concentration derivatives and both Euler updates fused in one kernel.
 * itime is the 1-based step number; *time and *dtime are updated as the
 * reference updates them; *taumin carries the previous step's value in and
 * this step's value out. */
int wpf_step(wpf_ctx* ctx, const wpf_solver_opts* opts, int32_t itime,
             double* time, double* dtime, double* taumin,
             wpf_solve_stats* stats);

/* Breakthrough: first step at which a non-well interior cell in the plane
 * ix_g = nx-2 holds phase > 0.  Returns 0 while it has not happened.
 * Checked by wpf_step after every Euler update.  The plane is two before
 * the outlet wells: wells hold their phase, and an outlet plane held at
 * phase = -1 keeps the interface's zero level about 1.5 cells away from
 * it, so the plane next to it approaches 0 from below without crossing. */
int wpf_breakthrough_step(wpf_ctx* ctx, int32_t* step);
/* Move the breakthrough plane to ix_g (1-based) and clear the recorded step.
 * The reference has no breakthrough event; this is the workload's own
 * definition, applied identically to the oracle. */
int wpf_set_breakthrough_plane(wpf_ctx* ctx, int32_t ix_g);

/* ---- input pipeline (native callers) -------------------------------------
 * loadConfig (wPFAW3D.f90:264-505) in two calls: wpf_config_load parses the
 * .cfg and its well file into per-thread staging and returns the parameters,
 * the well count, the start time (from the field file header when "init"
 * names one, else 0) and time_total; wpf_config_wells / wpf_config_fields
 * copy the wells and the initial haloed state (lw, permeab_rock, phase,
 * concent, as loadConfig + loadField leave them, 1:440-557) into arrays the
 * caller allocated from those sizes.  Returns WPF_ERR_IO where loadConfig
 * calls exit(1). */
int wpf_config_load(const char* cfg_path, wpf_params* p, int32_t* nwell,
                    double* time_start, double* time_total);
int wpf_config_wells(wpf_well* wells);
int wpf_config_fields(int32_t* lw_haloed, double* krock_haloed,
                      double* phase_haloed, double* concent_haloed);
/* The output counts of the last loaded .cfg: gid_output (n_res_out, results
 * per time_total, 1:318-320) and scr_output (n_dat_out, 1:321-323); 0 when
 * the key is absent. */
int wpf_config_output(int32_t* n_res_out, int32_t* n_dat_out);
/* saveField, wPFAW3D.f90:2096-2141 (haloed host arrays in). */
int wpf_field_write(const char* path, double time, const wpf_params* p,
                    const double* krock_haloed, const double* phase_haloed,
                    const double* concent_haloed);
/* saveResult, wPFAW3D.f90:1842-2091: the GiD post-processing binary
 * (GiDPostEx1.1).  first != 0 creates the file with the hexahedral mesh
 * (nodes every dx, one cell-centre Gauss point); every call then appends the
 * results Pressure, Phase-field, Concentration and Permeability (the node
 * value permeab(0,...), as FP32) at `time`.  Haloed host arrays in; the
 * file is closed after each call, which makes doneResult a no-op. */
int wpf_result_write(const char* path, double time, const wpf_params* p, int32_t first,
                     const double* pressure_haloed, const double* phase_haloed,
                     const double* concent_haloed, const double* permeab0_haloed);

/* ---- learned pressure correction ------------------------------------------
 * No counterpart in the reference (its solve is PETSc's).  A 3-D conv
 * encoder-decoder (4 levels, 16/32/48/64 channels, 3x3x3, BF16 on MFMA with
 * FP32 accumulation) maps (log k, pressure, Jacobi-scaled residual) at the
 * start of a solve to a pressure correction.  With WPF_FLAG_LEARNED_WARM and
 * a network attached, the solver takes that correction as a first search
 * direction with an exact line search (the FP64 solvers on the FP64
 * operator, keeping the step only if the FP64 residual drops; the mixed
 * solver on its FP32 CG operator, whose exact line search cannot raise the
 * error's A-norm), and then runs its ordinary iteration to the same FP64
 * tolerance: the network can speed a solve up, never decide it.  With an
 * extrapolated guess the mixed solver gates the stage: it runs in the first
 * and third of every 100 solves and otherwise only while those solves took
 * at least 3% less time than the second and fourth.
 * With WPF_SOLVER_CG and WPF_FLAG_LEARNED_FCG the solve is flexible CG
 * (truncation 1) with the Jacobi preconditioner, where every period-th
 * preconditioner application (wpf_set_fcg_period, default 16; the first
 * included) is the learned stage: y = network correction for the current
 * residual, s = (r, y)/(y, Ay), z = s y + D^-1 (r - s A y).  It stops on the
 * same FP64 residual test; if that pass stalls (the residual drops by less
 * than 10% over 32 iterations) or does not converge, plain CG continues from
 * its iterate without the network.
 * Grid sizes used with the network must be multiples of 32.
 * Fields for wpf_net_forward / _gradient / _train are planar FP32
 * [3][nx*ny*nz] (x fastest); targets [nx*ny*nz]; the output is the level-1
 * correction at (nx/4, ny/4, nz/4). */
typedef struct wpf_net wpf_net;
int64_t wpf_net_num_params(void);
int wpf_net_create(int32_t seed, wpf_net** out);
int wpf_net_destroy(wpf_net* net);
int wpf_net_get_params(wpf_net* net, float* host, int64_t count);
int wpf_net_set_params(wpf_net* net, const float* host, int64_t count);
int wpf_net_forward(wpf_net* net, int32_t nx, int32_t ny, int32_t nz,
                    const float* fields, float* out);
int wpf_net_gradient(wpf_net* net, int32_t nx, int32_t ny, int32_t nz,
                     const float* fields, const float* target, double* loss,
                     float* grad);
int wpf_net_train(wpf_net* net, int32_t nx, int32_t ny, int32_t nz,
                  int32_t nsamples, const float* fields, const float* targets,
                  int32_t epochs, float lr, double* epoch_losses);
/* Energy-norm training.  For a correction u (the output upsampled to full
 * resolution, 0 on Dirichlet cells) and the error t it should approximate,
 * the exact line search along u removes the share (u, At)^2/((u, Au)(t, At))
 * of the error's A-norm; the loss is 1 minus that share, the quantity the
 * solver's learned stage gains.  A is the 7-point operator of the pair:
 * faces [3][nx*ny*nz] (+x/+y/+z face of each cell, 0 on the global edge, any
 * common scale) and fixed[n] (1 on Dirichlet cells).  wpf_net_gradient_energy
 * is the loss and gradient of one host-given pair. */
int wpf_net_gradient_energy(wpf_net* net, int32_t nx, int32_t ny, int32_t nz, const float* fields,
                            const float* target, const float* faces, const uint8_t* fixed,
                            double* loss, float* grad);
/* A training set kept on the device: room for max_pairs pairs and max_ops
 * operators on one grid; wpf_net_dataset_add appends every stride-th pair
 * captured by the context's last flexible-CG solve (WPF_FLAG_LEARNED_FCG |
 * WPF_FLAG_CAPTURE) with the operator of that solve (call it before the
 * context's next permeability pass), and returns the pair count; wpf_net_train_energy runs epochs over the set in a seeded shuffled
 * order, one Adam step per pair, and returns each epoch's mean loss. */
int wpf_net_dataset_reserve(wpf_net* net, int32_t nx, int32_t ny, int32_t nz, int32_t max_pairs,
                            int32_t max_ops);
int wpf_net_dataset_add(wpf_net* net, wpf_ctx* ctx, int32_t stride, int32_t* npairs);
int wpf_net_train_energy(wpf_net* net, int32_t epochs, float lr, int32_t seed, double* epoch_losses);
int wpf_attach_net(wpf_ctx* ctx, wpf_net* net);
/* After a solve with WPF_FLAG_CAPTURE: the network input of its start
 * state (planar [3][n]) and the correction the solver found (n values). */
int wpf_capture_get(wpf_ctx* ctx, float* fields, float* target);
/* After a CG + Jacobi solve with WPF_FLAG_LEARNED_FCG | WPF_FLAG_CAPTURE:
 * the network input at learned-stage points of flexible CG and the
 * correction from the iterate there to the final solution, the training
 * pairs of the stage: of every second stage point, the first and a uniform
 * sample (fixed generator) of up to 15 later ones; count, then pair j
 * (fields [3][n], target n). */
int wpf_capture_fcg_count(wpf_ctx* ctx, int32_t* count);
int wpf_capture_fcg_get(wpf_ctx* ctx, int32_t j, float* fields, float* target);

/* ---- measurement ---------------------------------------------------------
 * Time every launch of one kernel with HIP events on the context's stream
 * between wpf_profile_start and wpf_profile_stop; *total_ms is the summed
 * kernel time and *count the number of launches timed. */
#define WPF_KERNEL_PERM        0  /* k_perm: permeability + operator        */
#define WPF_KERNEL_STEP_FIELDS 1  /* k_step_fields: fused stencil update    */
#define WPF_KERNEL_SMOOTH      2  /* fine-level smoothing steps (post, and
                                     the first pre-step of each solve)      */
#define WPF_KERNEL_MATVEC      3  /* fine-level matvec + dot                */
#define WPF_KERNEL_TRUE_RESID  4  /* FP64 true residual                     */
#define WPF_KERNEL_RESTRICT    5  /* fine-level residual + restriction      */
#define WPF_KERNEL_PROLONG     6  /* (fused into the smoothing: no launches) */
#define WPF_KERNEL_CG_UPDATE   7  /* fine-level CG update (fused with the
                                     first pre-smoothing step)              */
int wpf_profile_start(wpf_ctx* ctx, int32_t kernel);
int wpf_profile_stop(wpf_ctx* ctx, double* total_ms, int32_t* count);
/* Streaming bandwidth of the device: a kernel that reads `nread` (0..10)
 * and writes `nwrite` (1..3) arrays of n floats (n a multiple of 4), four
 * 16-byte vectors per thread and array with every load issued before its
 * use and non-temporal stores, timed over `reps` launches after 3 untimed
 * ones.  *us is the average launch time.  A measurement utility with no
 * counterpart in the reference: it gives the rate a multi-array kernel can
 * reach, beside the 8 TB/s data-sheet peak. */
int wpf_stream_probe(int32_t device, int32_t nread, int32_t nwrite, int64_t n, int32_t reps, double* us);
/* Planes per tile of the context's vector solver kernels: > 0 fixes it,
 * 0 restores the default (8; 4 below 64 planes), -1 asks for the autotuner
 * (the fastest of 2..32 planes, timed on the next solve).  Reductions are
 * partitioned by tile, so the depth changes results in the last bits; the
 * default keeps them independent of timing. */
int wpf_set_tile_depth(wpf_ctx* ctx, int32_t planes);
/* Every period-th preconditioner application of flexible CG with the
 * learned stage (WPF_SOLVER_CG_JACOBI + WPF_FLAG_LEARNED_FCG) is the stage
 * (16 by default; 1 = every application). */
int wpf_set_fcg_period(wpf_ctx* ctx, int32_t period);

/* ---- ensembles -------------------------------------------------------------
 * No counterpart in the reference (which runs one realisation per job): many
 * independent grids of different sizes on one GPU, stepped together.  Every
 * operation of a time step is one launch over the blocks of all members
 * (a fixed, tuned grid of workgroups walking a table of member blocks); the
 * pressure solve runs a schedule of FP64 refinement stages and FP32 MG-PCG
 * iterations (fixed, or adapted between steps) with per-member convergence
 * decided on the device, so a whole step is replayed from captured HIP
 * graphs with one host synchronisation.  A member the schedule leaves unconverged is finished by
 * the single-grid solver before its fields are advanced.  Each member
 * follows the arithmetic of the same grid stepped alone with wpf_step and
 * the same options (WPF_SOLVER_MIXED; WPF_FLAG_CHEBYSHEV, _OP16,
 * _EXTRAPOLATE and _PROJECT honoured; FP64 stencils; no learned stage).
 *
 * Members are contexts created by wpf_create on one device, each nx a
 * multiple of 4; they stay owned by the caller and must outlive the
 * ensemble.  itime/time/dtime give each member's time-loop state (steps
 * done, time, dt of the last step); wpf_ensemble_sync returns it and hands
 * the members back for single-grid use. */
typedef struct wpf_ensemble wpf_ensemble;
typedef struct wpf_ensemble_opts {
    int32_t stages;       /* refinement stages per solve in the schedule      */
    int32_t inner_iters;  /* FP32 iterations per stage in the schedule; with
                             both 0 the schedule adapts between steps: it
                             starts at 3 x 16, a stage grows when a member
                             used all of it, and every 8 steps each stage
                             shrinks to its most used + 2 iterations, with
                             one spare stage of 4                           */
    int32_t grid_blocks;  /* workgroups of every batched launch (0: tune)     */
    int32_t tile_depth;   /* planes per tile of the vector kernels (0: tune)  */
} wpf_ensemble_opts;
typedef struct wpf_ensemble_stats {
    int32_t steps;        /* steps of every member                            */
    int32_t fallbacks;    /* member solves finished by the single-grid solver */
    int64_t iterations;   /* FP32 inner iterations, all members and steps     */
    double  seconds;      /* wall time of the steps                           */
    int32_t stages;       /* the schedule after the last step: stages and     */
    int32_t inner_iters;  /* FP32 iteration slots summed over them            */
} wpf_ensemble_stats;
int wpf_ensemble_create(wpf_ctx* const* members, int32_t n, const wpf_solver_opts* opts,
                        const wpf_ensemble_opts* eopts, const int32_t* itime,
                        const double* time, const double* dtime, wpf_ensemble** out);
/* Autotuner: times the batched smoothing sweep over the members for each
 * candidate (workgroups, planes per tile) and keeps the fastest. */
int wpf_ensemble_tune(wpf_ensemble* e, int32_t* grid_blocks, int32_t* tile_depth);
int wpf_ensemble_step(wpf_ensemble* e, int32_t nsteps, wpf_ensemble_stats* stats);
int wpf_ensemble_sync(wpf_ensemble* e, int32_t* itime, double* time, double* dtime,
                      double* taumin);
int wpf_ensemble_destroy(wpf_ensemble* e);

#ifdef __cplusplus
}
#endif
#endif /* WPF_H */
