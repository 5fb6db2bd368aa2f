/*
 * cfx.h — C ABI of the MI355X (gfx950) explicit step of the multiblock
 * compressible multispecies Navier–Stokes solver.
 *
 * The reference has no C ABI: its step path sits behind host free functions
 * launched from Driver::steady_simulation (Driver.cu:113-192) and device
 * "scheme" objects selected by integer tags in setup_schemes
 * (Driver.cu:293-327).  Each entry point below names the reference routine(s)
 * it replaces.  A caller keeps the reference's Driver loop shape: create a
 * context from Parameter/Mesh/ChemData-equivalent descriptors, upload the
 * initial primitive field, then call the step pieces (or cfx_step) per
 * iteration.
 *
 * Conventions
 *   - Every call returns int: 0 = OK, >0 = a hipError_t passed through,
 *     <0 = a domain error (CFX_E_*).  cfx_last_error() gives a message.
 *   - Host arrays are owned by the caller and only read during the call.
 *   - The context owns all device memory and one HIP stream; calls are
 *     stream-ordered and asynchronous, except the ones that hand data back
 *     to the host (cfx_download, cfx_residual, cfx_get_dt, cfx_check,
 *     cfx_time_step with a non-NULL dt pointer), which synchronise.
 *   - A context is not thread-safe: one host thread per context.
 *   - No C++ exception crosses the ABI.
 *
 * Data layout of host arrays handed to cfx_upload_* / returned by
 * cfx_download: structure-of-arrays over the ghost-padded block, i fastest,
 * then j, then k; point (i,j,k) with i in [-ngg, mx+ngg) etc. sits at
 *   ((k+ngg)*(my+2ngg) + (j+ngg))*(mx+2ngg) + (i+ngg)
 * and variable v of a multi-variable field at v*n_padded + that index.
 * Variable orders follow the reference's DZone (Field.h:46-48):
 *   cv: rho, rho*u, rho*v, rho*w, rho*E, rho*Y_1..rho*Y_ns
 *   bv: rho, u, v, w, p, T
 */
#ifndef CFX_H
#define CFX_H

#ifdef __cplusplus
extern "C" {
#endif

#define CFX_ABI_VERSION 1
#define CFX_MAX_SPEC 16

/* domain errors */
#define CFX_E_ARG        (-1) /* bad argument                                  */
#define CFX_E_NEWTON     (-2) /* Newton temperature inversion did not converge */
#define CFX_E_NONFINITE  (-3) /* non-finite or non-positive rho / T detected   */
#define CFX_E_SURROGATE  (-4) /* surrogate transport requested but not ready   */
#define CFX_E_UNSUPPORTED (-5) /* configuration not compiled in (e.g. n_spec) */

/* boundary condition type labels, as in the reference (BoundCond.cu:254) */
#define CFX_BC_WALL      2
#define CFX_BC_SYMMETRY  3 /* slip wall: mirror with the normal velocity negated */
#define CFX_BC_INFLOW    5
#define CFX_BC_OUTFLOW   6

/* quirk bits: 1 = reproduce the reference text, 0 = corrected form */
#define CFX_Q_METRIC_AVG   0x1 /* Q5: fv/hv average only xi_x over the face (ViscousScheme.cu:16-24,238-246) */
#define CFX_Q_HV_RHOD_J    0x2 /* Q6: hv averages rho*D with (i,j+1,k) (ViscousScheme.cu:317) */
#define CFX_Q_NEWTON_REL   0x4 /* Q8: Newton stops at |1-T1/T|<=1e-3 (Thermo.cu:104,119) instead of |dT|<=1e-10 K */
#define CFX_QUIRKS_REFERENCE (CFX_Q_METRIC_AVG | CFX_Q_HV_RHOD_J)

/* precision modes */
#define CFX_FP64   0 /* plain: everything in FP64 */
#define CFX_MIXED  1 /* FP32 reconstruction/fluxes/transport; FP64 Newton, RK and reductions;
                     * Y is carried in the FP64 conserved state only (rho Y_l) */

/* transport models */
#define CFX_TRANSPORT_WILKE 0
#define CFX_TRANSPORT_MLP   1

/* what cfx_download returns */
#define CFX_CV     0 /* n_var x padded          */
#define CFX_BV     1 /* 6 x padded              */
#define CFX_YK     2 /* n_spec x padded; CFX_MIXED keeps no FP64 copy of Y
                        * after the initial upload and returns (rho Y_l) / rho
                        * from the conserved state (equal, at updated points,
                        * to the value an FP64 copy would hold) */
#define CFX_PROPS  3 /* (4 + n_spec) x padded: gamma, c, mu, kappa, rho*D_1..ns */
#define CFX_DQ     4 /* n_var x interior (mx*my*mz, i fastest) */
#define CFX_METRIC 5 /* 10 x padded: xi_x,xi_y,xi_z,eta_x..,zeta_z, jacobian */
#define CFX_DTLOCAL 6 /* interior: local time step of the last cfx_time_step */
#define CFX_COORD  7 /* 3 x padded: x, y, z vertex coordinates with the ghost grid */

/* One physical boundary face, as a line of the reference's boundary file
 * (Mesh.cpp:616-636, Boundary ctor Mesh.cpp:10-12): 0-based vertex ranges,
 * start == end in the face-normal direction. */
typedef struct {
  int range_start[3];
  int range_end[3];
  int type;  /* CFX_BC_* */
  int cond;  /* index into cfx_params.inflow / cfx_params.wall for that type */
} cfx_bc;

/* One block-to-block interface, as two lines of the reference's inner file
 * (Mesh.cpp:649-664, InnerFace ctor Mesh.cpp:34-42): 1-based, signed ranges
 * in the Plot3D/Gridgen convention, target block 1-based. */
typedef struct {
  int src[6];       /* x1 x2 y1 y2 z1 z2 of this block            */
  int tar[6];       /* x1 x2 y1 y2 z1 z2 of the target block       */
  int target_block; /* 1-based, as in the file                     */
} cfx_inner;

/* One grid block (Block, Mesh.h:109-159): vertex coordinates of the
 * mx*my*mz points (i fastest), boundaries and interfaces. */
typedef struct {
  int mx, my, mz;
  const double *x, *y, *z;
  int n_bc;
  const cfx_bc *bc;
  int n_inner;
  const cfx_inner *inner;
} cfx_block;

/* Inflow record (Inflow, BoundCond.h:14-48): give (mach, pressure,
 * temperature) or (density, velocity, pressure); unset fields are < 0 as in
 * the reference defaults. (u,v,w) is the unit direction. */
typedef struct {
  double mach, pressure, temperature, velocity, density;
  double u, v, w;
  const double *yk; /* n_spec mass fractions; NULL on the air path */
} cfx_inflow;

/* Wall record (Wall, BoundCond.h:50-58) */
typedef struct {
  int adiabatic;      /* 0 isothermal, 1 adiabatic */
  double temperature; /* wall temperature for isothermal */
} cfx_wall;

/* Scheme tags and constants (DParameter, DParameter.h:9-39) plus the
 * build's own switches. */
typedef struct {
  int inviscid_scheme;  /* 3 = AUSM+ (only scheme of the reference)      */
  int reconstruction;   /* 2 = MUSCL kappa=1/3, otherwise first order     */
  int limiter;          /* 0 = minmod, 1 = Van Albada                      */
  int viscous_order;    /* 2 = 2nd-order central, otherwise inviscid      */
  int temporal_scheme;  /* kept for the reference's tag; the stepping is chosen by rk_stages */
  int rk_stages;        /* 1 = explicit Euler with local dt (reference), 3 = SSP-RK3 with global CFL dt */
  double cfl, prandtl, schmidt;
  double gamma_air, r_air; /* air path (n_spec == 0) constants; reference: 1.4, R_u/28.9647 */
  int precision;        /* CFX_FP64 | CFX_MIXED */
  int transport;        /* CFX_TRANSPORT_WILKE | CFX_TRANSPORT_MLP */
  int quirks;           /* CFX_Q_* bits */
  int n_inflow;
  const cfx_inflow *inflow;
  int n_wall;
  const cfx_wall *wall;
  /* Implicit step: lusgs = 1 replaces the explicit update by one matrix-free
   * LU-SGS solve per step with the local time step (rk_stages must be 1).
   * The reference declares the scheme (temporal_scheme 1, TemporalScheme.cuh:19-21,
   * Field.cu:138-141) but gives it no body, so its tag alone keeps the explicit
   * Euler step there and here.  lusgs_omega is the over-relaxation factor of
   * the inviscid spectral radius (<= 0: 1.0). */
  int lusgs;
  double lusgs_omega;
} cfx_params;

/* Species data (Species, ChemData.h:11-46).  NASA-7 coefficient arrays are
 * n_spec x 7 row-major as read by read_therm (ChemData.cpp:106-228); eps_k
 * and sigma are the Lennard-Jones well depth [K] and diameter [Angstrom]
 * read by read_tran (ChemData.cpp:230-252).  n_spec == 0 selects the
 * calorically perfect air path. */
typedef struct {
  int n_spec;
  const double *mw;                       /* kg/kmol */
  const double *t_low, *t_mid, *t_high;   /* K */
  const double *coeff_lo, *coeff_hi;      /* n_spec*7 */
  const double *eps_k, *sigma;
} cfx_species;

typedef struct cfx_ctx cfx_ctx;

/* Library version / capability query. */
int cfx_version(void);

/* Host-only helpers (no device, no context): what the reference's setup
 * computes before the step loop, for callers that build the initial field
 * themselves (an input-deck reader).
 *
 * cfx_ghost_grid: the ghost-padded vertex coordinates of every block after
 * the ghost-grid construction (Mesh::init_ghost_grid, Mesh.cpp:718-922) and
 * the trim to ngg layers (Block::trim_abundant_ghost_mesh, Mesh.cpp:501-515):
 * xyz[b] receives 3 x n_padded doubles (x, then y, then z) in the padded
 * layout above.  Blocks are described as for cfx_create.
 * cfx_inflow_state: the primitive state (rho, u, v, w, p, T) and mass
 * fractions of inflow record `index` as the reference derives it
 * (Inflow::register_boundary_condition, BoundCond.cpp:10-106; var_info,
 * BoundCond.cpp:108-110); yk may be NULL. */
int cfx_ghost_grid(const cfx_block *blocks, int n_block, double *const *xyz);
int cfx_inflow_state(const cfx_params *params, const cfx_species *species, int index, double bv[6],
                     double *yk);
/* cfx_parse_numbers: the whitespace-separated decimal numbers of text[0..len)
 * as operator>> reads the grid file's coordinates (Mesh.cpp:587-607), up to
 * n_max of them, into out; *n_out = the count.  A token that is not a whole
 * number gives CFX_E_ARG with *n_out = its index.  Host only, multi-threaded. */
int cfx_parse_numbers(const char *text, long long len, double *out, long long n_max, long long *n_out);

/* Driver ctor + Field::setup_device_memory + DParameter + ghost grid and
 * metrics (Driver.cu:17-57, Field.cu:90-148, DParameter.cu:4-45,
 * Mesh.cpp:718-872 and 226-488). */
int cfx_create(const cfx_params *params, const cfx_species *species,
               const cfx_block *blocks, int n_block, cfx_ctx **out);
void cfx_destroy(cfx_ctx *ctx);
const char *cfx_last_error(const cfx_ctx *ctx);

/* Upload the primitive state of one block over the padded range (the
 * reference fills ghosts too: Field.cu:48-75). bv is 6 x padded, yk is
 * n_spec x padded (ignored on the air path). */
int cfx_upload_primitive(cfx_ctx *ctx, int blk, const double *bv, const double *yk);
/* Upload the conservative field directly (checkpoint restore), n_var x padded,
 * together with bv and yk. */
int cfx_upload_state(cfx_ctx *ctx, int blk, const double *cv, const double *bv, const double *yk);
int cfx_download(cfx_ctx *ctx, int blk, int what, double *dst);

/* Driver::initialize_computation (Driver.cu:59-97): cv from bv, boundary
 * conditions, inter-block exchange, physical properties. */
int cfx_init_conservative(cfx_ctx *ctx);

/* store_last_step (TimeAdvanceFunc.cu:13-24) for every block. */
int cfx_store_last_step(cfx_ctx *ctx);
/* compute_inviscid_flux + compute_viscous_flux (TimeAdvanceFunc.cu:38-243)
 * for one block: dq = -(dF_inv) + (dF_visc), overwritten (set_dq_to_0 is
 * folded in).  blk < 0 means every block. */
int cfx_rhs(cfx_ctx *ctx, int blk);
/* local_time_step (TimeAdvanceFunc.cu:245-253, TemporalScheme.cu:10-43) for
 * every block; global != 0 also reduces min(dt_local) over all blocks into the
 * device-resident dt used by cfx_update.  dt_out (may be NULL) receives it. */
int cfx_time_step(cfx_ctx *ctx, int global, double *dt_out);
/* update_cv_and_bv (TimeAdvanceFunc.cu:255-290) for every block.  With
 * rk_stages == 1 the local dt is used (reference); with rk_stages == 3,
 * stage = 0,1,2 applies the SSP-RK3 combination with the device dt; with
 * lusgs = 1, dq holds the increment of cfx_lusgs and cv += dq. */
int cfx_update(cfx_ctx *ctx, int stage);
/* The implicit treatment the reference leaves empty ("implicit treatment if
 * needed", Driver.cu:156; LUSGS, TemporalScheme.cuh:19-21): replaces dq (the
 * RHS) of every block by the LU-SGS increment dQ of
 *   (D + L) D^-1 (D + U) dQ = dq,  D = J/dt_local + sum_d (omega r_d + 2 v_d),
 * L dQ = -sum_d [ (dG_d + omega r_d dQ)/2 + v_d dQ ] at the lower neighbour,
 * U dQ =  sum_d [ (dG_d - omega r_d dQ)/2 - v_d dQ ] at the upper neighbour,
 * with r_d = |V.S_d| + c|S_d| and v_d = max(gamma, 4/3) mu/(rho Pr) J |grad xi_d|^2
 * at that neighbour (S_d = J grad xi_d), dG_d = G_d(Q + dQ) - G_d(Q) the
 * change of its inviscid flux through S_d (pressure change at frozen gamma),
 * and dQ = 0 outside the block (Yoon & Jameson, AIAA J. 26 (1988) 1025).
 * Needs cfx_time_step's local dt.  Requires lusgs = 1 and FP64. */
int cfx_lusgs(cfx_ctx *ctx);
/* DBoundCond::apply_boundary_conditions (BoundCond.cu:251-313) + slip wall. */
int cfx_apply_bc(cfx_ctx *ctx);
/* Driver::data_communication (Driver.cu:194-217), inner faces on device. */
int cfx_exchange(cfx_ctx *ctx);
/* eliminate_k_gradient (Field.cu:268-292), 2-D only. */
int cfx_eliminate_k_gradient(cfx_ctx *ctx);
/* update_physical_properties (Field.cu:178-210) over the padded range. */
int cfx_update_properties(cfx_ctx *ctx);
/* compute_residual (Driver.cu:219-277) without the normalisation:
 * res[0..3] = sqrt(sum over all blocks of (d rho, d|V|, d p, d T)^2 / n_grid_total). */
int cfx_residual(cfx_ctx *ctx, double res[4]);
/* The loop body (Driver.cu:135-190) of Driver::steady_simulation
 * (Driver.cu:113-192), n_steps times, without residual output. */
int cfx_step(cfx_ctx *ctx, int n_steps);
/* Read back the device dt of the last global reduction. */
int cfx_get_dt(cfx_ctx *ctx, double *dt);
/* Set the device dt (fixed-dt runs). */
int cfx_set_dt(cfx_ctx *ctx, double dt);
/* Health check: synchronises and reports non-finite state (CFX_E_NONFINITE)
 * and Newton failures (CFX_E_NEWTON, count in *n_newton_fail if non-NULL). */
int cfx_check(cfx_ctx *ctx, long long *n_newton_fail);
/* Synchronise the context's stream. */
int cfx_sync(cfx_ctx *ctx);

/* Time k steps with HIP events on the context's stream (no host sync inside).
 * ms_total receives the elapsed time.  Used by the benchmark. */
int cfx_time_steps(cfx_ctx *ctx, int n_steps, double *ms_total);

/* Per-kernel timing of the last cfx_time_steps call: for kernel class c
 * (0 rhs, 1 update, 2 ghost fill = boundary conditions and inner-face
 * exchange, 3 props, 4 dt, 5 other) the summed milliseconds and launch
 * count, measured with events on the context's stream (eager launches only:
 * a context with profiling on does not replay the step graph). */
int cfx_kernel_times(cfx_ctx *ctx, int cls, double *ms, long long *launches);
int cfx_set_profiling(cfx_ctx *ctx, int on);

/* Execution of the step (no counterpart in the reference, whose Driver
 * launches every kernel block by block, Driver.cu:135-190).
 *   batched = 0: one launch per block and kernel, one launch per boundary
 *                face and per inner face, in the reference's order;
 *   batched = 1 (default): one launch per kernel over all blocks, and the
 *                boundary/exchange face operations grouped into levels of
 *                operations that touch disjoint points (order kept between
 *                any two that overlap), one launch per level;
 *   graph = 1:   cfx_step replays one captured step graph (batched or not).
 * Results are identical in every mode. */
int cfx_set_execution(cfx_ctx *ctx, int batched, int graph);
/* Ghost-fill operations (boundary faces + inner faces) of one step stage and
 * the number of levels (launches) they take when batched.  An outflow face
 * that covers a whole block face and is not touched by an earlier operation
 * is applied by the update kernel to its interior rows (same values, same
 * order), so only its edge rows count here; cfx_apply_bc and
 * cfx_init_conservative still run every face in full. */
int cfx_schedule_info(cfx_ctx *ctx, int *n_ops, int *n_levels);

/* RHS tile selection and autotuner (no counterpart in the reference, whose
 * inviscid kernels use fixed 64-point lines, TimeAdvanceFunc.cu:38-154).
 * The RHS runs one kernel per face direction; each direction has
 * cfx_rhs_variants() tile shapes, and a persistence factor p (0: one
 * workgroup per tile; p > 0: p x resident-workgroups-per-CU x CUs
 * workgroups looping over the tiles). */
int cfx_rhs_variants(cfx_ctx *ctx);
/* Same tile variant in every direction, persistence 0. */
int cfx_set_rhs_variant(cfx_ctx *ctx, int variant);
int cfx_set_rhs_tiles(cfx_ctx *ctx, const int variant[3], const int persist[3]);
int cfx_get_rhs_tiles(cfx_ctx *ctx, int variant[3], int persist[3]);
/* Times every (variant, persistence) of every direction on the context's
 * blocks, reps launches each, and keeps the fastest; best_ms[3] (may be
 * NULL) receives the per-direction time of the choice.  Leaves dq undefined. */
int cfx_autotune(cfx_ctx *ctx, int reps, double *best_ms);

/* ------------------------------------------------ transport surrogate ---
 * No counterpart in the reference (its transport is Transport.cu:43-96,
 * which generates the training samples and stays the fallback).  A 3 x 64
 * SiLU MLP maps (T, log p, Y_1..Y_ns) to log(mu, kappa, rho*D_1..ns), BF16
 * weights on MFMA with FP32 accumulation, trained on the GPU.  Runs on the
 * mixed-precision path with 1..14 species, selected by
 * cfx_params.transport = CFX_TRANSPORT_MLP or cfx_set_transport.  Points
 * outside the training envelope (T / p box, Mahalanobis distance in the
 * principal axes of the training inputs) take the Wilke routine; one
 * wavefront in check_every also evaluates Wilke and keeps it where the
 * surrogate is off by more than check_tol.  Until trained or loaded, a
 * context with the surrogate selected returns CFX_E_SURROGATE from the calls
 * that evaluate properties. */
typedef struct {
  long long n_samples;    /* training set including the held-out part (2,000,000) */
  double holdout;         /* held-out fraction for validation (0.02) */
  int epochs;             /* passes over the training part (60) */
  int batch;              /* samples per Adam step, multiple of 32 (4096) */
  double lr, lr_final;    /* Adam step size, cosine-decayed (4e-3 -> 1e-5) */
  double t_min, t_max;    /* sampled temperature range [K] (250, 3000) */
  double p_min, p_max;    /* sampled pressure range [Pa], log-uniform (0.5, 2 atm) */
  double dirichlet_alpha; /* mass fractions ~ Dirichlet(alpha) (0.5) */
  double check_tol;       /* relative bound of the sampled Wilke check (0.05) */
  int check_every;        /* one wavefront in check_every is checked (64) */
  unsigned long long seed;
  double envelope_quantile; /* the envelope admits inputs up to this quantile of the
                               training inputs' Mahalanobis distance (0.99; 1: all) */
} cfx_mlp_config;

typedef struct {
  double samples_ms, train_ms; /* GPU time: sample generation, training */
  int epochs, steps;
  double loss_first, loss_last; /* mean squared error of the normalised log outputs */
  double val_rmse;              /* the same on the held-out samples */
  double val_max_rel[3];        /* held-out max relative error: mu, kappa, max_l rho*D_l */
  double val_p99_rel[3];        /* 99th percentile of the same */
  long long n_train, n_val;
  int n_params;
  double val_answered; /* fraction of the held-out samples the surrogate answers (inside the
                          envelope, below the near-pure limit); the val_ statistics are over these */
} cfx_mlp_report;

void cfx_mlp_default_config(cfx_mlp_config *cfg);
/* Generates the training set on the GPU with the FP64 transport routine,
 * trains, validates, and makes the surrogate available (cfg NULL: defaults). */
int cfx_mlp_train(cfx_ctx *ctx, const cfx_mlp_config *cfg, cfx_mlp_report *report);
/* Number of FP32 parameters: W1[64][ns+2] b1[64] W2[64][64] b2 W3[64][64] b3
 * W4[ns+2][64] b4[ns+2], in that order. */
int cfx_mlp_n_params(cfx_ctx *ctx);
/* norm: 4 x 16 floats: input mean, input 1/std, output mean, output std
 * (inputs T, log p, Y_l; outputs log mu, log kappa, log rho*D_l).  Loaded
 * parameters get an envelope that admits every input. */
int cfx_mlp_set_params(cfx_ctx *ctx, const float *params, const float *norm);
int cfx_mlp_get_params(cfx_ctx *ctx, float *params, float *norm);
int cfx_mlp_set_check(cfx_ctx *ctx, double check_tol, int check_every);
/* Gradient of loss = sum (y - t)^2 / (n (ns+2)) over n (a multiple of 32)
 * normalised samples x[n][16], t[n][16] (zero padded) with the GPU training
 * kernels; grad receives cfx_mlp_n_params floats. */
int cfx_mlp_gradient(cfx_ctx *ctx, const float *x, const float *t, int n, float *grad, double *loss);
/* Normalised outputs y[n][16] of normalised inputs x[n][16]. */
int cfx_mlp_predict(cfx_ctx *ctx, const float *x, int n, float *y);
/* cfx_mlp_predict without device allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (serving loops, code running
 * while a graph is captured on the device).  Stages x and y through a fixed
 * device array of 2 x 65536 x 16 floats (8 MiB per device, held until the
 * process ends) instead of buffers allocated per call; n > 65536 returns
 * CFX_E_ARG.  Results equal cfx_mlp_predict's.  Every call on one device
 * shares that array: calls that overlap on one device race. */
int cfx_mlp_predict_noalloc(cfx_ctx *ctx, const float *x, int n, float *y);
int cfx_set_transport(cfx_ctx *ctx, int transport);
/* Counters since training/loading: [0] points outside the envelope,
 * [1] points checked against Wilke, [2] checks failed. */
int cfx_mlp_stats(cfx_ctx *ctx, long long counts[3]);

#ifdef __cplusplus
}
#endif

#endif /* CFX_H */
