/* collop.h -- C-ABI of the MI355X (gfx950) Landau collision library.
 *
 * The library replaces the CUDA.jl hot loops of junyixu/CollisionOperators.jl
 * behind the reference's own swappable hooks (solver.jl:30-34).  Every entry
 * point takes plain pointers and sizes; no Julia, torch or HIP types cross
 * this boundary.
 *
 * Conventions
 *   - Marker arrays are Julia's column-major N x 2 Float64 matrices:
 *     v[0..N) = v1, v[N..2N) = v2 (the same for G, dot_v; A is N x 3).
 *   - Host pointers belong to the caller and are not retained after a call.
 *     The library owns the device buffers cached in the context.
 *   - Calls are synchronous.  One context per caller thread; a context is not
 *     thread-safe.  Several contexts may share one device.
 *   - Every function returns a collop_status; on failure the message is in
 *     collop_last_error(ctx).  The reference raises a Julia error in the same
 *     places (e.g. P_DEG != 2 at projection_gpu.jl:189).
 */
#ifndef COLLOP_H
#define COLLOP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct collop_ctx collop_ctx;

typedef enum {
  COLLOP_OK = 0,
  COLLOP_ERR_INVALID = 1,      /* bad argument or size                        */
  COLLOP_ERR_UNSUPPORTED = 2,  /* e.g. p_deg != 2 (projection_gpu.jl:189)      */
  COLLOP_ERR_HIP = 3,          /* HIP runtime error (message has the detail)  */
  COLLOP_ERR_NO_MESH = 4,      /* collop_set_mesh not called                  */
  COLLOP_ERR_NONFINITE = 5,    /* NaN/Inf reached the residual of a map       */
  COLLOP_ERR_NO_STATE = 6,     /* device-resident call before collop_state_set */
  COLLOP_ERR_NO_MODEL = 7      /* warm start requested without loaded weights  */
} collop_status;

/* ---- context ------------------------------------------------------------ */
int collop_create(int device, collop_ctx** out);
int collop_destroy(collop_ctx* ctx);
const char* collop_last_error(const collop_ctx* ctx);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int collop_version(void);

/* ---- mesh (the Workspace fields the reference's GPU code reads) ----------
 * Replaces the cached mesh tables of projection_gpu.jl:186-211 (bp1, bp2,
 * ext1d_1/2, basis_start_1d_1/2, n_dofs_1d_1) and the mass-matrix LU of
 * MantisWrappers.jl:107-120, which the library rebuilds as the Kronecker
 * product M1 (x) M2 of banded 1D factors.
 *   bp1[nb1], bp2[nb2]        breakpoints (strictly increasing)
 *   C1[9*(nb1-1)], C2[...]    Bezier extraction, C[i + 3*j + 9*cell] with
 *                             phi_j = sum_i C[i,j] B_i  (MantisWrappers.jl:237)
 *   bs1[nb1-1], bs2[nb2-1]    1-based first basis index per cell
 *   nd1, nd2                  1D basis counts (n_dofs = nd1*nd2)
 *   p_deg                     must be 2 (the kernels are P_DEG=2 specialised)
 *   n_quad                    Gauss-Legendre points per direction for the
 *                             entropy / r quadrature (N_QUAD, Parameters.jl:27) */
int collop_set_mesh(collop_ctx* ctx, const double* bp1, int nb1, const double* bp2, int nb2, const double* C1,
                    const double* C2, const int32_t* bs1, const int32_t* bs2, int nd1, int nd2, int p_deg,
                    int n_quad);

/* ---- hook-level entry points (host arrays in / out) --------------------- */

/* COLLISION_FN (solver.jl:30): the O(N^2) Landau pair sum of functions.jl:205-237
 * on the GPU.  Replaces compute_collision_gpu! (collision_gpu.jl:258-280). */
int collop_collision_f64(collop_ctx* ctx, const double* v, const double* w, const double* G, double* dot_v,
                         int64_t N);

/* COLLISION_FN with gpu_fp32=true: pair math in FP32 from FP32 copies of
 * (v, G, w), per-marker sums with two-sum compensated accumulation, result
 * returned in FP64.  Replaces compute_collision_gpu32! (collision_gpu.jl:233-256). */
int collop_collision_f32c(collop_ctx* ctx, const double* v, const double* w, const double* G, double* dot_v,
                          int64_t N);

/* COLLMETRIC_FN (solver.jl:34): pair sum plus the metric A = sum w U(d)
 * (A11, A12, A22), functions.jl:249-290.  Replaces
 * compute_collision_metric_gpu! (collision_gpu.jl:377-402). */
int collop_collision_metric_f64(collop_ctx* ctx, const double* v, const double* w, const double* G, double* dot_v,
                                double* A, int64_t N);
/* The same from FP32 pair terms with compensated sums (no counterpart in the
 * reference, whose metric kernel is FP64 only): the outer maps of the
 * FP32-compensated defect-correction solve. */
int collop_collision_metric_f32c(collop_ctx* ctx, const double* v, const double* w, const double* G, double* dot_v,
                                 double* A, int64_t N);

/* Allocation-free variant of collop_collision_f64 (no counterpart in the
 * reference; for callers that cannot allocate: code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  v, w, G and dot_v are
 * DEVICE arrays on the context's GPU, in the same layouts; the scratch (packed
 * sources and slice partials, 21 MB) is a static device array of the library,
 * so N is at most 65536 (COLLOP_ERR_INVALID beyond).  The work is launched on
 * `stream` (a hipStream_t; NULL = the context's stream) without
 * synchronisation.  Every call on one GPU shares that scratch: calls that
 * overlap on one GPU race. */
int collop_collision_f64_noalloc(collop_ctx* ctx, const double* v, const double* w, const double* G, double* dot_v,
                                 int64_t N, void* stream);

/* L2PROJ_FN, scatter only: rhs_k = sum_alpha w_alpha phi_k(v_alpha) over
 * in-domain markers (functions.jl:121-131; _l2_scatter_kernel!,
 * projection_gpu.jl:47-71).  A caller wanting bitwise host-solve parity
 * applies its own M_lu \ rhs. */
int collop_l2_rhs(collop_ctx* ctx, const double* v, const double* w, int64_t N, double* rhs);

/* L2PROJ_FN: scatter plus the mass solve on the device, f = M^{-1} rhs.
 * Replaces l2_project_gpu! (projection_gpu.jl:213-226). */
int collop_l2_project(collop_ctx* ctx, const double* v, const double* w, int64_t N, double* f_coeffs);

/* COMPG_FN (solver.jl:32): G_alpha = grad(sum_k L_k phi_k)(v_alpha), zero out of
 * domain (functions.jl:182-202).  Replaces compute_G_gpu! (projection_gpu.jl:228-246). */
int collop_grad_gather(collop_ctx* ctx, const double* v, int64_t N, const double* L, double* G);

/* LOGGRAD_FN (solver.jl:33): g = clamp(grad f / max(|f|, fs_floor), +-g_max),
 * zero out of domain (functions.jl:37-64).  Replaces eval_loggrad_gpu!
 * (projection_gpu.jl:251-269); the reference passes fs_floor=1e-30, g_max=100. */
int collop_loggrad_gather(collop_ctx* ctx, const double* v, int64_t N, const double* f_coeffs, double fs_floor,
                          double g_max, double* g);

/* compute_r! and compute_entropy (functions.jl:137-179) on the device: the
 * N_QUAD^2-point element quadrature the reference runs on the CPU for every
 * map.  Either output pointer may be NULL. */
int collop_quadrature(collop_ctx* ctx, const double* f_coeffs, int use_logsq, double* r, double* entropy);

/* x = M^{-1} b (the reference's ws.M_lu \ b, functions.jl:132, solver.jl:67). */
int collop_mass_solve(collop_ctx* ctx, const double* b, double* x);

/* gmres! (newton_krylov.jl:60-131) of the device Newton-Krylov engine on a
 * dense operator: x = the GMRES solution of A x = b (A row-major n x n) from
 * x0 = 0 with basis size mem (<= 64), stopping at ||b - A x|| <= atol + rtol
 * ||b||, on breakdown or after itmax products.  products and rnorm (the GMRES
 * residual norm) may be NULL.  For small systems and for testing the engine. */
int collop_gmres_dense(collop_ctx* ctx, int n, const double* A, const double* b, double rtol, double atol, int itmax,
                       int mem, double* x, int* products, double* rnorm);

/* Lenard-Bernstein right-hand side at markers v with log-density gradient g
 * (N x 2, e.g. from collop_loggrad_gather): the moments n, U, Q of (v, w)
 * (compute_moments, functions.jl:66-82), the drift multipliers (A1, A2, B) of
 * the 3x3 system enforcing momentum and energy conservation
 * (compute_drift_multipliers, functions.jl:84-106) and dot_v = -nu (g + A +
 * B v) (compute_LB_velocity!, functions.jl:108-115).  mult[3] may be NULL. */
int collop_lb_rhs(collop_ctx* ctx, const double* v, const double* w, const double* g, int64_t N, double nu,
                  double* dot_v, double* mult);

/* Run-record diagnostics, the conservation CSV's neg_part and fp_minus_fs
 * columns (io.jl:106-107, written per step at main.jl:306-307):
 *   collop_negative_part_l1: int max(-f_s, 0) dv by the N_QUAD^2 Gauss rule per
 *     cell; replaces compute_negative_part_l1 (functions.jl:300-313).
 *   collop_fs_minus_fp_l2: ||f_s - f_p||_2 with f_p = m_e / |e| the cellwise
 *     histogram density of the markers (v, w) in the open box; replaces
 *     compute_fs_minus_fp_l2 (functions.jl:320-343). */
int collop_negative_part_l1(collop_ctx* ctx, const double* f_coeffs, double* neg_part);
int collop_fs_minus_fp_l2(collop_ctx* ctx, const double* v, const double* w, int64_t N, const double* f_coeffs,
                          double* fs_minus_fp);

/* ---- device-resident implicit step ---------------------------------------
 * The step of main.jl:226-307 (Landau branch, solver = :anderson) with the
 * Picard map of solver.jl:176-227 and the Anderson solve of solver.jl:300-423
 * running on the device; the host only reads a handful of reduced scalars per
 * Anderson iteration. */

typedef enum {
  COLLOP_PREC_F64 = 0,      /* every map in FP64 (the plain FP64 GPU path)     */
  COLLOP_PREC_F32C = 1      /* FP32-compensated pair sum for the early maps,
                               FP64 maps to finish; last map always FP64      */
} collop_precision;

typedef enum {
  COLLOP_WARM_EULER = 0,    /* explicit Euler predictor (main.jl:239-245)       */
  COLLOP_WARM_MLP = 1       /* Euler + per-marker MLP correction (main.jl:252-254) */
} collop_warmstart;

typedef enum {
  COLLOP_SOLVER_ANDERSON = 0, /* step_anderson! (solver.jl:300-423)              */
  COLLOP_SOLVER_DEFECT = 1,   /* frozen-metric defect correction, step_defect!
                                 (solver.jl:569-634): each outer iteration is one
                                 full map with the metric A; an inner Anderson
                                 solve on the frozen-metric map needs no pair sum */
  COLLOP_SOLVER_NEWTON = 2    /* Jacobian-free Newton-Krylov, step_newton!
                                 (solver.jl:437-486, newton_krylov.jl): GMRES on
                                 forward-difference products, one Picard map each;
                                 FP64 maps                                       */
} collop_solver;

typedef enum {
  COLLOP_MODEL_LANDAU = 0,  /* the Landau pair sum (collision_model = :landau)  */
  COLLOP_MODEL_LB = 1       /* Lenard-Bernstein drift, O(N) per map
                               (collision_model = :lb; functions.jl:66-115,
                               solver.jl:185-222): Anderson solver, FP64,
                               Euler predictor only                             */
} collop_model;

typedef struct {
  double dt;                /* DT                                               */
  int32_t m_anderson;       /* window m (<= 8)                                  */
  int32_t max_iter;
  double tol;               /* relative tol on ||r|| / ||v||                     */
  double abs_floor;
  int32_t stag_window;
  double stag_rel_tol;
  int32_t damp_decay_start;
  double damp_decay_factor;
  double restart_factor;    /* +inf = never (the reference default)             */
  double damping;           /* beta                                             */
  double reg_factor;        /* 1e-10 in the reference                           */
  int32_t use_anderson;
  int32_t use_gonzalez;
  int32_t exit_picard_step;
  int32_t use_logsq;
  int32_t precision;        /* collop_precision                                 */
  int32_t warmstart;        /* collop_warmstart                                 */
  double f32_floor_factor;  /* F32C: FP32 maps run until ||r|| < factor * 2^-24
                               * ||Gv - v0|| (0 = 16), then FP64 maps finish;
                               DEFECT also switches one outer map early when the
                               contraction so far predicts the floor next      */
  int32_t f32_keep_window;  /* F32C: 1 = keep the Anderson window across the
                               switch (the last FP32 iterate is re-evaluated in
                               FP64 first), 0 = restart it                      */
  int32_t record_training;  /* 1 = append (features, target) of every marker of
                               every step to the warm-start training dataset   */
  int32_t solver;           /* collop_solver                                    */
  double dc_eta;            /* DEFECT: inner tolerance relative to the outer
                               residual (Parameters.jl:89, 0.05)                */
  int32_t dc_max_inner;     /* DEFECT: inner maps per outer iteration (60)      */
  int32_t dc_stag_window;   /* DEFECT: outer iterations between stagnation
                               checks (5)                                       */
  int32_t dc_outer_m;       /* DEFECT extension: Anderson window (<= 8) on the
                               outer iteration v -> (inner solution); 0 = the
                               reference's plain outer iteration                */
  double dc_near_theta;     /* DEFECT extension: pairs with w_a / |d| > theta at
                               the step's first outer iterate are evaluated
                               exactly inside the frozen-metric map (0 = none);
                               the fixed point is unchanged                     */
  int32_t dc_metric_reuse;  /* DEFECT extension: which outer maps compute the
                               metric A (the others run the plain pair sum and
                               keep the last A): 0 = every one (the reference),
                               1 = all but the FP64 maps after an FP32c phase,
                               2 = the step's first only                        */
  int32_t collision_model;  /* collop_model (Parameters.jl:50)                  */
  double nu;                /* LB collision frequency (Parameters.jl:53, 1.0)   */
  double dc_eta_f64;        /* DEFECT extension: inner tolerance of the FP64
                               outer iterations (0 = dc_eta); a tighter one
                               trades inner maps for FP64 pair sums            */
  int32_t nk_krylov_max;    /* NEWTON: GMRES basis per Newton step (<= 64; 30)  */
  double nk_eta_max;        /* NEWTON: Eisenstat-Walker forcing cap (0.9)       */
  double nk_fd_h;           /* NEWTON: absolute finite-difference step of J u
                               (Parameters.jl:98; 0 = 1e-6)                     */
  double nk_fd_rel;         /* NEWTON: legacy relative step nk_fd_rel max(|v|, 1);
                               > 0 overrides nk_fd_h                            */
  double cl_theta;          /* ANDERSON extension (Landau): markers with
                               max(w_i, w_j) / |d| > cl_theta at the start of
                               the step form stiff clusters whose rows every
                               FP64 map solves exactly (nonlinear elimination);
                               the exit test stays on the plain map, so the
                               root is the same; 0 = off                       */
  int32_t cl_kmax;          /* largest stiff cluster (markers, <= 32)          */
  int32_t dc_fallback;      /* DEFECT extension: 1 = hand the step over to the
                               Anderson solve (step_anderson!, FP64 maps) with
                               the iterations left when (a) an inner solve ends
                               unconverged at dc_max_inner maps or (b) an FP32c
                               outer iteration, above 4x the FP32 floor, does
                               not reduce the residual by 0.4 (both from the
                               outer iterate), (c) an FP64 outer iteration does
                               not reduce it by 0.4 or (d) the outer stagnation
                               test fires (both from the best outer iterate's
                               map); 0 = off (the reference's step_defect!)    */
  int32_t dc_sticky;        /* DEFECT with dc_fallback, across the steps of a
                               state: 1 = after a step that handed over within
                               its first 2 outer iterations, the next steps run
                               the Anderson solve from the start (FP64 maps and
                               predictor, no stiff clusters) until one of them
                               converges; 0 = every step starts with the defect
                               correction.  No counterpart in the reference.   */
} collop_step_params;

typedef struct {
  int32_t iters;            /* Picard-map evaluations (the reference's `iter`)   */
  int32_t iters_f32;        /* of which FP32-compensated maps                    */
  int32_t n_restart;
  int32_t exit_reason;      /* 0 converged, 1 stagnation, 2 max_iter            */
  double residual;          /* ||r|| at exit (the reference's `res_final`)       */
  double r0;                /* ||r|| of the first map (`r0_init`)               */
  double entropy;           /* S_h of the new state                            */
  double energy;            /* E = 1/2 sum w |v|^2                              */
  double momentum[2];       /* P = sum w v                                     */
  double seconds;           /* wall time of the step                           */
  int32_t inner_maps;       /* DEFECT: frozen-metric maps (no pair sum)         */
  int32_t n_clusters;       /* ANDERSON: stiff clusters solved in each FP64 map */
  int32_t cl_fail;          /* cluster solves that reached no root (those maps
                               kept the plain rows of the cluster)             */
  int32_t stat_syncs;       /* maps whose statistics needed a stream
                               synchronisation (the 20 ms poll ran out)        */
  int32_t handover;         /* DEFECT with dc_fallback: outer maps before the
                               step continued with the Anderson solve (0 = the
                               defect correction finished the step)           */
  int32_t dc_skipped;       /* DEFECT with dc_sticky: 1 = the step ran the
                               Anderson solve from the start                   */
} collop_step_stats;

/* Upload markers (v: N x 2 column-major, w: N) and project them; keeps S0. */
int collop_state_set(collop_ctx* ctx, const double* v, const double* w, int64_t N);
/* Copy the current markers back (N x 2). */
int collop_state_get(collop_ctx* ctx, double* v);
/* Current projection coefficients f_s (n_dofs). */
int collop_state_coeffs(collop_ctx* ctx, double* f_coeffs);
/* Diagnostics of the resident state and its projection (the same two as
 * collop_fs_minus_fp_l2 and collop_negative_part_l1, main.jl:306-307):
 * out[0] = fp_minus_fs, out[1] = neg_part. */
int collop_state_diagnostics(collop_ctx* ctx, double* out);
/* Advance n_steps implicit steps on the device; stats[n_steps] optional. */
int collop_run_steps(collop_ctx* ctx, const collop_step_params* params, int n_steps, collop_step_stats* stats);
/* One step on host arrays: state_set + run_steps(1) + state_get. */
int collop_step(collop_ctx* ctx, const collop_step_params* params, double* v_inout, const double* w, int64_t N,
                collop_step_stats* stats);

/* ---- learned warm start ----------------------------------------------------
 * The reference calls nn_warmstart_correct!(v1, model, ...) after its Euler
 * predictor (main.jl:252-254, hook site only: warmstart_nn.jl is absent).  Here
 * the model is a per-marker MLP 6 -> 64 -> 64 -> 2 (GELU) run on the matrix
 * cores (BF16 operands, FP32 accumulation).  Features of marker a at the start
 * of a step: v_a (2), G_a = -(dS/dv_a)/w_a (2), f_s(v_a), |grad f_s(v_a)|.  The
 * output y corrects the predictor: v1 = v + dt (vdot_euler + y * out_scale),
 * i.e. it predicts the step's midpoint velocity.  Training targets are
 * (v_{n+1} - v_n)/dt - vdot_euler from the solver's own converged steps.
 * Parameters: float32[4738] = W1[64][6] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]. */
int collop_mlp_set_params(collop_ctx* ctx, const float* params, int n);
int collop_mlp_get_params(collop_ctx* ctx, float* params, int n);
/* features are normalised as (x - mu) / sd; targets scaled by 1 / out_scale */
int collop_mlp_set_norm(collop_ctx* ctx, const float* mu, const float* sd, const float* out_scale);
/* (Re)allocate and clear the device-resident training dataset (samples). */
int collop_mlp_dataset_alloc(collop_ctx* ctx, int64_t capacity);
int collop_mlp_dataset_size(collop_ctx* ctx, int64_t* n);
/* Copy samples [first, first + count) as float32[count][8] = (x0..x5, t0, t1). */
int collop_mlp_dataset_read(collop_ctx* ctx, int64_t first, int64_t count, float* out);
/* Adam (beta 0.9 / 0.999, eps 1e-8) on the batch-mean squared error over
 * shuffled 16-sample tiles; loss_per_epoch[epochs] optional. */
int collop_mlp_train(collop_ctx* ctx, int epochs, int batch, double lr, uint64_t seed, double* loss_per_epoch);
/* Run the network on host records float32[n][8]: mode 0 -> y[n][2]; mode 1 ->
 * mean loss; mode 2 -> mean loss and its gradient grad[4738]. */
int collop_mlp_eval(collop_ctx* ctx, const float* rec, int64_t n, int mode, float* y, double* loss, float* grad);

/* ---- batched ensemble ----------------------------------------------------
 * Independent runs with uneven N on the context's mesh, advanced together (the
 * reference advances one run per process: main.jl:226-307).  Each Anderson
 * iteration of the whole ensemble is one replay of a captured hipGraph whose
 * pair sum is ONE persistent launch over the tiles of every active run.  Each
 * run follows exactly the solve of collop_run_steps (own Anderson window,
 * precision switch and exit).  The pair-sum form is the context's at the time
 * of this call: symmetric with the context's tiles (4 when automatic) if every
 * run fits in 512 partial slots, else ordered.  A run's markers then equal,
 * bit for bit, those of the same run stepped alone on a context with
 * collop_set_pair_mode(ctx, COLLOP_PAIR_SYMMETRIC, tiles) (symmetric), or with
 * collop_set_pair_tile(ctx, 1, collop_ensemble_slices(N), 1) (ordered).
 * v: the runs' N_r x 2 column-major blocks one after another (2 sum N_r
 * doubles); w: their weights (sum N_r). */
int collop_ensemble_set(collop_ctx* ctx, int n_runs, const int64_t* n, const double* v, const double* w);
/* current markers, same layout as collop_ensemble_set's v */
int collop_ensemble_get(collop_ctx* ctx, double* v);
/* projection coefficients of every run, [n_runs][n_dofs] */
int collop_ensemble_coeffs(collop_ctx* ctx, double* f_coeffs);
/* n_steps implicit steps of every run (Euler predictor); stats[step * n_runs + run]
 * optional, `seconds` = wall time of the ensemble step. */
int collop_ensemble_step(collop_ctx* ctx, const collop_step_params* params, int n_steps, collop_step_stats* stats);
/* pair-sum source slices the ensemble gives a run of n markers */
int collop_ensemble_slices(int64_t n);
/* launch shape of the persistent pair sum: gamma_tile targets per workgroup
 * (64, 128, 256, 512) and waves_per_cu resident waves per CU (sets the grid);
 * use_graph = 0 launches the iteration's kernels one by one (same results). */
int collop_ensemble_set_tile(collop_ctx* ctx, int gamma_tile, int waves_per_cu, int use_graph);
/* autotuner: times the persistent pair sum (one FP32c and one FP64 launch with
 * every run active) for gamma_tile in {64, 128, 256, 512} x waves_per_cu in
 * {4, 8, 16, 32} on the current ensemble, keeps the fastest; ms[16] optional
 * receives each candidate's time in that order.  In the symmetric form the
 * gamma tile is the super-row of 128 R markers, R in {2, 4, 8} waves per block
 * (ms rows 0..2; *gamma_tile = 128 R): each run is then bitwise the same run
 * stepped alone after collop_set_pair_mode(ctx, COLLOP_PAIR_SYMMETRIC, R).
 * collop_ensemble_set_tile ignores gamma_tile in the symmetric form. */
int collop_ensemble_tune(collop_ctx* ctx, int* gamma_tile, int* waves_per_cu, double* ms);

/* ---- tuning and measurement (no counterpart in the reference) -----------
 * Ordered pair-sum tile (and selects the ordered sum): gamma_per_lane in {1, 2}
 * targets per lane, slices in [1, 64] source slices (0 = automatic), newton in
 * {1, 2} FP64 rsqrt refinements. */
int collop_set_pair_tile(collop_ctx* ctx, int gamma_per_lane, int slices, int newton);

typedef enum {
  COLLOP_PAIR_ORDERED = 0,   /* every ordered pair evaluated, as collision_gpu.jl:49-107 */
  COLLOP_PAIR_SYMMETRIC = 1  /* every unordered pair evaluated once and applied to
                                both markers (the default)                           */
} collop_pair_mode;
/* Pair-sum form for every later pair sum of the context (hooks, steps);
 * tiles = waves per block of the symmetric sum, 0 (automatic), 2, 4 or 8.
 * Both forms sum the same terms in different orders: results agree to
 * rounding.  The symmetric form needs ceil(N / (128 tiles)) <= 512 partial
 * slots and falls back to the ordered sum beyond that.  No counterpart in the
 * reference. */
int collop_set_pair_mode(collop_ctx* ctx, int mode, int tiles);
/* Fused (on = 1, the default) or launch-per-stage (0) small kernels in the
 * frozen-metric inner maps of defect correction and in the Lenard-Bernstein
 * Anderson maps.  Both compute the same bits.  The fused maps need at most 256
 * basis functions per direction (larger meshes always run the staged maps).
 * No counterpart in the reference. */
int collop_set_fused_maps(collop_ctx* ctx, int on);
/* Control of the Anderson solve of Landau steps without stiff clusters: on the
 * device (on = 1, the default: a control kernel per map runs the window's
 * least squares, the exit and stagnation tests and plans the next map, and the
 * host launches the next map before the current one has finished once the maps
 * are FP64) or on the host (0: the host reads every map's statistics and
 * decides).  Both compute the same bits.  The host-side counterpart is
 * step_anderson! (solver.jl:300-423), whose least squares is solver.jl:396-412. */
int collop_set_device_solve(collop_ctx* ctx, int on);
/* The equal-weight form of the symmetric pair sums (on = 1, the default): when
 * every marker carries the same weight w = 2^k, 2^-60 <= w <= 2^60 (the state's
 * markers in the step, the call's markers in the collision hooks), w scales G
 * exactly and the weight products per pair drop out (FP64: the plain sum;
 * FP32-compensated: the plain sum and the metric sum without the near-pair
 * search); off (0) runs the general form.  Both compute the same bits.  No
 * counterpart in the reference (its weights enter every term of
 * functions.jl:230-231). */
int collop_set_weight_path(collop_ctx* ctx, int on);
/* The form (mode), waves per block (tiles, 0 for the ordered sum) and partial
 * slots a pair sum over N markers runs with under the current settings; each
 * output pointer may be NULL. */
int collop_get_pair_mode(collop_ctx* ctx, int64_t N, int* mode, int* tiles, int* slots);
/* HIP-event timing of every pair-kernel launch on the context's stream.
 * Returns the totals accumulated since the previous call (ms_total[4],
 * launches[4]; index 0 = FP64, 1 = FP32c, 2 = FP64 with the metric A,
 * 3 = FP32c with the metric), then resets them and sets the enable flag. */
int collop_profile(collop_ctx* ctx, int enable, double* ms_total, long* launches);

#ifdef __cplusplus
}
#endif

#endif /* COLLOP_H */
