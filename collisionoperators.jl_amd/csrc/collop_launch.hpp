// collop_launch.hpp -- host-side launchers of the gfx950 kernels.  Each .hip
// translation unit owns its kernels; the C-ABI layer calls them through these.
#pragma once
#include "collop_dev.hpp"

namespace collop {

// Block counts of the grid-stride scatter and gather launches for N markers.
__host__ __device__ inline int scatter_blocks(int N) {
  const int b = (N + kScatterThreads * 4 - 1) / (kScatterThreads * 4);
  return b < 1 ? 1 : (b > 128 ? 128 : b);
}
__host__ __device__ inline int gather_blocks(int N) {
  const int b = (N + 256 * 2 - 1) / (256 * 2);
  return b < 1 ? 1 : (b > 512 ? 512 : b);
}

// ---- batched ensemble: run tables ---------------------------------------------
// Runs are concatenated: run r holds markers [off, off + n) of arrays whose
// component stride is NT (off is a multiple of 256).  Each batched kernel gives
// run r the same number of blocks, walking the markers in the same order, as
// the single-run kernel at N = n, so per-run results match it bit for bit.
struct EnsRun {
  int n, off;
  int S;             // pair slices
  int sblk0, nsb;    // scatter blocks: first row, count
  int gblk0, ngb;    // gather blocks
  int pblk0, npb;    // blocks of 256 markers
  int eblk0, neb;    // blocks of 256 over the flattened 2n components
  long long poff;    // pair partials [S][2n] at part + poff
  double scale, inv_scale;  // scatter fixed-point scale 2^62 / sum|w| and its inverse
};

// Per-run control of one Anderson iteration (uploaded before each replay).
struct RunCtl {
  int active;       // the run maps in this iteration
  int kind;         // 0 FP64 pair sum, 1 FP32c
  int col;          // Anderson column written by the map (-1: none)
  int hist;         // window columns dotted with the new column / residual
  int cur;          // ping-pong buffer the map writes (r, Gv); 1 - cur holds (rp, Gp)
  int upd;          // before the map: 0 none, 1 v1 <- beta (Gv - dG gamma) + (1 - beta) v1,
                    // 2 v1 <- Gv, 3 v1 <- best   (Gv of the previous map: buffer 1 - cur)
  int upd_hist;     // columns of the upd = 1 step
  int copy_best;    // before the update: best <- Gv (exit_picard) or v1
  int exit_picard;
  int pad;
  double beta;
  double gamma[kMaxWindow];
};

struct EnsView {
  int R, NT;
  int n_sblk, n_gblk, n_pblk, n_eblk;
  const EnsRun* runs;
  const RunCtl* ctl;  // nullptr: every run takes part (predictor, projection)
  const int2* sblk;   // {run, local block} per block of each family
  const int2* gblk;
  const int2* pblk;
  const int2* eblk;
};

// ---- pair sum (pair.hip) ----------------------------------------------------
// Near-pair detection during a metric pass: every ordered in-domain pair with
// w_a / |d| > theta (th2 = theta^2) is appended to list[] as (gamma, alpha);
// count receives the number found (entries beyond cap are dropped, the count
// still grows).  qmax = max(w)^2 / theta^2: no pair with |d|^2 >= qmax is near
// in either direction, so the kernels test that first.
struct NearOut {
  double th2;
  int* count;
  int2* list;
  int cap;
  double qmax;
  // device flag of the device-resident solve: 0 = the solve has ended, the
  // launch returns at once (nullptr: always run)
  const int* live = nullptr;
};
constexpr int kNearCap = 1 << 16;  // near pairs kept per step (beyond it the near field is skipped)
struct PairTile {
  int gamma_per_lane;  // 1 or 2 targets per lane
  int slices;          // alpha slices (partials per target)
  int newton;          // FP64 rsqrt refinement steps (1 or 2)
  int sym;             // 0: ordered sum; R in {2, 4, 8}: symmetric sum, R waves per block (slices = sym_rows)
};

// ---- symmetric pair sum (pair_sym.hip) -----------------------------------------
// Each unordered pair once; super-rows of 128 R markers, one block of R waves
// per pair of super-rows.  Writes part[S][C][N] with S = sym_rows(N, R) slots
// (C = 2, or 5 with the metric), every slot of every marker exactly once.
__host__ __device__ inline int sym_rows(int N, int R) { return (N + 128 * R - 1) / (128 * R); }
constexpr int kSymMaxSlots = 512;    // larger N falls back to the ordered sum
constexpr int kSymPreferSlots = 256;  // the automatic choice takes 4 waves per block up to this many slots
// pow2w: every marker has the same weight, a power of two (FP64: the plain sum
// only; FP32c: the plain and the metric sums without the near-pair search; the
// same bits as without it, with fewer multiplies per pair)
void launch_pair_sym_f64(const Src64* src, int N, int R, bool metric, const NearOut* near, double* part,
                         hipStream_t st, const int* live = nullptr, bool pow2w = false);
void launch_pair_sym_f32c(const Src32* src, int N, int R, bool metric, const NearOut* near, double* part,
                          hipStream_t st, const int* live = nullptr, bool pow2w = false);

// Pack (v, G, w) into Src64/Src32 records.  Markers outside the open box get
// w_masked = 0 and in_domain = 0 (functions.jl:212-213, :219-220).
void launch_pack_src(const double* v, const double* G, const double* w, int N, double lo1, double hi1, double lo2,
                     double hi2, Src64* s64, Src32* s32, hipStream_t st);
// part[S][2N] <- slice sums of the FP64 pair sum
void launch_pair_f64(const Src64* src, int N, const PairTile& t, double* part, hipStream_t st,
                     const int* live = nullptr, bool pow2w = false);
// part[S][2N] <- slice sums of the FP32 pair sum (compensated, written in FP64)
void launch_pair_f32c(const Src32* src, int N, const PairTile& t, double* part, hipStream_t st,
                      const int* live = nullptr, bool pow2w = false);
// part[S][5N] <- slice sums of F (2 comps) and A (3 comps); near != null also
// collects the near pairs (one target per lane, one Newton step)
void launch_pair_metric_f64(const Src64* src, int N, const PairTile& t, double* part, hipStream_t st,
                            const NearOut* near = nullptr);
// the same from FP32 pair terms (compensated sums, written in FP64)
void launch_pair_metric_f32c(const Src32* src, int N, const PairTile& t, double* part, hipStream_t st,
                             const NearOut* near = nullptr, bool pow2w = false);
// out[ncomp*N] <- sum over S slices in fixed order
void launch_slice_sum(const double* part, int S, int64_t len, double* out, hipStream_t st);

// ---- FEM side (fem.hip) ------------------------------------------------------
// Scatter: fields = 1: rhs partials of x (= vin); fields = 2: v_mid = (v0+vin)/2
// and dv = vin - v0 are written, and rhs partials of v_mid (field 0) and vin
// (field 1); grid_fields (<= fields) says how many of them are scattered.
// part layout [field][blocks][n_dofs], 64-bit fixed point with scale
// 2^62 / w_abs_sum (w_abs_sum = sum |w|).  When mom != nullptr the blocks of
// field 0 also write partial sums (w v1, w v2, w |x|^2) of the scattered
// positions to mom[blocks][3].
double scatter_scale(double w_abs_sum);  // 2^62 / W (fixed-point scale of the scatter)
// An Anderson update v1 <- beta (Gv - sum_j dG_j g_j) + (1 - beta) v1
// (solver.jl:383-412) that a scatter applies to its input markers as it reads
// them (single run: components N apart, dG columns 2N apart).
// The sum over S partial slots part[k * stride + i], k = 0..S-1, in that order
// (the bits of the plain loop), with eight loads in flight instead of one.
__device__ __forceinline__ double slot_sum(const double* __restrict__ part, int S, size_t stride, size_t i) {
  double s = 0.0;
  int k = 0;
  for (; k + 8 <= S; k += 8) {
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) t[u] = part[(size_t)(k + u) * stride + i];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += t[u];
  }
  for (; k < S; ++k) s += part[(size_t)k * stride + i];
  return s;
}

struct UpdateArgs {
  const double* Gv;
  const double* dG;
  double g[kMaxWindow];
  int hist;
  double beta;
};
void launch_scatter(const MeshView& m, const double* v0, const double* vin, const double* w, int N, int fields,
                    int grid_fields, double w_abs_sum, double* vmid, double* dv, long long* part, double* mom,
                    hipStream_t st, const UpdateArgs* upd = nullptr, double* vin_out = nullptr);
// c[field] = M^{-1} rhs[field]; rhs is the exact sum of the fixed-point partial
// rows, or read from `direct[field * n_dofs]` when part == nullptr.  One
// workgroup per field.
void launch_reduce_solve(const MeshView& m, const long long* part, int nblocks, double w_abs_sum,
                         const double* direct, int fields, double* out, hipStream_t st);
// The same solve as two launches over nd2 and nd1 workgroups per field (tmp:
// fields x n_dofs doubles), bitwise equal to launch_reduce_solve; with sum_out,
// also *sum_out = the fixed-order sum of sum_part[0, sum_n) (as launch_sum_rows).
// Needs nd1, nd2 <= 256.
void launch_solve_split(const MeshView& m, const long long* part, int nblocks, double w_abs_sum,
                        const double* direct, int fields, double* out, double* tmp, hipStream_t st,
                        const double* sum_part = nullptr, int sum_n = 0, double* sum_out = nullptr);
// rhs (no solve) from the fixed-point partial rows of field 0
void launch_rhs(const MeshView& m, const long long* part, int nblocks, double w_abs_sum, double* out,
                hipStream_t st);
// Quadrature pass 1: per element and Gauss point.  For field y: iota[y] (if
// non-null) gets (1 + log f) w_q |e| (0 where f is below the floor) and spart[y]
// (if non-null) partial sums of -f log f w_q |e|.  coeff[y] points at n_dofs.
int quad_blocks(const MeshView& m);
// Lenard-Bernstein right-hand side at x (device pointers): with coeff, g = the
// clamped log-gradient of f = sum coeff phi at x (fs_floor 1e-30, g_max 100),
// else g is given; then the block sums (part: lb_blocks(N) x kLbSums), the
// reduction, lambda (qpart: nq entropy partials of S(f_in), nq = 0: none), the
// 3x3 drift multipliers (coef = lambda, A1, A2, B, S1) and dot = -nu (G_eff + A + B x).
constexpr int kLbSums = 12;
int lb_blocks(int N);
void launch_lb_rhs(const MeshView& m, const double* x, const double* coeff, const double* w, const double* dv, int N,
                   const double* qpart, int nq, const double* S0, double nu, double* g, double* part, double* coef,
                   double* dot, hipStream_t st);
// The gather of launch_lb_rhs (coeff given) in one launch with, when cin is
// given, the quadrature partials qpart (quad_blocks(m)) of S(cin); its
// reduction, lambda and v_dot then come from launch_epilogue_lb.
void launch_lb_gather_q(const MeshView& m, const double* x, const double* coeff, const double* cin, int use_logsq,
                        const double* w, const double* dv, int N, double* g, double* part, double* qpart,
                        hipStream_t st);
// Run-record diagnostics of f_s = sum c_k phi_k: out[0] = int max(-f_s, 0) and,
// when v is given, out[1] = ||f_s - f_p||_2^2 with f_p the markers' cell
// histogram density (device pointers; part: scatter_blocks(N) x ne1*ne2,
// cell_mass: ne1*ne2, spart: 2 x quad_blocks).
void launch_fs_diagnostics(const MeshView& m, const double* v, const double* w, int N, double w_abs_sum,
                           const double* c, long long* part, double* cell_mass, double* spart, double* out,
                           hipStream_t st);
void launch_quad1(const MeshView& m, const double* c0, const double* c1, int fields, int use_logsq, double* iota,
                  double* spart, hipStream_t st);
// r_k = sum over cells / points of phi_k iota  (compute_r!, functions.jl:158-179)
void launch_quad2(const MeshView& m, const double* iota, double* r, hipStream_t st);
// Quadrature passes 1 and 2 in one launch: r = compute_r!(c0) with the pass-1
// integrand evaluated where pass 2 reads it (the same arithmetic), and, when
// c1 is given, the block partials of S(c1) into spart[0 .. quad_blocks).
void launch_quad_r(const MeshView& m, const double* c0, const double* c1, int use_logsq, double* r, double* spart,
                   hipStream_t st);
// out[i] = sum_{j<n} part[i * istride + j * jstride] in fixed order, i < count
void launch_sum_rows(const double* part, int n, int count, int istride, int jstride, double* out, hipStream_t st);
// Gather: G (mode 0), or dS = -w G plus partial (dv.dS, |dv|^2) (mode 1),
// or clamped log-gradient (mode 2), or (f, |grad f|) (mode 3).
void launch_gather(const MeshView& m, const double* x, const double* coeff, const double* w, const double* dv,
                   int N, int mode, double fs_floor, double g_max, double* out, double* dpart, hipStream_t st);

// Batched forms (one launch for every run; see EnsView):
//   scatter:  part [field][n_sblk][n_dofs], mom [n_sblk][3]
//   solve:    grid (R, fields), from scatter rows or vec slot in_slot, to vec slot
//             out_slot (or fc[r] when out_slot < 0); vec is [R][4][n_dofs]
//   quad1:    coefficients c + r * rstride + field * n_dofs; iota [R][points],
//             spart [R][2][quad blocks];  quad2: vec slot 2 <- r(iota)
//   sum_rows: out[r * ostride + c] = sum_{j < cnt[r]} part[row0[r] + c istride + j jstride]
//   gather:   coefficients vec slot `slot`, dpart [n_gblk][2]
void launch_scatter_b(const MeshView& m, const EnsView& E, const double* v0, const double* vin, const double* w,
                      int fields, int grid_fields, double* vmid, double* dv, long long* part, double* mom,
                      hipStream_t st);
void launch_solve_b(const MeshView& m, const EnsView& E, const long long* part, int fields, int in_slot, int out_slot,
                    double* vec, double* fc, hipStream_t st);
void launch_quad1_b(const MeshView& m, const EnsView& E, const double* c, int rstride, int fields, int use_logsq,
                    double* iota, double* spart, hipStream_t st);
void launch_quad2_b(const MeshView& m, const EnsView& E, const double* iota, double* vec, hipStream_t st);
void launch_sum_rows_b(const EnsView& E, const double* part, const int* row0, const int* cnt, int count, int istride,
                       int jstride, double* out, int ostride, hipStream_t st);
void launch_gather_b(const MeshView& m, const EnsView& E, const double* x, const double* vec, int slot,
                     const double* w, const double* dv, int mode, double* out, double* dpart, hipStream_t st);

// ---- solver side (step.hip) -------------------------------------------------
// Gonzalez lambda from reduced partials, then G_eff and the pair-source pack
// (s64 / s32 may be null).  Optional: geff_out[2N] <- G_eff, and the frozen-
// metric product D[2N] <- A (G_eff - Gk) with A = [A11, A12, A22] (3N).
void launch_geff_pack(const double* vmid, const double* dS, const double* dv, const double* w, int N,
                      const double* dpart, int nparts, const double* S1, const double* S0, int use_gonzalez,
                      double lo1, double hi1, double lo2, double hi2, Src64* s64, Src32* s32, double* lambda_out,
                      hipStream_t st, double* geff_out = nullptr, const double* Gk = nullptr,
                      const double* A = nullptr, double* D = nullptr);

struct EpiArgs {
  const double* part; int S;
  const double* v0; const double* v1; double dt;
  double* Gv; double* r; const double* rp; const double* Gp;
  double* dF; double* dG; int col; int hist;
  double* stat_part;  // [blocks][kStatWidth]: |r|^2, |v1|^2, |Gv-v0|^2, ATA row, ATr, |r_map|^2 - |r|^2
  // stiff clusters (null: none): row of marker i in the cluster members (-1:
  // none); the map value of such a row is cl_x[2 row + comp] (the cluster
  // solve's); the plain map's residual there enters statistic kStatTrue
  const int* cl_row; const double* cl_x;
  // device-resident solve (null: host control): gate = {live, kind, col, hist}
  // of the map, written by the control kernel of the previous map; the launch
  // returns at once when live = 0, takes col / hist from it, and S32 slices
  // instead of S when kind = 1 (FP32c)
  const int* gate; int S32;
};
int epi_blocks(int N);
constexpr int kStatTrue = 3 + 2 * kMaxWindow;  // |Gv_map - v1|^2 - |Gv - v1|^2 over the cluster rows
constexpr int kStatWidth = kStatTrue + 1;
void launch_epilogue(const EpiArgs& a, int N, hipStream_t st);
// The statistics of n rows of epilogue partials (launch_sum_rows' sums and
// order) into the pinned host buffer out (device address), then seq into
// *seq_out (pinned, device address) once they are visible to the host.
void launch_stat_publish(const double* part, int n, double* out, unsigned* seq_out, unsigned seq, hipStream_t st);
// v1 = beta (Gv - sum_j dG_j gamma_j) + (1 - beta) v1; with best_dst, first
// best_dst <- best_src (2N values; best_src may be v1 itself)
void launch_update(double* v1, const double* Gv, const double* dG, const double* gam, int hist, double beta, int N,
                   hipStream_t st, const double* best_src = nullptr, double* best_dst = nullptr);
// Near-field correction of the frozen-metric map (pairs in CSR order: targets
// tg[nt], partners al[off[t] .. off[t+1])): D -= sum (exact - frozen) pair terms.
void launch_near(int nt, const int* tg, const int* off, const int* al, int N, const double* w, const double* vm,
                 const double* Ge, const double* vmk, const double* Gk, double* D, hipStream_t st);
// The frozen-metric map's G_eff, D = A (G_eff - G_k) and near-pair correction
// (k_geff_pack + k_near), computed inside the Anderson epilogue instead of in
// launches of their own.  nidx[a] = the near-target row of marker a, or -1.
struct FrozenArgs {
  const double* dS;
  const double* dv;
  const double* w;
  const double* Gk;
  const double* A;      // metric, 3N: A11, A12, A22
  const double* dpart;  // gather partials (dv . dS, |dv|^2)
  int nparts;
  const double* S1;
  const double* S0;
  int use_gonzalez;
  const int* nidx;      // null: no near field
  const int* off;
  const int* al;
  const double* vm;     // v_mid of the inner iterate
  const double* vmk;    // v_mid of v_k
};
void launch_epilogue_frozen(const EpiArgs& a, const FrozenArgs& f, int N, hipStream_t st);
// The Lenard-Bernstein map's value inside the Anderson epilogue: every block
// reduces the gather partials (and the entropy partials) in the fixed order of
// launch_lb_rhs, solves for lambda and the drift multipliers, and forms
// v_dot = -nu (G_eff + A + B x) of its elements; block 0 also writes coef =
// (lambda, A1, A2, B, S1).  The same bits as launch_lb_rhs + launch_epilogue.
struct LbEpiArgs {
  const double* x;      // v_mid (2N)
  const double* g;      // clamped log-gradient at v_mid (2N)
  const double* dv;     // v_in - v (2N)
  const double* w;
  const double* part;   // nparts = lb_blocks(N) rows of kLbSums gather partials
  int nparts;
  const double* qpart;  // nq partials of S(f_in); nq = 0: no Gonzalez lambda
  int nq;
  const double* S0;
  double nu;
  double* coef;
};
void launch_epilogue_lb(const EpiArgs& a, const LbEpiArgs& l, int N, hipStream_t st);
// nidx[tg[t]] = t for t < nt (nidx cleared to -1 first)
void launch_near_index(int nt, const int* tg, int N, int* nidx, hipStream_t st);
// out = v + dt * (sum over slices of part)  (the Euler predictor, main.jl:245)
void launch_euler(const double* part, int S, const double* v, double dt, int N, double* out, double* dotv,
                  hipStream_t st);

// Batched forms of the solver-side kernels.  scal is [R][8]: S0, S1, lambda,
// P1, P2, sum w |v|^2.
struct EnsBufs {
  const double* v0;    // state
  double* v1;          // iterate
  double* R[2];        // residual ping-pong
  double* GV[2];       // map value ping-pong
  double* dF;          // [kMaxWindow][2][NT]
  double* dG;
  double* best;
  const double* part;  // pair partials (run r at poff)
  double* statp;       // [n_eblk][kStatWidth]
};
void launch_geff_pack_b(const EnsView& E, const double* vmid, const double* dS, const double* dv, const double* w,
                        const double* dpart, double* scal, int use_gonzalez, double lo1, double hi1, double lo2,
                        double hi2, Src64* s64, Src32* s32, hipStream_t st);
void launch_epilogue_b(const EnsView& E, const EnsBufs& B, double dt, hipStream_t st);
void launch_update_b(const EnsView& E, const EnsBufs& B, hipStream_t st);
void launch_euler_b(const EnsView& E, const double* part, const double* v, double dt, double* out, double* dotv,
                    hipStream_t st);
void launch_copy_b(const EnsView& E, const double* src, double* dst, hipStream_t st);

// ---- batched pair sum (pair.hip) ---------------------------------------------------
// Persistent launch over a tile list: tile t = (run, first target, slice),
// `bt` targets per workgroup (64, 128, 256 or 512), `grid` workgroups striding
// over the list.  force_f64: every run, FP64 (predictor); otherwise the runs
// active in E.ctl, in their own precision.
struct EnsTiles {
  const int4* tiles;  // ordered: {run, first target, slice, -}; symmetric: {run, a, b, -}
  int ntiles;
  int bt;
  int grid;
  int sym;  // 0: ordered tiles; R: symmetric blocks of R waves
};
void launch_pack_b(const EnsView& E, const double* v, const double* G, const double* w, double lo1, double hi1,
                   double lo2, double hi2, Src64* s64, Src32* s32, hipStream_t st);
// the symmetric blocks of T (T.sym != 0) over every active run
void launch_pair_sym_b(const EnsView& E, const EnsTiles& T, const Src64* s64, const Src32* s32, int force_f64,
                       double* part, hipStream_t st);
void launch_pair_b(const EnsView& E, const EnsTiles& T, const Src64* s64, const Src32* s32, int force_f64,
                   double* part, hipStream_t st);

// ---- warm-start MLP (mlp.hip) --------------------------------------------------
constexpr int kMlpParams = 4738;
struct MlpArgs {
  const float* params;  // kMlpParams
  const float* rec;     // [samples][8]: x0..x5 raw features, t0, t1 raw targets
  long long nsamples;
  const int* tiles;     // tile ids (16 samples each), or nullptr = 0..ntiles-1
  int ntiles;
  float mu[6], isd[6], oscale[2];
  float inv_batch;      // 2 / B: gradient of the batch-mean squared error
  int mode;             // 0: y only, 1: + loss, 2: + gradients
  float* y;             // [samples][2] (mode 0)
  double* loss_part;    // [grid]
  float* grad_part;     // [grid][kMlpParams]
};
void launch_mlp(const MlpArgs& a, int grid, hipStream_t st);
// grad = row sum of part; apply != 0: Adam step (lr, step >= 1) on params
void launch_adam(const float* part, int rows, float* params, float* m, float* v, float* grad_out, float lr, int step,
                 int apply, hipStream_t st);
void launch_features(const double* v, const double* G, const double* fg, int N, float* rec, hipStream_t st);
void launch_targets(const double* v_old, const double* v_new, const double* dotv, double dt, int N, float* rec,
                    hipStream_t st);
void launch_warm_apply(const double* v, const double* dotv, const float* y, float os0, float os1, double dt, int N,
                       double* v1, hipStream_t st);


// ---- stiff clusters (cluster.hip) --------------------------------------------
constexpr int kClusterMax = 32;         // largest cluster solved exactly (markers)
constexpr int kLinkCap = 1 << 17;       // near links kept per step (beyond it: no clusters that step)
constexpr int kClMembersCap = 1 << 15;  // cluster members per step (beyond it: no clusters that step)
constexpr int kClusterTries = 2;        // failed solves of a cluster per step before it keeps the plain rows
constexpr int kClusterPathIts = 60;     // Newton iterations along a cluster's homotopy path
constexpr int kClusterIts = 90;         // ... including the eigen-line starts
struct NearLink {
  int i, j;  // i < j
  double q;  // |v_i - v_j|^2
};
// x[0 .. n) <- value
void launch_fill_int(int* x, int n, int value, hipStream_t st);
// Markers of v (N) bucketed by mesh cell: cell_of[N] (-1 outside the box),
// start[ncell + 1] (cnt[ncell] is scratch), sorted[] = marker ids cell by cell.
void launch_cell_sort(const MeshView& m, const double* v, int N, int* cell_of, int* cnt, int* start, int* sorted,
                      hipStream_t st);
// Every unordered in-box pair with |d|^2 >= 1e-24 and max(w_i, w_j) / |d| >
// theta (w_max = max w bounds the reach across cell edges) into out[] (count
// receives the number found; entries beyond cap are dropped).
void launch_near_cells(const MeshView& m, const double* v, const double* w, int N, double theta, double w_max,
                       const int* start, const int* sorted, NearLink* out, int* count, int cap, hipStream_t st);
// row[idx[r]] = r (set) or -1 (reset) for r < total
void launch_cluster_rows(const int* idx, int total, bool set, int* row, hipStream_t st);
struct ClusterArgs {
  MeshView m;
  const int* off;        // clusters: members idx[off[c] .. off[c + 1]) (<= kClusterMax each)
  const int* idx;
  int N;
  double dt;
  const double* v0;      // state (start of the step)
  const double* w;
  const double* vmid;    // the map's v_mid, dS = -w grad L(v_mid), dv = v_in - v0
  const double* dS;
  const double* dv;
  const double* lam;     // the map's Gonzalez lambda (device scalar)
  const double* L;       // the map's M^-1 r (n_dofs)
  const double* part;    // the map's pair-sum partials [S][2N]
  int S;
  double* mroot;         // 2 per member: the cluster's last root (warm start)
  int* have_root;        // per cluster: 1 root of an earlier map, 0 none, -f after f failed solves
  double* x_out;         // 2 per member: the map value of the member's rows
  int* fail_count;       // += clusters whose solve reached no root
  int* newton_its;       // per cluster: += Newton iterations
};
void launch_cluster_solve(const ClusterArgs& a, int ncl, hipStream_t st);

}  // namespace collop
