// anderson.hpp -- the state machine of the implicit solve (step_anderson!,
// solver.jl:300-423), shared by the single-run step and every run of the
// batched ensemble.  It decides, from the reduced statistics of each Picard
// map, what the device does next; it owns no device memory.  The same code runs
// on the host (ensemble, traced and cluster solves) and inside the control
// kernel of the device-resident solve (anderson_dev.hip); floating-point
// contraction is off in it, so both give bitwise the same decisions.
//
// Per map k the caller:
//   1. asks plan() for the map's precision, Anderson column and window width,
//   2. runs the map (epilogue statistics st[kStatWidth]: |r|^2, |v1|^2,
//      |Gv - v0|^2, new Gram row dF_col . dF_j, dF_j . r),
//   3. passes st to after_map(), which returns the device actions in order:
//      copy_best (best <- Gv or v1 of this map), then either an exit (final
//      iterate: Gv, v1 or best) or an update of v1 (Anderson or damped Picard
//      step with gamma, beta), then whether (r, Gv) rotate into (rp, Gp).
#pragma once
#include <algorithm>
#include <cmath>
#include <limits>

#include "../../include/collop.h"
#include "collop_launch.hpp"

namespace collop {

// Small SPD solve (Cholesky; Julia's factorize picks it for the exactly
// symmetric ATA), with partial-pivot LU as the fallback.  n <= kMaxWindow.
__host__ __device__ inline void small_solve(int n, const double* A0, const double* b0, double* x) {
#pragma clang fp contract(off)
  double A[kMaxWindow * kMaxWindow], L[kMaxWindow * kMaxWindow], b[kMaxWindow];
  for (int i = 0; i < n * n; ++i) A[i] = L[i] = A0[i];
  for (int i = 0; i < n; ++i) b[i] = b0[i];
  bool ok = true;
  for (int j = 0; j < n && ok; ++j) {
    double d = L[j * n + j];
    for (int k = 0; k < j; ++k) d -= L[j * n + k] * L[j * n + k];
    if (!(d > 0.0)) {
      ok = false;
      break;
    }
    d = sqrt(d);
    L[j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = L[i * n + j];
      for (int k = 0; k < j; ++k) t -= L[i * n + k] * L[j * n + k];
      L[i * n + j] = t / d;
    }
  }
  if (ok) {
    double y[kMaxWindow];
    for (int i = 0; i < n; ++i) {
      double t = b[i];
      for (int k = 0; k < i; ++k) t -= L[i * n + k] * y[k];
      y[i] = t / L[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double t = y[i];
      for (int k = i + 1; k < n; ++k) t -= L[k * n + i] * x[k];
      x[i] = t / L[i * n + i];
    }
    return;
  }
  for (int k = 0; k < n; ++k) {
    int pr = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + k]) > fabs(A[pr * n + k])) pr = i;
    if (pr != k) {
      for (int j = 0; j < n; ++j) {
        const double t = A[k * n + j];
        A[k * n + j] = A[pr * n + j];
        A[pr * n + j] = t;
      }
      const double t = b[k];
      b[k] = b[pr];
      b[pr] = t;
    }
    for (int i = k + 1; i < n; ++i) {
      const double f = A[i * n + k] / A[k * n + k];
      for (int j = k; j < n; ++j) A[i * n + j] -= f * A[k * n + j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double t = b[i];
    for (int j = i + 1; j < n; ++j) t -= A[i * n + j] * x[j];
    x[i] = t / A[i * n + i];
  }
}

// The window's least squares: the same solve on an n x n matrix stored with
// row stride kMaxWindow (n <= kMaxWindow), padded to kMaxWindow with the
// identity, so that every loop has constant bounds and the device keeps the
// factor in registers.  The padding adds exact zeros to every sum of the
// leading n x n solve, whose result it therefore leaves unchanged.  A matrix
// that is not positive definite goes to small_solve's pivoted LU.
__host__ __device__ __forceinline__ void window_solve(int n, const double* A0, const double* b0, double* x) {
#pragma clang fp contract(off)
  constexpr int M = kMaxWindow;
  double L[M * M], y[M], xs[M];
  bool ok = true;
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < M; ++j) L[i * M + j] = (i < n && j < n) ? A0[i * M + j] : (i == j ? 1.0 : 0.0);
#pragma unroll
  for (int j = 0; j < M; ++j) {
    double d = L[j * M + j];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= L[j * M + k] * L[j * M + k];
    ok = ok && d > 0.0;
    d = sqrt(d);
    L[j * M + j] = d;
#pragma unroll
    for (int i = j + 1; i < M; ++i) {
      double t = L[i * M + j];
#pragma unroll
      for (int k = 0; k < j; ++k) t -= L[i * M + k] * L[j * M + k];
      L[i * M + j] = t / d;
    }
  }
  if (ok) {
#pragma unroll
    for (int i = 0; i < M; ++i) {
      double t = i < n ? b0[i] : 0.0;
#pragma unroll
      for (int k = 0; k < i; ++k) t -= L[i * M + k] * y[k];
      y[i] = t / L[i * M + i];
    }
#pragma unroll
    for (int i = M - 1; i >= 0; --i) {
      double t = y[i];
#pragma unroll
      for (int k = i + 1; k < M; ++k) t -= L[k * M + i] * xs[k];
      xs[i] = t / L[i * M + i];
    }
#pragma unroll
    for (int i = 0; i < M; ++i)
      if (i < n) x[i] = xs[i];
    return;
  }
  double A[M * M], b[M];
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) A[i * n + j] = A0[i * M + j];
    b[i] = b0[i];
  }
  small_solve(n, A, b, x);
}


struct AndersonPlan {
  int kind;  // 0 FP64 map, 1 FP32c map
  int col;   // Anderson column the map writes (-1: none)
  int hist;  // window columns dotted with the new column and the residual
};

enum AndersonExit { kContinue = 0, kConverged = 1, kStagnated = 2, kMaxIter = 3 };

struct AndersonAct {
  bool copy_best = false;  // best <- (exit_picard_step ? Gv : v1) of this map
  int exit = kContinue;    // kConverged: final = exit_picard_step ? Gv : v1; otherwise final = best
  bool update = false;     // v1 <- beta (Gv - dG gamma) + (1 - beta) v1
  int h = 0;
  double beta = 0.0;
  double gamma[kMaxWindow] = {0};
  bool rotate = false;     // rp <- r, Gp <- Gv
};

class AndersonRun {
 public:
  // clusters: the maps carry stiff-cluster rows (statistic kStatTrue); the
  // FP32 window is then not kept across the switch (its maps had none)
  void start(const collop_step_params& p, bool clusters = false) {
    p_ = p;
    clusters_ = clusters;
    mw_ = p.m_anderson;
    history_ = slot_ = n_restart = 0;
    nrm_r0 = 0.0;
    nrm_best_ = nrm_best_window_ = std::numeric_limits<double>::infinity();
    std::fill(ATA_, ATA_ + kMaxWindow * kMaxWindow, 0.0);
    use32_ = p.precision == COLLOP_PREC_F32C;
    iters32 = 0;
    floor_factor_ = p.f32_floor_factor > 0.0 ? p.f32_floor_factor : 16.0;
    exit_reason = 2;
    k_done = p.max_iter;
    res_out = 0.0;
    since_reset_ = 0;
    best32_ = std::numeric_limits<double>::infinity();
    since_best32_ = 0;
    rebase_ = false;
    k_ = 0;
    done = false;
  }

  __host__ __device__ __forceinline__ AndersonPlan plan() {
    const bool anderson_step = p_.use_anderson && since_reset_ > 0 && !rebase_;
    AndersonPlan pl;
    pl.col = anderson_step ? slot_ % mw_ : -1;
    pl.hist = anderson_step ? (history_ + 1 < mw_ ? history_ + 1 : mw_) : (rebase_ ? history_ : 0);
    pl.kind = use32_ ? 1 : 0;
    plan_ = pl;
    return pl;
  }

  // Returns COLLOP_OK or COLLOP_ERR_NONFINITE.
  __host__ __device__ __forceinline__ int after_map(const double* st, AndersonAct& a) {
#pragma clang fp contract(off)
    a = AndersonAct{};
    const int k = ++k_;
    const int col = plan_.col, hist_new = plan_.hist;
    since_reset_++;
    const double nrm_r = sqrt(st[0]);
    const double nrm_v1 = sqrt(st[1]);
    last_nrm_r = nrm_r;
    last_eff_tol = 0.0;
    if (!__builtin_isfinite(nrm_r) || !__builtin_isfinite(nrm_v1)) return COLLOP_ERR_NONFINITE;
    // the plain map's residual: the exit, best-iterate and stagnation tests
    // (with stiff clusters, nrm_r is that of the map Anderson mixes)
    const double nrm_map = sqrt(fmax(st[0] + st[kStatTrue], 0.0));
    last_nrm_map = nrm_map;
    last_f32 = use32_;
    if (k == 1) nrm_r0 = nrm_map;
    const double eff_tol = fmax(p_.tol * (nrm_v1 + 1e-30), p_.abs_floor);
    last_eff_tol = eff_tol;
    if (use32_) {
      iters32++;
      // FP32 maps run until their own rounding floor (a few FP32 ulps of the
      // step increment |Gv - v0|), or until four maps bring no new best
      // residual; FP64 maps then finish, so every accepted iterate and the
      // exit map are FP64.  Switching well inside the basin keeps the FP64
      // maps on the root the plain FP64 solve finds (the implicit system can
      // have several roots at large dt).
      const double floor32 = floor_factor_ * 5.96e-8 * sqrt(st[2]);
      last_floor32 = floor32;
      if (nrm_r < best32_) {
        best32_ = nrm_r;
        since_best32_ = 0;
      } else {
        since_best32_++;
      }
      const bool switch_now = nrm_r < fmax(floor32, eff_tol) || since_best32_ >= 4 || k >= p_.max_iter / 2;
      if (switch_now && k >= p_.max_iter) return finish_max_iter(a);
      if (switch_now && p_.f32_keep_window && !clusters_ && p_.use_anderson && history_ > 0) {
        // Keep the secant window built from FP32 maps (its columns are
        // differences far above FP32 rounding) and re-evaluate this iterate
        // with an FP64 map before the next update: no column mixes the two.
        use32_ = false;
        rebase_ = true;
        return COLLOP_OK;
      }
      if (switch_now) {
        use32_ = false;
        history_ = 0;
        slot_ = 0;
        since_reset_ = 0;
        a.update = true;
        a.h = 0;
        a.beta = p_.damping;
        a.rotate = true;
        return COLLOP_OK;
      }
    } else {
      if (nrm_map < nrm_best_) {
        nrm_best_ = nrm_map;
        a.copy_best = true;
      }
      if (nrm_map < eff_tol) {
        a.exit = kConverged;
        exit_reason = 0;
        k_done = k;
        res_out = nrm_map;
        done = true;
        return COLLOP_OK;
      }
      if (k > p_.stag_window && k % p_.stag_window == 0) {
        const double rel = (nrm_best_window_ - nrm_best_) / (nrm_best_window_ + 1e-30);
        if (rel < p_.stag_rel_tol) {
          a.exit = kStagnated;
          exit_reason = 1;
          k_done = k;
          res_out = nrm_best_;
          done = true;
          return COLLOP_OK;
        }
        nrm_best_window_ = nrm_best_;
      }
    }
    if (k >= p_.max_iter) return finish_max_iter(a);
    bool just_restarted = false;
    if (!use32_ && k > 1 && nrm_r > p_.restart_factor * nrm_best_) {
      history_ = 0;
      slot_ = 0;
      n_restart++;
      just_restarted = true;
    }
    const double beta = k > p_.damp_decay_start ? p_.damping * p_.damp_decay_factor : p_.damping;
    const bool rebased = rebase_;
    rebase_ = false;
    a.update = true;
    a.beta = beta;
    a.rotate = true;
    if (just_restarted || (col < 0 && !(rebased && history_ > 0))) {
      a.h = 0;
      return COLLOP_OK;
    }
    if (!rebased) {
      slot_ = slot_ % mw_ + 1;
      history_ = hist_new;
      // new Gram row and column (loops over the full window with constant
      // indices, so that the device keeps the matrix in registers)
#pragma unroll
      for (int i = 0; i < kMaxWindow; ++i)
#pragma unroll
        for (int j = 0; j < kMaxWindow; ++j) {
          if (j < history_ && (i == col)) ATA_[i * kMaxWindow + j] = st[3 + j];
          if (i < history_ && (j == col)) ATA_[i * kMaxWindow + j] = st[3 + i];
        }
    }
    const int h = history_;
    double A[kMaxWindow * kMaxWindow], b[kMaxWindow];
    double dm = 0.0;
#pragma unroll
    for (int i = 0; i < kMaxWindow; ++i) {
#pragma unroll
      for (int c = 0; c < kMaxWindow; ++c) A[i * kMaxWindow + c] = (i < h && c < h) ? ATA_[i * kMaxWindow + c] : 0.0;
      b[i] = i < h ? st[3 + kMaxWindow + i] : 0.0;
      if (i < h) dm += A[i * kMaxWindow + i];
    }
    dm /= h;
    const double l2 = p_.reg_factor * dm + 1e-30;
#pragma unroll
    for (int i = 0; i < kMaxWindow; ++i)
      if (i < h) A[i * kMaxWindow + i] += l2;
    window_solve(h, A, b, a.gamma);
    a.h = h;
    return COLLOP_OK;
  }

  __host__ __device__ int maps() const { return k_; }
  __host__ __device__ double best32() const { return best32_; }

  // results of the step
  int iters32 = 0, n_restart = 0, exit_reason = 2, k_done = 0;
  double nrm_r0 = 0.0, res_out = 0.0;
  bool done = false;
  // the last map's figures, for tracing
  double last_nrm_r = 0.0, last_nrm_map = 0.0, last_eff_tol = 0.0, last_floor32 = 0.0;
  bool last_f32 = false;

 private:
  // The map budget is spent: the best iterate (the initial guess if no FP64
  // map improved on it) is the result.
  __host__ __device__ __forceinline__ int finish_max_iter(AndersonAct& a) {
    a.exit = kMaxIter;
    exit_reason = 2;
    k_done = p_.max_iter;
    res_out = nrm_best_;
    done = true;
    return COLLOP_OK;
  }

  collop_step_params p_{};
  int mw_ = 1, history_ = 0, slot_ = 0, since_reset_ = 0, since_best32_ = 0, k_ = 0;
  double nrm_best_ = 0.0, nrm_best_window_ = 0.0, best32_ = 0.0, floor_factor_ = 16.0;
  double ATA_[kMaxWindow * kMaxWindow] = {0};
  bool use32_ = false, rebase_ = false, clusters_ = false;
  AndersonPlan plan_{0, -1, 0};
};

// Control block of the device-resident solve (one per context, in device
// memory): the kernels of map k read gate = {live, kind, col, hist}, which the
// control kernel of map k - 1 wrote.
struct AndCtl {
  int gate[4];
  int maps;  // maps whose control ran
  int err;   // non-finite residual in map `maps`
  AndersonRun run;
  AndersonAct act;  // the action of map `maps`
  double st[kStatWidth];
};

// Control kernel of map k (one workgroup): reduce the epilogue partials
// part[n][kStatWidth], run the state machine, plan map k + 1; publishes
// pub[1..3] = {live, kind, err}, then pub[0] = seq0 + k (pinned memory,
// device-side address).
void launch_anderson_ctl(const double* part, int n, AndCtl* c, int k, unsigned* pub, unsigned seq0, hipStream_t st);
// The action of map k on the 2N components (v1, rp, Gp, best).
void launch_anderson_apply(const AndCtl* c, int k, double* v1, const double* Gv, const double* dG, const double* r,
                           double* rp, double* Gp, double* best, int N, int exit_picard, hipStream_t st);

}  // namespace collop
