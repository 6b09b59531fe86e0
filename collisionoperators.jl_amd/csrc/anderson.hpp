// anderson.hpp -- host-side state machine of the implicit solve
// (step_anderson!, solver.jl:300-423), shared by the single-run step and every
// run of the batched ensemble.  It decides, from the reduced statistics of each
// Picard map, what the device does next; it owns no device memory.
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
#include <cstdio>
#include <limits>

#include "../../include/collop.h"
#include "collop_launch.hpp"

namespace collop {

// Small SPD solve (Cholesky; Julia's factorize picks it for the exactly
// symmetric ATA), with partial-pivot LU as the fallback.
void small_solve(int n, const double* A0, const double* b0, double* x);

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

  AndersonPlan plan() {
    const bool anderson_step = p_.use_anderson && since_reset_ > 0 && !rebase_;
    AndersonPlan pl;
    pl.col = anderson_step ? slot_ % mw_ : -1;
    pl.hist = anderson_step ? std::min(history_ + 1, mw_) : (rebase_ ? history_ : 0);
    pl.kind = use32_ ? 1 : 0;
    plan_ = pl;
    return pl;
  }

  // Returns COLLOP_OK or COLLOP_ERR_NONFINITE.
  int after_map(const double* st, AndersonAct& a, const char* trace_tag = nullptr) {
    a = AndersonAct{};
    const int k = ++k_;
    const int col = plan_.col, hist_new = plan_.hist;
    since_reset_++;
    const double nrm_r = std::sqrt(st[0]);
    const double nrm_v1 = std::sqrt(st[1]);
    if (!std::isfinite(nrm_r) || !std::isfinite(nrm_v1)) return COLLOP_ERR_NONFINITE;
    // the plain map's residual: the exit, best-iterate and stagnation tests
    // (with stiff clusters, nrm_r is that of the map Anderson mixes)
    const double nrm_map = std::sqrt(std::max(st[0] + st[kStatTrue], 0.0));
    if (k == 1) nrm_r0 = nrm_map;
    const double eff_tol = std::max(p_.tol * (nrm_v1 + 1e-30), p_.abs_floor);
    if (use32_) {
      iters32++;
      // FP32 maps run until their own rounding floor (a few FP32 ulps of the
      // step increment |Gv - v0|), or until four maps bring no new best
      // residual; FP64 maps then finish, so every accepted iterate and the
      // exit map are FP64.  Switching well inside the basin keeps the FP64
      // maps on the root the plain FP64 solve finds (the implicit system can
      // have several roots at large dt).
      const double floor32 = floor_factor_ * 5.96e-8 * std::sqrt(st[2]);
      if (nrm_r < best32_) {
        best32_ = nrm_r;
        since_best32_ = 0;
      } else {
        since_best32_++;
      }
      const bool switch_now = nrm_r < std::max(floor32, eff_tol) || since_best32_ >= 4 || k >= p_.max_iter / 2;
      if (trace_tag)
        std::fprintf(stderr, "  [%s f32] k=%d |r|=%.3e floor=%.3e best=%.3e\n", trace_tag, k, nrm_r, floor32,
                     best32_);
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
      if (trace_tag)
        std::fprintf(stderr, "  [%s f64] k=%d |r|=%.3e (mixed %.3e) tol=%.3e\n", trace_tag, k, nrm_map, nrm_r, eff_tol);
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
      for (int j = 0; j < history_; ++j) {
        ATA_[col * kMaxWindow + j] = st[3 + j];
        ATA_[j * kMaxWindow + col] = st[3 + j];
      }
    }
    const int h = history_;
    double A[kMaxWindow * kMaxWindow], b[kMaxWindow];
    double dm = 0.0;
    for (int i = 0; i < h; ++i) {
      for (int c = 0; c < h; ++c) A[i * h + c] = ATA_[i * kMaxWindow + c];
      b[i] = st[3 + kMaxWindow + i];
      dm += A[i * h + i];
    }
    dm /= h;
    const double l2 = p_.reg_factor * dm + 1e-30;
    for (int i = 0; i < h; ++i) A[i * h + i] += l2;
    small_solve(h, A, b, a.gamma);
    a.h = h;
    return COLLOP_OK;
  }

  int maps() const { return k_; }

  // results of the step
  int iters32 = 0, n_restart = 0, exit_reason = 2, k_done = 0;
  double nrm_r0 = 0.0, res_out = 0.0;
  bool done = false;

 private:
  // The map budget is spent: the best iterate (the initial guess if no FP64
  // map improved on it) is the result.
  int finish_max_iter(AndersonAct& a) {
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

}  // namespace collop
