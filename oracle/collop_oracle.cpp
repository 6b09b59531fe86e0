// collop_oracle.cpp -- CPU restatement of the reference's Landau collision path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle: the checker that the HIP
// library is compared against.  Only the tests, the smoke entry point and the
// CPU-baseline leg of the benchmark may load it.  Nothing in the shipped
// library calls it, links it or falls back to it.
//
// Every routine restates, in plain FP64 C++, one function of the reference
// (junyixu/CollisionOperators.jl) and cites the file:line it follows.  The
// arithmetic order follows the reference CPU path where that is cheap to do,
// and the file is compiled with -ffp-contract=off so that, like Julia without
// @fastmath/muladd, no multiply-add is fused behind our back.
//
// Mantis.jl 0.6.0 (the FEM dependency of the reference, not vendored in the
// reference tree) supplies the B-spline space, Bezier extraction, Gauss-Legendre
// rules and mass-matrix assembly.  Those pieces are restated from their
// published algorithms:
//   * open knot vector with interior multiplicity P_DEG-K_REG (Piegl & Tiller,
//     "The NURBS Book", ch. 2),
//   * Bezier extraction operators, Borden, Scott, Evans & Hughes (2011),
//     Int. J. Numer. Meth. Engng 87:15-47, Algorithm 1,
//   * Gauss-Legendre nodes by Newton iteration on P_n (Golub-Welsch would do
//     as well), mapped to [0,1],
//   * SuiteSparse `lu(M)` (MantisWrappers.jl:119) is replaced by a banded
//     Cholesky factorisation of the same assembled (SPD) matrix.
// Parity status of each function (what pins it) is listed in the tests that
// exercise it; every function here has at least one independent check.
//
// Layout conventions (mirroring the reference's Julia arrays):
//   v, G, dot_v: N x 2 column-major  -> v[0..N) = v1, v[N..2N) = v2
//   A (metric):  N x 3 column-major
//   extraction C: C[i + (p+1)*j + (p+1)^2*e]  (row i = Bernstein, col j = local
//                 basis), i.e. phi_j = sum_i C[i,j] B_i   (MantisWrappers.jl:237)
//   dof index:   gid = (s1 + j1) + (s2 + j2) * nd1   (0-based; the 1-based form
//                 is projection_gpu.jl:66)

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Gauss-Legendre rule on [0,1] (Mantis Quadrature.gauss_legendre, used at
// MantisWrappers.jl:109-110,122-123).
void gauss_legendre01(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) {
    // Chebyshev-like initial guess for the i-th root of P_n on [-1,1].
    double t = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) { p1 = t; p0 = 1.0; }
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      double dt = p1 / dp;
      t -= dt;
      if (std::fabs(dt) < 1e-16) break;
    }
    // recompute derivative at the converged root
    double p0 = 1.0, p1 = t;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) { p1 = t; p0 = 1.0; }
    dp = n * (t * p1 - p0) / (t * t - 1.0);
    double wt = 2.0 / ((1.0 - t * t) * dp * dp);
    // roots come out descending in t; store ascending on [0,1]
    x[n - 1 - i] = 0.5 * (1.0 + t);
    w[n - 1 - i] = 0.5 * wt;
  }
}

double binom(int n, int k) {
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// MantisWrappers.jl:201-216 (_bernstein_eval!)
void bernstein_eval(int p, double xi, double* B, double* dB) {
  const double om = 1.0 - xi;
  for (int i = 0; i <= p; ++i) B[i] = binom(p, i) * std::pow(xi, i) * std::pow(om, p - i);
  if (p == 0) {
    dB[0] = 0.0;
    return;
  }
  for (int i = 0; i <= p; ++i) {
    double left = i > 0 ? binom(p - 1, i - 1) * std::pow(xi, i - 1) * std::pow(om, p - i) : 0.0;
    double right = i < p ? binom(p - 1, i) * std::pow(xi, i) * std::pow(om, p - 1 - i) : 0.0;
    dB[i] = p * (left - right);
  }
}

// 1D spline data for one velocity direction.
struct Dir {
  int p = 2, k = 1;
  int ne = 0, nd = 0;
  std::vector<double> bp;     // ne+1 breakpoints
  std::vector<double> C;      // (p+1)^2 * ne, reference convention
  std::vector<int> bs;        // 0-based first basis index per element
};

// Open knot vector + Borden et al. (2011) Algorithm 1.  Produces the
// extraction operators of MantisWrappers.jl:128-133 and the basis starts.
void build_dir(Dir& d, const double* bp, int nb, int p, int k) {
  d.p = p;
  d.k = k;
  d.ne = nb - 1;
  d.bp.assign(bp, bp + nb);
  const int mult_int = p - k;
  std::vector<double> U;
  for (int i = 0; i <= p; ++i) U.push_back(bp[0]);
  for (int e = 1; e < nb - 1; ++e)
    for (int r = 0; r < mult_int; ++r) U.push_back(bp[e]);
  for (int i = 0; i <= p; ++i) U.push_back(bp[nb - 1]);
  const int m = (int)U.size();  // = n + p + 1
  d.nd = m - p - 1;
  const int q = p + 1;
  // 1-based helpers for the algorithm as published
  auto Uat = [&](int i) { return U[i - 1]; };
  std::vector<std::vector<double>> Cs;  // Borden convention: rows local basis a, cols Bernstein b
  auto eye = [&]() {
    std::vector<double> I(q * q, 0.0);
    for (int i = 0; i < q; ++i) I[i * q + i] = 1.0;
    return I;
  };
  auto at = [&](std::vector<double>& M, int r, int c) -> double& { return M[(r - 1) * q + (c - 1)]; };
  int a = p + 1, b = a + 1;
  Cs.push_back(eye());
  int nb_ = 1;
  std::vector<double> alphas(q + 1, 0.0);
  while (b < m) {
    Cs.push_back(eye());
    int i = b;
    while (b < m && Uat(b + 1) == Uat(b)) b++;
    int mult = b - i + 1;
    if (mult < p) {
      double numer = Uat(b) - Uat(a);
      for (int j = p; j >= mult + 1; --j) alphas[j - mult] = numer / (Uat(a + j) - Uat(a));
      int r = p - mult;
      for (int j = 1; j <= r; ++j) {
        int save = r - j + 1;
        int s = mult + j;
        for (int kk = p + 1; kk >= s + 1; --kk) {
          double alpha = alphas[kk - s];
          for (int row = 1; row <= q; ++row)
            at(Cs[nb_ - 1], row, kk) = alpha * at(Cs[nb_ - 1], row, kk) + (1.0 - alpha) * at(Cs[nb_ - 1], row, kk - 1);
        }
        if (b < m) {
          for (int t = 0; t <= j; ++t) at(Cs[nb_], save + t, save) = at(Cs[nb_ - 1], p - j + 1 + t, p + 1);
        }
      }
    }
    nb_++;
    if (b < m) {
      a = b;
      b = b + 1;
    }
  }
  d.C.assign((size_t)q * q * d.ne, 0.0);
  for (int e = 0; e < d.ne; ++e)
    for (int i = 0; i < q; ++i)      // Bernstein
      for (int j = 0; j < q; ++j)    // local basis
        d.C[i + q * j + q * q * e] = Cs[e][j * q + i];
  // basis start: span index of element e minus p (Mantis get_basis_indices)
  d.bs.assign(d.ne, 0);
  int e = 0;
  for (int l = p; l < m - p - 1 && e < d.ne; ++l) {
    if (U[l] < U[l + 1]) {
      d.bs[e] = l - p;
      e++;
    }
  }
}

// phi_j = sum_i C[i,j] B_i  (MantisWrappers.jl:237-238, mul!(phi, transpose(C), B))
void local_basis(const Dir& d, int e, double xi, double* phi, double* dphi) {
  const int q = d.p + 1;
  double B[16], dB[16];
  bernstein_eval(d.p, xi, B, dB);
  const double* C = &d.C[(size_t)q * q * e];
  for (int j = 0; j < q; ++j) {
    double s = 0.0, ds = 0.0;
    for (int i = 0; i < q; ++i) {
      s += C[i + q * j] * B[i];
      ds += C[i + q * j] * dB[i];
    }
    phi[j] = s;
    if (dphi) dphi[j] = ds;
  }
}

// searchsortedlast (MantisWrappers.jl:171), 0-based result in [0, ne-1]
int search_last(const std::vector<double>& bp, double x) {
  auto it = std::upper_bound(bp.begin(), bp.end(), x);
  return (int)(it - bp.begin()) - 1;
}

// Banded symmetric positive definite matrix, lower band storage:
// L[i*(bw+1) + (i-j)] for 0 <= i-j <= bw.
struct Band {
  int n = 0, bw = 0;
  std::vector<double> a;
  double& at(int i, int j) { return a[(size_t)i * (bw + 1) + (i - j)]; }
  double get(int i, int j) const {
    if (i < j) std::swap(i, j);
    if (i - j > bw) return 0.0;
    return a[(size_t)i * (bw + 1) + (i - j)];
  }
};

bool band_cholesky(Band& M) {
  const int n = M.n, bw = M.bw;
  for (int j = 0; j < n; ++j) {
    double s = M.at(j, j);
    for (int k = std::max(0, j - bw); k < j; ++k) s -= M.at(j, k) * M.at(j, k);
    if (!(s > 0.0)) return false;
    double ljj = std::sqrt(s);
    M.at(j, j) = ljj;
    for (int i = j + 1; i <= std::min(n - 1, j + bw); ++i) {
      double t = M.at(i, j);
      for (int k = std::max(0, i - bw); k < j; ++k) t -= M.at(i, k) * M.at(j, k);
      M.at(i, j) = t / ljj;
    }
  }
  return true;
}

void band_solve(const Band& L, const double* b, double* x) {
  const int n = L.n, bw = L.bw;
  std::vector<double> y(b, b + n);
  for (int i = 0; i < n; ++i) {
    double s = y[i];
    for (int k = std::max(0, i - bw); k < i; ++k) s -= L.get(i, k) * y[k];
    y[i] = s / L.get(i, i);
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k <= std::min(n - 1, i + bw); ++k) s -= L.get(k, i) * y[k];
    y[i] = s / L.get(i, i);
  }
  std::memcpy(x, y.data(), sizeof(double) * n);
}

}  // namespace

// ---------------------------------------------------------------------------
// Workspace (MantisWrappers.jl:39-80, build_workspace :90-158)
struct orc_mesh {
  Dir d1, d2;
  int p = 2, k = 1, nq = 6;
  int n_dofs = 0, n_elem = 0;
  std::vector<double> qx, qw;   // N_QUAD rule on [0,1]
  Band M;                       // Cholesky factor of the assembled 2D mass matrix
  std::vector<double> Mdense;   // assembled matrix (kept only for small meshes)
};

struct orc_solver_params {
  double dt;
  int m_anderson;
  int max_iter;
  double tol;
  double abs_floor;
  int stag_window;
  double stag_rel_tol;
  int damp_decay_start;
  double damp_decay_factor;
  double restart_factor;
  double damping;
  double reg_factor;
  int use_anderson;
  int use_gonzalez;
  int exit_picard_step;
  int use_logsq;
  int nthreads;
  int solver;          // 0 step_anderson!, 1 step_defect!
  double dc_eta;       // Parameters.jl:89-91
  int dc_max_inner;
  int dc_stag_window;
  int dc_outer_m;      // extension: Anderson window of the outer iteration (0 = plain outer iteration)
  double dc_near_theta;  // extension: pairs with w_a / |d| > theta are exact in the frozen map (0 = none)
  int collision_model;   // 0 :landau, 1 :lb (Parameters.jl:50)
  double nu;             // LB collision frequency (Parameters.jl:53)
  int nk_krylov_max;     // solver 2 (step_newton!): GMRES basis per Newton step (Parameters.jl:92)
  double nk_eta_max;     // Eisenstat-Walker cap (Parameters.jl:93)
  double nk_fd_h;        // absolute finite-difference step (Parameters.jl:98; 0 = 1e-6)
  double nk_fd_rel;      // legacy relative step (> 0 overrides nk_fd_h)
  double cl_theta;       // extension: stiff clusters of pairs with max(w_i, w_j) / |d| > cl_theta (0 = off)
  int cl_kmax;           // largest cluster solved exactly (markers)
  int dc_fallback;       // extension: step_defect! hands a step whose inner solve fails (or whose outer
                         // iteration stagnates) to step_anderson! from the outer iterate
};

struct orc_step_stats {
  int iters;
  int n_restart;
  double residual;
  double r0;
  int exit_reason;  // 0 converged, 1 stagnation, 2 max_iter
  double entropy, energy, p1, p2;
  int inner;        // step_defect!: frozen-metric maps
  int n_clusters;   // stiff clusters solved exactly inside each map
  int cl_fail;      // cluster solves that failed (their rows keep the plain map)
  int handover;     // step_defect! with dc_fallback: outer maps before step_anderson! took over (0 = none)
};

namespace {

struct Loc {
  int i, j;        // 0-based element indices per direction
  double xi1, xi2, h1, h2;
};

// MantisWrappers.jl:168-179 (locate_particle)
bool locate(const orc_mesh* ms, double v1, double v2, Loc& L) {
  const auto& b1 = ms->d1.bp;
  const auto& b2 = ms->d2.bp;
  if (v1 <= b1.front() || v1 >= b1.back() || v2 <= b2.front() || v2 >= b2.back()) return false;
  int i = search_last(b1, v1);
  int j = search_last(b2, v2);
  L.i = i;
  L.j = j;
  L.h1 = b1[i + 1] - b1[i];
  L.h2 = b2[j + 1] - b2[j];
  L.xi1 = (v1 - b1[i]) / L.h1;
  L.xi2 = (v2 - b2[j]) / L.h2;
  return true;
}

// MantisWrappers.jl:252-289 (fast_eval_particle_grad!); loop order j2 outer, j1 inner
int eval_particle(const orc_mesh* ms, const Loc& L, double* vals, double* dx1, double* dx2, int* gids) {
  const int q = ms->p + 1;
  double phi1[16], dphi1[16], phi2[16], dphi2[16];
  local_basis(ms->d1, L.i, L.xi1, phi1, dphi1);
  local_basis(ms->d2, L.j, L.xi2, phi2, dphi2);
  const int s1 = ms->d1.bs[L.i], s2 = ms->d2.bs[L.j];
  int kk = 0;
  for (int j2 = 0; j2 < q; ++j2)
    for (int j1 = 0; j1 < q; ++j1) {
      vals[kk] = phi1[j1] * phi2[j2];
      if (dx1) dx1[kk] = dphi1[j1] * phi2[j2];
      if (dx2) dx2[kk] = phi1[j1] * dphi2[j2];
      gids[kk] = (s1 + j1) + (s2 + j2) * ms->d1.nd;
      kk++;
    }
  return kk;
}

// Field values at the N_QUAD^2 points of element (e1,e2) together with the
// local basis table: the equivalent of Forms.evaluate(field, e, nodes) and
// Forms.evaluate(space, e, nodes) (MantisWrappers.jl:182-188).
void element_eval(const orc_mesh* ms, int e1, int e2, const double* fc, std::vector<double>& fq,
                  std::vector<double>& basis, std::vector<int>& gids) {
  const int q = ms->p + 1, nq = ms->nq;
  const int nloc = q * q, npt = nq * nq;
  fq.assign(npt, 0.0);
  basis.assign((size_t)npt * nloc, 0.0);
  gids.resize(nloc);
  const int s1 = ms->d1.bs[e1], s2 = ms->d2.bs[e2];
  for (int j2 = 0; j2 < q; ++j2)
    for (int j1 = 0; j1 < q; ++j1) gids[j1 + q * j2] = (s1 + j1) + (s2 + j2) * ms->d1.nd;
  std::vector<double> p1(q), p2(q);
  for (int b = 0; b < nq; ++b)
    for (int a = 0; a < nq; ++a) {
      const int pt = a + nq * b;
      local_basis(ms->d1, e1, ms->qx[a], p1.data(), nullptr);
      local_basis(ms->d2, e2, ms->qx[b], p2.data(), nullptr);
      double f = 0.0;
      for (int j2 = 0; j2 < q; ++j2)
        for (int j1 = 0; j1 < q; ++j1) {
          double phi = p1[j1] * p2[j2];
          basis[(size_t)pt * nloc + j1 + q * j2] = phi;
          f += fc[gids[j1 + q * j2]] * phi;
        }
      fq[pt] = f;
    }
}

inline bool in_box(double a, double b, double lo1, double hi1, double lo2, double hi2) {
  return !(a <= lo1 || a >= hi1 || b <= lo2 || b >= hi2);
}

void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

}  // namespace

extern "C" {

void orc_gauss_legendre01(int n, double* x, double* w) { gauss_legendre01(n, x, w); }

// Borden extraction for one direction (exposed for the scipy cross-check).
int orc_extraction_1d(const double* bp, int nb, int p, int k, double* C_out, int* bs_out, int* nd_out) {
  Dir d;
  build_dir(d, bp, nb, p, k);
  std::memcpy(C_out, d.C.data(), sizeof(double) * d.C.size());
  for (int e = 0; e < d.ne; ++e) bs_out[e] = d.bs[e];
  *nd_out = d.nd;
  return 0;
}

orc_mesh* orc_mesh_build(const double* bp1, int nb1, const double* bp2, int nb2, int p_deg, int k_reg,
                         int n_quad) {
  if (nb1 < 2 || nb2 < 2 || p_deg < 1 || p_deg > 8 || k_reg < 0 || k_reg >= p_deg || n_quad < 1) return nullptr;
  auto* ms = new orc_mesh();
  ms->p = p_deg;
  ms->k = k_reg;
  ms->nq = n_quad;
  build_dir(ms->d1, bp1, nb1, p_deg, k_reg);
  build_dir(ms->d2, bp2, nb2, p_deg, k_reg);
  ms->n_dofs = ms->d1.nd * ms->d2.nd;
  ms->n_elem = ms->d1.ne * ms->d2.ne;
  ms->qx.resize(n_quad);
  ms->qw.resize(n_quad);
  gauss_legendre01(n_quad, ms->qx.data(), ms->qw.data());
  // Mass matrix M_ij = \int phi_i phi_j dv with a (p+1)^2 Gauss rule per
  // element (MantisWrappers.jl:107-120), assembled in 2D, dof-by-dof.
  const int q = p_deg + 1;
  const int nd1 = ms->d1.nd;
  std::vector<double> gx(q), gw(q);
  gauss_legendre01(q, gx.data(), gw.data());
  Band& M = ms->M;
  M.n = ms->n_dofs;
  M.bw = p_deg * nd1 + p_deg;
  M.a.assign((size_t)M.n * (M.bw + 1), 0.0);
  const bool keep_dense = ms->n_dofs <= 1600;
  if (keep_dense) ms->Mdense.assign((size_t)ms->n_dofs * ms->n_dofs, 0.0);
  std::vector<double> p1(q), p2(q), vals(q * q);
  std::vector<int> gid(q * q);
  for (int e2 = 0; e2 < ms->d2.ne; ++e2)
    for (int e1 = 0; e1 < ms->d1.ne; ++e1) {
      const double jac = (ms->d1.bp[e1 + 1] - ms->d1.bp[e1]) * (ms->d2.bp[e2 + 1] - ms->d2.bp[e2]);
      const int s1 = ms->d1.bs[e1], s2 = ms->d2.bs[e2];
      for (int b = 0; b < q; ++b)
        for (int a = 0; a < q; ++a) {
          local_basis(ms->d1, e1, gx[a], p1.data(), nullptr);
          local_basis(ms->d2, e2, gx[b], p2.data(), nullptr);
          const double wq = gw[a] * gw[b] * jac;
          for (int j2 = 0; j2 < q; ++j2)
            for (int j1 = 0; j1 < q; ++j1) {
              vals[j1 + q * j2] = p1[j1] * p2[j2];
              gid[j1 + q * j2] = (s1 + j1) + (s2 + j2) * nd1;
            }
          for (int x = 0; x < q * q; ++x)
            for (int y = 0; y < q * q; ++y) {
              const double c = vals[x] * vals[y] * wq;
              if (keep_dense) ms->Mdense[(size_t)gid[x] * ms->n_dofs + gid[y]] += c;
              if (gid[x] >= gid[y]) M.at(gid[x], gid[y]) += c;
            }
        }
    }
  if (!band_cholesky(M)) {
    delete ms;
    return nullptr;
  }
  return ms;
}

void orc_mesh_free(orc_mesh* ms) { delete ms; }

// info[0..7] = n_dofs, n_elem, nd1, nd2, ne1, ne2, p, n_quad
void orc_mesh_info(const orc_mesh* ms, int* info) {
  info[0] = ms->n_dofs;
  info[1] = ms->n_elem;
  info[2] = ms->d1.nd;
  info[3] = ms->d2.nd;
  info[4] = ms->d1.ne;
  info[5] = ms->d2.ne;
  info[6] = ms->p;
  info[7] = ms->nq;
}

// Extraction tables in the layout the reference hands its GPU kernels
// (projection_gpu.jl:191-206): C[i,j,cell] column-major, 1-based basis starts.
void orc_mesh_tables(const orc_mesh* ms, double* C1, double* C2, int* bs1, int* bs2) {
  std::memcpy(C1, ms->d1.C.data(), sizeof(double) * ms->d1.C.size());
  std::memcpy(C2, ms->d2.C.data(), sizeof(double) * ms->d2.C.size());
  for (int e = 0; e < ms->d1.ne; ++e) bs1[e] = ms->d1.bs[e] + 1;
  for (int e = 0; e < ms->d2.ne; ++e) bs2[e] = ms->d2.bs[e] + 1;
}

// Dense assembled mass matrix (only for n_dofs <= 1600), row-major.
int orc_mass_dense(const orc_mesh* ms, double* M) {
  if (ms->Mdense.empty()) return -1;
  std::memcpy(M, ms->Mdense.data(), sizeof(double) * ms->Mdense.size());
  return 0;
}

// M^{-1} b through the factored assembled matrix (the reference's `ws.M_lu \ b`,
// functions.jl:132, solver.jl:67, main.jl:241).
void orc_mass_solve(const orc_mesh* ms, const double* b, double* x) { band_solve(ms->M, b, x); }

// Particle-side basis values and gradients (for tests).
// Returns number of local functions, or 0 when outside the open domain.
int orc_eval_particle(const orc_mesh* ms, double v1, double v2, double* vals, double* dx1, double* dx2, int* gids,
                      double* h) {
  Loc L;
  if (!locate(ms, v1, v2, L)) return 0;
  h[0] = L.h1;
  h[1] = L.h2;
  return eval_particle(ms, L, vals, dx1, dx2, gids);
}

// functions.jl:121-131: b_k = sum_alpha w_alpha phi_k(v_alpha) over in-domain markers
void orc_l2_rhs(const orc_mesh* ms, const double* v, const double* w, int64_t N, double* rhs) {
  std::fill(rhs, rhs + ms->n_dofs, 0.0);
  double vals[81], dx1[81], dx2[81];
  int gids[81];
  for (int64_t a = 0; a < N; ++a) {
    Loc L;
    if (!locate(ms, v[a], v[N + a], L)) continue;
    int nl = eval_particle(ms, L, vals, dx1, dx2, gids);
    for (int j = 0; j < nl; ++j) rhs[gids[j]] += w[a] * vals[j];
  }
}

// functions.jl:121-134 (l2_project!)
void orc_l2_project(const orc_mesh* ms, const double* v, const double* w, int64_t N, double* fc) {
  std::vector<double> rhs(ms->n_dofs);
  orc_l2_rhs(ms, v, w, N, rhs.data());
  band_solve(ms->M, rhs.data(), fc);
}

// functions.jl:137-155 (compute_entropy)
double orc_entropy(const orc_mesh* ms, const double* fc, int use_logsq) {
  double S = 0.0;
  std::vector<double> fq, basis;
  std::vector<int> gids;
  for (int e2 = 0; e2 < ms->d2.ne; ++e2)
    for (int e1 = 0; e1 < ms->d1.ne; ++e1) {
      const double jac = (ms->d1.bp[e1 + 1] - ms->d1.bp[e1]) * (ms->d2.bp[e2 + 1] - ms->d2.bp[e2]);
      element_eval(ms, e1, e2, fc, fq, basis, gids);
      for (int b = 0; b < ms->nq; ++b)
        for (int a = 0; a < ms->nq; ++a) {
          const double f = fq[a + ms->nq * b];
          const double wq = ms->qw[a] * ms->qw[b];
          if (use_logsq) {
            if (f * f > 1e-60) S -= f * 0.5 * std::log(f * f) * wq * jac;
          } else if (f > 1e-30) {
            S -= f * std::log(f) * wq * jac;
          }
        }
    }
  return S;
}

// functions.jl:300-313 (compute_negative_part_l1): sum over cells and Gauss
// points of (-f) w_q |e| where f < 0.
double orc_negative_part_l1(const orc_mesh* ms, const double* fc) {
  double neg = 0.0;
  std::vector<double> fq, basis;
  std::vector<int> gids;
  for (int e2 = 0; e2 < ms->d2.ne; ++e2)
    for (int e1 = 0; e1 < ms->d1.ne; ++e1) {
      const double jac = (ms->d1.bp[e1 + 1] - ms->d1.bp[e1]) * (ms->d2.bp[e2 + 1] - ms->d2.bp[e2]);
      element_eval(ms, e1, e2, fc, fq, basis, gids);
      for (int b = 0; b < ms->nq; ++b)
        for (int a = 0; a < ms->nq; ++a) {
          const double f = fq[a + ms->nq * b];
          if (f < 0.0) neg += (-f) * (ms->qw[a] * ms->qw[b]) * jac;
        }
    }
  return neg;
}

// functions.jl:320-343 (compute_fs_minus_fp_l2): the cell masses m_e = sum of
// w over the markers locate_particle places in cell e, in marker order; then
// sqrt(sum over cells and Gauss points of (f - m_e / |e|)^2 w_q |e|).
double orc_fs_minus_fp_l2(const orc_mesh* ms, const double* fc, const double* v, const double* w, int64_t N) {
  const int ne1 = ms->d1.ne, ne2 = ms->d2.ne;
  std::vector<double> mass((size_t)ne1 * ne2, 0.0);
  for (int64_t a = 0; a < N; ++a) {
    Loc L;
    if (!locate(ms, v[a], v[N + a], L)) continue;
    mass[(size_t)L.i + (size_t)ne1 * L.j] += w[a];
  }
  double sumsq = 0.0;
  std::vector<double> fq, basis;
  std::vector<int> gids;
  for (int e2 = 0; e2 < ne2; ++e2)
    for (int e1 = 0; e1 < ne1; ++e1) {
      const double jac = (ms->d1.bp[e1 + 1] - ms->d1.bp[e1]) * (ms->d2.bp[e2 + 1] - ms->d2.bp[e2]);
      const double fp = mass[(size_t)e1 + (size_t)ne1 * e2] / jac;
      element_eval(ms, e1, e2, fc, fq, basis, gids);
      for (int b = 0; b < ms->nq; ++b)
        for (int a = 0; a < ms->nq; ++a) {
          const double d = fq[a + ms->nq * b] - fp;
          sumsq += d * d * (ms->qw[a] * ms->qw[b]) * jac;
        }
    }
  return std::sqrt(sumsq);
}

// functions.jl:158-179 (compute_r!)
void orc_compute_r(const orc_mesh* ms, const double* fc, int use_logsq, double* r) {
  std::fill(r, r + ms->n_dofs, 0.0);
  const int q = ms->p + 1, nloc = q * q;
  std::vector<double> fq, basis;
  std::vector<int> gids;
  for (int e2 = 0; e2 < ms->d2.ne; ++e2)
    for (int e1 = 0; e1 < ms->d1.ne; ++e1) {
      const double jac = (ms->d1.bp[e1 + 1] - ms->d1.bp[e1]) * (ms->d2.bp[e2 + 1] - ms->d2.bp[e2]);
      element_eval(ms, e1, e2, fc, fq, basis, gids);
      for (int b = 0; b < ms->nq; ++b)
        for (int a = 0; a < ms->nq; ++a) {
          const int pt = a + ms->nq * b;
          const double f = fq[pt];
          double integrand;
          if (use_logsq)
            integrand = f * f > 1e-60 ? (1.0 + 0.5 * std::log(f * f)) : 0.0;
          else
            integrand = f > 1e-30 ? (1.0 + std::log(f)) : 0.0;
          const double wq = ms->qw[a] * ms->qw[b];
          for (int j = 0; j < nloc; ++j) r[gids[j]] += integrand * basis[(size_t)pt * nloc + j] * wq * jac;
        }
    }
}

// functions.jl:182-202 (compute_G!)
void orc_compute_G(const orc_mesh* ms, const double* v, int64_t N, const double* Lv, double* G) {
  double vals[81], dx1[81], dx2[81];
  int gids[81];
  for (int64_t a = 0; a < N; ++a) {
    G[a] = 0.0;
    G[N + a] = 0.0;
    Loc L;
    if (!locate(ms, v[a], v[N + a], L)) continue;
    int nl = eval_particle(ms, L, vals, dx1, dx2, gids);
    const double ih1 = 1.0 / L.h1, ih2 = 1.0 / L.h2;
    double acc1 = 0.0, acc2 = 0.0;
    for (int j = 0; j < nl; ++j) {
      const double c = Lv[gids[j]];
      acc1 += c * dx1[j] * ih1;
      acc2 += c * dx2[j] * ih2;
    }
    G[a] = acc1;
    G[N + a] = acc2;
  }
}

// functions.jl:37-64 (eval_loggrad_at_particles!)
void orc_loggrad(const orc_mesh* ms, const double* v, int64_t N, const double* fc, double fs_floor, double g_max,
                 double* g) {
  double vals[81], dx1[81], dx2[81];
  int gids[81];
  for (int64_t a = 0; a < N; ++a) {
    Loc L;
    if (!locate(ms, v[a], v[N + a], L)) {
      g[a] = 0.0;
      g[N + a] = 0.0;
      continue;
    }
    int nl = eval_particle(ms, L, vals, dx1, dx2, gids);
    const double ih1 = 1.0 / L.h1, ih2 = 1.0 / L.h2;
    double f = 0.0, d1 = 0.0, d2 = 0.0;
    for (int j = 0; j < nl; ++j) {
      const double c = fc[gids[j]];
      f += c * vals[j];
      d1 += c * dx1[j] * ih1;
      d2 += c * dx2[j] * ih2;
    }
    const double invf = 1.0 / std::max(std::fabs(f), fs_floor);
    g[a] = std::min(std::max(d1 * invf, -g_max), g_max);
    g[N + a] = std::min(std::max(d2 * invf, -g_max), g_max);
  }
}

// functions.jl:205-237 (compute_collision!).  box = {lo1, hi1, lo2, hi2}.
void orc_collision(const double* box, const double* v, const double* w, const double* G, int64_t N, double* dot_v,
                   int nthreads) {
  const double lo1 = box[0], hi1 = box[1], lo2 = box[2], hi2 = box[3];
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t g = 0; g < N; ++g) {
    const double vg1 = v[g], vg2 = v[N + g];
    dot_v[g] = 0.0;
    dot_v[N + g] = 0.0;
    if (!in_box(vg1, vg2, lo1, hi1, lo2, hi2)) continue;
    const double Gg1 = G[g], Gg2 = G[N + g];
    double acc1 = 0.0, acc2 = 0.0;
    for (int64_t a = 0; a < N; ++a) {
      if (a == g) continue;
      const double va1 = v[a], va2 = v[N + a];
      if (!in_box(va1, va2, lo1, hi1, lo2, hi2)) continue;
      const double d1 = vg1 - va1;
      const double d2 = vg2 - va2;
      const double dist2 = d1 * d1 + d2 * d2;
      if (dist2 < 1e-24) continue;
      const double dist = std::sqrt(dist2);
      const double g1 = G[a] - Gg1;
      const double g2 = G[N + a] - Gg2;
      const double dvg = (d1 * g1 + d2 * g2) / dist2;
      const double inv = 1.0 / dist;
      acc1 += w[a] * (g1 - d1 * dvg) * inv;
      acc2 += w[a] * (g2 - d2 * dvg) * inv;
    }
    dot_v[g] = acc1;
    dot_v[N + g] = acc2;
  }
}

// Same sum restricted to the targets gamma in [g0, g1) (the CPU baseline times a
// bounded sample of targets; the result rows outside the range are untouched).
void orc_collision_rows(const double* box, const double* v, const double* w, const double* G, int64_t N, int64_t g0,
                        int64_t g1, double* dot_v, int nthreads) {
  const double lo1 = box[0], hi1 = box[1], lo2 = box[2], hi2 = box[3];
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t g = g0; g < g1; ++g) {
    const double vg1 = v[g], vg2 = v[N + g];
    dot_v[g] = 0.0;
    dot_v[N + g] = 0.0;
    if (!in_box(vg1, vg2, lo1, hi1, lo2, hi2)) continue;
    const double Gg1 = G[g], Gg2 = G[N + g];
    double acc1 = 0.0, acc2 = 0.0;
    for (int64_t a = 0; a < N; ++a) {
      if (a == g) continue;
      const double va1 = v[a], va2 = v[N + a];
      if (!in_box(va1, va2, lo1, hi1, lo2, hi2)) continue;
      const double d1 = vg1 - va1;
      const double d2 = vg2 - va2;
      const double dist2 = d1 * d1 + d2 * d2;
      if (dist2 < 1e-24) continue;
      const double dist = std::sqrt(dist2);
      const double g1 = G[a] - Gg1;
      const double g2 = G[N + a] - Gg2;
      const double dvg = (d1 * g1 + d2 * g2) / dist2;
      const double inv = 1.0 / dist;
      acc1 += w[a] * (g1 - d1 * dvg) * inv;
      acc2 += w[a] * (g2 - d2 * dvg) * inv;
    }
    dot_v[g] = acc1;
    dot_v[N + g] = acc2;
  }
}

// functions.jl:249-290 (compute_collision_metric!); A is N x 3 (A11, A12, A22).
void orc_collision_metric(const double* box, const double* v, const double* w, const double* G, int64_t N,
                          double* dot_v, double* A, int nthreads) {
  const double lo1 = box[0], hi1 = box[1], lo2 = box[2], hi2 = box[3];
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t g = 0; g < N; ++g) {
    const double vg1 = v[g], vg2 = v[N + g];
    dot_v[g] = dot_v[N + g] = 0.0;
    A[g] = A[N + g] = A[2 * N + g] = 0.0;
    if (!in_box(vg1, vg2, lo1, hi1, lo2, hi2)) continue;
    const double Gg1 = G[g], Gg2 = G[N + g];
    double acc1 = 0.0, acc2 = 0.0, a11 = 0.0, a12 = 0.0, a22 = 0.0;
    for (int64_t a = 0; a < N; ++a) {
      if (a == g) continue;
      const double va1 = v[a], va2 = v[N + a];
      if (!in_box(va1, va2, lo1, hi1, lo2, hi2)) continue;
      const double d1 = vg1 - va1;
      const double d2 = vg2 - va2;
      const double dist2 = d1 * d1 + d2 * d2;
      if (dist2 < 1e-24) continue;
      const double dist = std::sqrt(dist2);
      const double g1 = G[a] - Gg1;
      const double g2 = G[N + a] - Gg2;
      const double dvg = (d1 * g1 + d2 * g2) / dist2;
      const double inv = 1.0 / dist;
      acc1 += w[a] * (g1 - d1 * dvg) * inv;
      acc2 += w[a] * (g2 - d2 * dvg) * inv;
      const double s = w[a] * inv;
      a11 += s * (1 - d1 * d1 / dist2);
      a12 -= s * d1 * d2 / dist2;
      a22 += s * (1 - d2 * d2 / dist2);
    }
    dot_v[g] = acc1;
    dot_v[N + g] = acc2;
    A[g] = a11;
    A[N + g] = a12;
    A[2 * N + g] = a22;
  }
}

// functions.jl:349-358
void orc_momentum_energy(const double* v, const double* w, int64_t N, double* out3) {
  double p1 = 0.0, p2 = 0.0, e = 0.0;
  for (int64_t a = 0; a < N; ++a) {
    p1 += w[a] * v[a];
    p2 += w[a] * v[N + a];
    e += w[a] * (v[a] * v[a] + v[N + a] * v[N + a]);
  }
  out3[0] = p1;
  out3[1] = p2;
  out3[2] = 0.5 * e;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Implicit step (solver.jl)
namespace {

struct Scratch {
  std::vector<double> vmid, dv, dS, Geff, dotv, fbuf, r, L, Gbuf;
  explicit Scratch(int64_t N, int nd)
      : vmid(2 * N), dv(2 * N), dS(2 * N), Geff(2 * N), dotv(2 * N), fbuf(nd), r(nd), L(nd), Gbuf(2 * N) {}
};

void box_of(const orc_mesh* ms, double* box) {
  box[0] = ms->d1.bp.front();
  box[1] = ms->d1.bp.back();
  box[2] = ms->d2.bp.front();
  box[3] = ms->d2.bp.back();
}

// solver.jl:62-74 (compute_entropy_gradient!)
void entropy_gradient(const orc_mesh* ms, double* dS, const double* v, const double* w, int64_t N, Scratch& s,
                      int use_logsq) {
  orc_l2_project(ms, v, w, N, s.fbuf.data());
  orc_compute_r(ms, s.fbuf.data(), use_logsq, s.r.data());
  orc_mass_solve(ms, s.r.data(), s.L.data());
  orc_compute_G(ms, v, N, s.L.data(), s.Gbuf.data());
  for (int64_t a = 0; a < N; ++a) {
    dS[a] = -w[a] * s.Gbuf[a];
    dS[N + a] = -w[a] * s.Gbuf[N + a];
  }
}

// solver.jl:86-113 (landau_geff!); returns lambda
double landau_geff(const orc_mesh* ms, double* Geff, const double* vin, const double* v0, const double* w, int64_t N,
                   double S0, Scratch& s, int use_gonzalez, int use_logsq) {
  for (int64_t i = 0; i < 2 * N; ++i) {
    s.vmid[i] = 0.5 * (v0[i] + vin[i]);
    s.dv[i] = vin[i] - v0[i];
  }
  entropy_gradient(ms, s.dS.data(), s.vmid.data(), w, N, s, use_logsq);
  double corr = 0.0;
  if (use_gonzalez) {
    orc_l2_project(ms, vin, w, N, s.fbuf.data());
    const double S1 = orc_entropy(ms, s.fbuf.data(), use_logsq);
    double dot = 0.0, nrm2 = 0.0;
    for (int64_t a = 0; a < N; ++a) {
      dot += s.dv[a] * s.dS[a] + s.dv[N + a] * s.dS[N + a];
      nrm2 += s.dv[a] * s.dv[a] + s.dv[N + a] * s.dv[N + a];
    }
    corr = nrm2 > 1e-30 ? (S1 - S0 - dot) / nrm2 : 0.0;
  }
  for (int64_t a = 0; a < N; ++a) {
    const double iw = 1.0 / w[a];
    Geff[a] = -(s.dS[a] + corr * s.dv[a]) * iw;
    Geff[N + a] = -(s.dS[N + a] + corr * s.dv[N + a]) * iw;
  }
  return corr;
}

// ---- Lenard-Bernstein (functions.jl:66-115) ----------------------------------

// functions.jl:66-82 (compute_moments): n, U1, U2, Q in marker order
void lb_moments(const double* v, const double* w, int64_t N, double* m4) {
  double n = 0.0, U1 = 0.0, U2 = 0.0, Q = 0.0;
  for (int64_t a = 0; a < N; ++a) {
    const double wa = w[a], x = v[a], y = v[N + a];
    n += wa;
    U1 += wa * x;
    U2 += wa * y;
    Q += wa * (x * x + y * y);
  }
  m4[0] = n;
  m4[1] = U1;
  m4[2] = U2;
  m4[3] = Q;
}

// functions.jl:84-106 (compute_drift_multipliers): Sg, P in marker order, then
// Mmat \ rhs, which for a square dense matrix is LU with partial pivoting.
void lb_multipliers(const double* v, const double* w, const double* g, int64_t N, const double* m4, double* AB) {
  double Sg1 = 0.0, Sg2 = 0.0, Pd = 0.0;
  for (int64_t a = 0; a < N; ++a) {
    const double g1 = g[a], g2 = g[N + a], wa = w[a];
    Sg1 += wa * g1;
    Sg2 += wa * g2;
    Pd += wa * (v[a] * g1 + v[N + a] * g2);
  }
  double M[3][3] = {{m4[0], 0.0, m4[1]}, {0.0, m4[0], m4[2]}, {m4[1], m4[2], m4[3]}};
  double r[3] = {-Sg1, -Sg2, -Pd};
  for (int k = 0; k < 3; ++k) {
    int pr = k;
    for (int i = k + 1; i < 3; ++i)
      if (std::fabs(M[i][k]) > std::fabs(M[pr][k])) pr = i;
    if (pr != k) {
      for (int j = 0; j < 3; ++j) std::swap(M[k][j], M[pr][j]);
      std::swap(r[k], r[pr]);
    }
    for (int i = k + 1; i < 3; ++i) {
      const double f = M[i][k] / M[k][k];
      for (int j = k; j < 3; ++j) M[i][j] -= f * M[k][j];
      r[i] -= f * r[k];
    }
  }
  for (int i = 2; i >= 0; --i) {
    double t = r[i];
    for (int j = i + 1; j < 3; ++j) t -= M[i][j] * AB[j];
    AB[i] = t / M[i][i];
  }
}

// functions.jl:108-115 (compute_LB_velocity!)
void lb_velocity(double* dot_v, const double* v, const double* g, int64_t N, const double* AB, double nu) {
  for (int64_t a = 0; a < N; ++a) {
    dot_v[a] = -nu * (g[a] + AB[0] + AB[2] * v[a]);
    dot_v[N + a] = -nu * (g[N + a] + AB[1] + AB[2] * v[N + a]);
  }
}

// The right-hand side at v with gradient g: moments, multipliers, velocity.
void lb_rhs(double* dot_v, const double* v, const double* w, const double* g, int64_t N, double nu, double* AB) {
  double m4[4];
  lb_moments(v, w, N, m4);
  lb_multipliers(v, w, g, N, m4, AB);
  lb_velocity(dot_v, v, g, N, AB, nu);
}

// solver.jl:185-210 (picard_map!, LB branch): G_eff = clamped grad f / f of the
// projection of v_mid (LOGGRAD_FN), with the Gonzalez correction
// G_eff -= lambda dv / w, lambda = (S(v_in) - S0 + sum w dv.G_eff) / |dv|^2.
void lb_geff(const orc_mesh* ms, double* Geff, const double* vin, const double* v0, const double* w, int64_t N,
             double S0, Scratch& s, int use_gonzalez, int use_logsq) {
  for (int64_t i = 0; i < 2 * N; ++i) {
    s.vmid[i] = 0.5 * (v0[i] + vin[i]);
    s.dv[i] = vin[i] - v0[i];
  }
  orc_l2_project(ms, s.vmid.data(), w, N, s.fbuf.data());
  orc_loggrad(ms, s.vmid.data(), N, s.fbuf.data(), 1e-30, 100.0, Geff);
  if (!use_gonzalez) return;
  orc_l2_project(ms, vin, w, N, s.fbuf.data());
  const double S1 = orc_entropy(ms, s.fbuf.data(), use_logsq);
  double dot = 0.0, nrm2 = 0.0;
  for (int64_t a = 0; a < N; ++a) {
    dot -= w[a] * (s.dv[a] * Geff[a] + s.dv[N + a] * Geff[N + a]);
    nrm2 += s.dv[a] * s.dv[a] + s.dv[N + a] * s.dv[N + a];
  }
  const double lam = nrm2 > 1e-30 ? (S1 - S0 - dot) / nrm2 : 0.0;
  for (int64_t a = 0; a < N; ++a) {
    const double iw = 1.0 / w[a];
    Geff[a] -= lam * s.dv[a] * iw;
    Geff[N + a] -= lam * s.dv[N + a] * iw;
  }
}

// solver.jl:176-227 (picard_map!): the Landau branch, or the LB branch
// (solver.jl:185-210, 218-222) when collision_model = 1
// Returns the Gonzalez lambda of the Landau branch (0 for LB).
double picard_map(const orc_mesh* ms, double* vout, const double* vin, const double* v0, const double* w, int64_t N,
                  double S0, double dt, Scratch& s, const orc_solver_params& P) {
  if (P.collision_model == 1) {
    lb_geff(ms, s.Geff.data(), vin, v0, w, N, S0, s, P.use_gonzalez, P.use_logsq);
    double AB[3];
    lb_rhs(s.dotv.data(), s.vmid.data(), w, s.Geff.data(), N, P.nu, AB);
    for (int64_t i = 0; i < 2 * N; ++i) vout[i] = v0[i] + dt * s.dotv[i];
    return 0.0;
  }
  const double lam = landau_geff(ms, s.Geff.data(), vin, v0, w, N, S0, s, P.use_gonzalez, P.use_logsq);
  double box[4];
  box_of(ms, box);
  orc_collision(box, s.vmid.data(), w, s.Geff.data(), N, s.dotv.data(), P.nthreads);
  for (int64_t i = 0; i < 2 * N; ++i) vout[i] = v0[i] + dt * s.dotv[i];
  return lam;
}

double nrm(const double* x, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += x[i] * x[i];
  return std::sqrt(s);
}

// Solve the small SPD system (Julia's factorize picks Cholesky for the
// exactly-symmetric ATA; partial-pivot LU as a fallback).
void small_solve(int n, std::vector<double> A, std::vector<double> b, double* x) {
  std::vector<double> Lc(A);
  bool ok = true;
  for (int j = 0; j < n && ok; ++j) {
    double d = Lc[j * n + j];
    for (int k = 0; k < j; ++k) d -= Lc[j * n + k] * Lc[j * n + k];
    if (!(d > 0.0)) {
      ok = false;
      break;
    }
    d = std::sqrt(d);
    Lc[j * n + j] = d;
    for (int i = j + 1; i < n; ++i) {
      double t = Lc[i * n + j];
      for (int k = 0; k < j; ++k) t -= Lc[i * n + k] * Lc[j * n + k];
      Lc[i * n + j] = t / d;
    }
  }
  if (ok) {
    std::vector<double> y(n);
    for (int i = 0; i < n; ++i) {
      double t = b[i];
      for (int k = 0; k < i; ++k) t -= Lc[i * n + k] * y[k];
      y[i] = t / Lc[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double t = y[i];
      for (int k = i + 1; k < n; ++k) t -= Lc[k * n + i] * x[k];
      x[i] = t / Lc[i * n + i];
    }
    return;
  }
  std::vector<int> piv(n);
  for (int i = 0; i < n; ++i) piv[i] = i;
  for (int k = 0; k < n; ++k) {
    int pr = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[i * n + k]) > std::fabs(A[pr * n + k])) pr = i;
    if (pr != k) {
      for (int j = 0; j < n; ++j) std::swap(A[k * n + j], A[pr * n + j]);
      std::swap(b[k], b[pr]);
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

}  // namespace

extern "C" {

// solver.jl:86-113; Geff out (N x 2).  Returns the Gonzalez lambda.
double orc_landau_geff(const orc_mesh* ms, const double* vin, const double* v0, const double* w, int64_t N, double S0,
                       int use_gonzalez, int use_logsq, double* Geff, double* vmid_out) {
  Scratch s(N, ms->n_dofs);
  double lam = landau_geff(ms, Geff, vin, v0, w, N, S0, s, use_gonzalez, use_logsq);
  if (vmid_out) std::memcpy(vmid_out, s.vmid.data(), sizeof(double) * 2 * N);
  return lam;
}

// functions.jl:66-115: moments, drift multipliers (A1, A2, B) and the LB
// velocity at v with log-density gradient g.
void orc_lb_rhs(const double* v, const double* w, const double* g, int64_t N, double nu, double* dot_v, double* AB) {
  lb_rhs(dot_v, v, w, g, N, nu, AB);
}

// solver.jl:185-210: the LB G_eff at v_mid = (v0 + vin) / 2
void orc_lb_geff(const orc_mesh* ms, const double* vin, const double* v0, const double* w, int64_t N, double S0,
                 int use_gonzalez, int use_logsq, double* Geff) {
  Scratch s(N, ms->n_dofs);
  lb_geff(ms, Geff, vin, v0, w, N, S0, s, use_gonzalez, use_logsq);
}

// solver.jl:176-227
void orc_picard_map(const orc_mesh* ms, const orc_solver_params* P, const double* vin, const double* v0,
                    const double* w, int64_t N, double S0, double* vout) {
  Scratch s(N, ms->n_dofs);
  picard_map(ms, vout, vin, v0, w, N, S0, P->dt, s, *P);
}

// main.jl:239-245: Euler predictor v1 = v + dt * vdot(v), G = grad(M^{-1} r(f_s)).
// fc is the projection of v (main.jl:106, :298).
void orc_predictor(const orc_mesh* ms, const orc_solver_params* P, const double* v, const double* w, int64_t N,
                   const double* fc, double* v1, double* dot_v_out) {
  std::vector<double> r(ms->n_dofs), L(ms->n_dofs), G(2 * N), dv(2 * N);
  if (P->collision_model == 1) {  // main.jl:233-238: the direct clamped gradient at v
    orc_loggrad(ms, v, N, fc, 1e-30, 100.0, G.data());
    double AB[3];
    lb_rhs(dv.data(), v, w, G.data(), N, P->nu, AB);
    for (int64_t i = 0; i < 2 * N; ++i) v1[i] = v[i] + P->dt * dv[i];
    if (dot_v_out) std::memcpy(dot_v_out, dv.data(), sizeof(double) * 2 * N);
    return;
  }
  orc_compute_r(ms, fc, P->use_logsq, r.data());
  orc_mass_solve(ms, r.data(), L.data());
  orc_compute_G(ms, v, N, L.data(), G.data());
  double box[4];
  box_of(ms, box);
  orc_collision(box, v, w, G.data(), N, dv.data(), P->nthreads);
  for (int64_t i = 0; i < 2 * N; ++i) v1[i] = v[i] + P->dt * dv[i];
  if (dot_v_out) std::memcpy(dot_v_out, dv.data(), sizeof(double) * 2 * N);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Stiff clusters (extension; no counterpart in the reference).
//
// Markers closer than max(w_i, w_j) / theta form clusters.  Inside a cluster
// the pair terms w U(d) (G_a - G_g) turn the midpoint system very stiff (the
// stiffness of a pair grows like dt w |Hess L| / |d|), and Anderson over the
// whole map stalls once there are more such modes than window columns.  Each
// map therefore replaces a cluster's rows by the exact solution of the
// cluster's own midpoint equations, everything outside the cluster held at
// the map's values (nonlinear elimination):
//   m_i = v0_i + dt/2 (ext_i + sum_{j in C} w_j U(m_i - m_j) (g_j - g_i)),
//   g_i = grad L(m_i) - 2 lambda (m_i - v0_i) / w_i,   x_i = 2 m_i - v0_i,
// with ext_i = the map's vdot_i minus the cluster's own terms at the iterate,
// and L, lambda those of the map.  At a fixed point of the modified map the
// cluster rows equal the plain map's, so the root is the implicit system's.
// The cluster equations can have several roots; the one taken is the end of
// the homotopy in the coupling strength t (the cluster's own terms scaled by
// t) from t = 0, where m = v0 + dt/2 ext, to t = 1, each point corrected by
// Newton with the analytic Jacobian; a warm start from the previous map's root
// of the same cluster is tried first.  The weighted centre of m does not move
// along the path (the cluster's own terms carry no momentum).
namespace {

struct Link {
  double q;
  int i, j;
};

// Unordered in-domain pairs i < j with |d|^2 >= 1e-24 and max(w_i, w_j) / |d| > theta,
// ordered by (|d|^2, i, j).
std::vector<Link> find_links(const double* box, const double* v, const double* w, int64_t N, double theta) {
  std::vector<Link> out;
  if (!(theta > 0.0)) return out;
  double wmax = 0.0;
  for (int64_t a = 0; a < N; ++a) wmax = std::max(wmax, w[a]);
  const double qmax = wmax * wmax / (theta * theta);
  // sweep in v1 order: only pairs with |dv1| < sqrt(qmax) can be near
  std::vector<int> ord(N);
  for (int64_t a = 0; a < N; ++a) ord[a] = (int)a;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return v[a] != v[b] ? v[a] < v[b] : a < b; });
  const double r = std::sqrt(qmax);
  for (int64_t x = 0; x < N; ++x) {
    const int a = ord[x];
    if (!in_box(v[a], v[N + a], box[0], box[1], box[2], box[3])) continue;
    for (int64_t y = x + 1; y < N && v[ord[y]] - v[a] < r; ++y) {
      const int b = ord[y];
      if (!in_box(v[b], v[N + b], box[0], box[1], box[2], box[3])) continue;
      const double d1 = v[a] - v[b], d2 = v[N + a] - v[N + b];
      const double q = d1 * d1 + d2 * d2;
      const double wm = std::max(w[a], w[b]);
      if (q >= 1e-24 && theta * theta * q < wm * wm) out.push_back({q, std::min(a, b), std::max(a, b)});
    }
  }
  std::sort(out.begin(), out.end(), [](const Link& x, const Link& y) {
    return x.q != y.q ? x.q < y.q : (x.i != y.i ? x.i < y.i : x.j < y.j);
  });
  return out;
}

// Links joined closest first (union-find) unless the merged cluster would
// exceed kmax markers; clusters of >= 2 markers with members ascending, ordered
// by their first member.  CSR: off[c] .. off[c+1] into idx.
void build_clusters(const std::vector<Link>& links, int kmax, std::vector<int>& off, std::vector<int>& idx) {
  std::vector<int> nodes;
  for (const Link& l : links) {
    nodes.push_back(l.i);
    nodes.push_back(l.j);
  }
  std::sort(nodes.begin(), nodes.end());
  nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
  auto id = [&](int a) { return (int)(std::lower_bound(nodes.begin(), nodes.end(), a) - nodes.begin()); };
  std::vector<int> par(nodes.size()), sz(nodes.size(), 1);
  for (size_t t = 0; t < par.size(); ++t) par[t] = (int)t;
  std::function<int(int)> find = [&](int a) { return par[a] == a ? a : (par[a] = find(par[a])); };
  for (const Link& l : links) {
    int a = find(id(l.i)), b = find(id(l.j));
    if (a == b || sz[a] + sz[b] > kmax) continue;
    if (a > b) std::swap(a, b);
    par[b] = a;
    sz[a] += sz[b];
  }
  std::vector<std::vector<int>> groups(nodes.size());
  for (size_t t = 0; t < nodes.size(); ++t) groups[find((int)t)].push_back(nodes[t]);
  std::vector<std::vector<int>> cl;
  for (auto& g : groups)
    if (g.size() >= 2) cl.push_back(g);
  std::sort(cl.begin(), cl.end(), [](const std::vector<int>& a, const std::vector<int>& b) { return a[0] < b[0]; });
  off.assign(1, 0);
  idx.clear();
  for (auto& g : cl) {
    idx.insert(idx.end(), g.begin(), g.end());
    off.push_back((int)idx.size());
  }
}

// grad L and the Hessian (H11, H12, H22) of L = sum L_k phi_k at (x, y); false outside the box
bool grad_hess(const orc_mesh* ms, const double* L, double x, double y, double* g, double* H) {
  Loc lc;
  if (!locate(ms, x, y, lc)) return false;
  const int q = ms->p + 1;
  double B1[16], dB1[16], B2[16], dB2[16];
  bernstein_eval(ms->p, lc.xi1, B1, dB1);
  bernstein_eval(ms->p, lc.xi2, B2, dB2);
  // second derivatives of the Bernstein polynomials: p (p-1) (B^{p-2}_{i-2} - 2 B^{p-2}_{i-1} + B^{p-2}_i)
  auto d2b = [&](double xi, double* out) {
    const int p = ms->p;
    double Bm[16] = {0};
    for (int i = 0; i <= p - 2; ++i) Bm[i] = binom(p - 2, i) * std::pow(xi, i) * std::pow(1.0 - xi, p - 2 - i);
    for (int i = 0; i <= p; ++i) {
      const double a = (i >= 2 && i - 2 <= p - 2) ? Bm[i - 2] : 0.0;
      const double b = (i >= 1 && i - 1 <= p - 2) ? Bm[i - 1] : 0.0;
      const double c = (i <= p - 2) ? Bm[i] : 0.0;
      out[i] = p * (p - 1) * (a - 2.0 * b + c);
    }
  };
  double ddB1[16], ddB2[16];
  d2b(lc.xi1, ddB1);
  d2b(lc.xi2, ddB2);
  double f1[16], df1[16], ddf1[16], f2[16], df2[16], ddf2[16];
  const double* C1 = &ms->d1.C[(size_t)q * q * lc.i];
  const double* C2 = &ms->d2.C[(size_t)q * q * lc.j];
  for (int j = 0; j < q; ++j) {
    f1[j] = df1[j] = ddf1[j] = f2[j] = df2[j] = ddf2[j] = 0.0;
    for (int i = 0; i < q; ++i) {
      f1[j] += C1[i + q * j] * B1[i];
      df1[j] += C1[i + q * j] * dB1[i];
      ddf1[j] += C1[i + q * j] * ddB1[i];
      f2[j] += C2[i + q * j] * B2[i];
      df2[j] += C2[i + q * j] * dB2[i];
      ddf2[j] += C2[i + q * j] * ddB2[i];
    }
  }
  const int s1 = ms->d1.bs[lc.i], s2 = ms->d2.bs[lc.j];
  double gx = 0, gy = 0, hxx = 0, hxy = 0, hyy = 0;
  for (int j2 = 0; j2 < q; ++j2)
    for (int j1 = 0; j1 < q; ++j1) {
      const double c = L[(s1 + j1) + (s2 + j2) * ms->d1.nd];
      gx += c * df1[j1] * f2[j2];
      gy += c * f1[j1] * df2[j2];
      hxx += c * ddf1[j1] * f2[j2];
      hxy += c * df1[j1] * df2[j2];
      hyy += c * f1[j1] * ddf2[j2];
    }
  g[0] = gx / lc.h1;
  g[1] = gy / lc.h2;
  H[0] = hxx / (lc.h1 * lc.h1);
  H[1] = hxy / (lc.h1 * lc.h2);
  H[2] = hyy / (lc.h2 * lc.h2);
  return true;
}

// The cluster's equations Phi(m; t) (interleaved m[2i], m[2i+1]) and, if J is
// non-null, their Jacobian (n x n row-major, n = 2k).  False if a marker left
// the box or two coincide.
struct ClusterEq {
  const orc_mesh* ms;
  const double* L;
  double lam, dt;
  int k;
  const double* v0;   // 2k interleaved
  const double* w;    // k
  const double* ext;  // 2k interleaved

  bool eval(const double* m, double t, double* F, double* J) const {
    const int n = 2 * k;
    std::vector<double> g(n), H(3 * k);
    for (int i = 0; i < k; ++i) {
      if (!grad_hess(ms, L, m[2 * i], m[2 * i + 1], &g[2 * i], &H[3 * i])) return false;
      g[2 * i] -= 2.0 * lam * (m[2 * i] - v0[2 * i]) / w[i];
      g[2 * i + 1] -= 2.0 * lam * (m[2 * i + 1] - v0[2 * i + 1]) / w[i];
    }
    const double s = t * 0.5 * dt;
    if (J) std::fill(J, J + (size_t)n * n, 0.0);
    for (int i = 0; i < k; ++i) {
      double acc0 = 0.0, acc1 = 0.0;
      for (int j = 0; j < k; ++j) {
        if (j == i) continue;
        const double d0 = m[2 * i] - m[2 * j], d1 = m[2 * i + 1] - m[2 * j + 1];
        const double y0 = g[2 * j] - g[2 * i], y1 = g[2 * j + 1] - g[2 * i + 1];
        const double qd = d0 * d0 + d1 * d1;
        if (!(qd >= 1e-24)) return false;
        const double r = std::sqrt(qd), r3 = qd * r;
        const double dy = d0 * y0 + d1 * y1;
        // U(d) y = y / r - d (d.y) / r^3
        acc0 += w[j] * (y0 / r - d0 * dy / r3);
        acc1 += w[j] * (y1 / r - d1 * dy / r3);
        if (J) {
          // Dd[a][b] = d(U(d) y)_a / d d_b = -(y_a d_b + d_a y_b + (d.y) delta_ab) / r^3 + 3 d_a (d.y) d_b / r^5
          const double dd[2] = {d0, d1}, yy[2] = {y0, y1};
          double Dd[2][2], U[2][2];
          for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
              Dd[a][b] = -(yy[a] * dd[b] + dd[a] * yy[b] + (a == b ? dy : 0.0)) / r3 + 3.0 * dd[a] * dy * dd[b] / (r3 * qd);
              U[a][b] = ((a == b ? 1.0 : 0.0) - dd[a] * dd[b] / qd) / r;
            }
          // B_i = Hess L(m_i) - 2 lambda / w_i I
          const double Bi[2][2] = {{H[3 * i] - 2.0 * lam / w[i], H[3 * i + 1]}, {H[3 * i + 1], H[3 * i + 2] - 2.0 * lam / w[i]}};
          const double Bj[2][2] = {{H[3 * j] - 2.0 * lam / w[j], H[3 * j + 1]}, {H[3 * j + 1], H[3 * j + 2] - 2.0 * lam / w[j]}};
          for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
              const double UBi = U[a][0] * Bi[0][b] + U[a][1] * Bi[1][b];
              const double UBj = U[a][0] * Bj[0][b] + U[a][1] * Bj[1][b];
              // dPhi_i/dm_i -= s w_j (Dd - U B_i);  dPhi_i/dm_j -= s w_j (-Dd + U B_j)
              J[(size_t)(2 * i + a) * n + 2 * i + b] -= s * w[j] * (Dd[a][b] - UBi);
              J[(size_t)(2 * i + a) * n + 2 * j + b] -= s * w[j] * (-Dd[a][b] + UBj);
            }
        }
      }
      F[2 * i] = m[2 * i] - v0[2 * i] - 0.5 * dt * ext[2 * i] - s * acc0;
      F[2 * i + 1] = m[2 * i + 1] - v0[2 * i + 1] - 0.5 * dt * ext[2 * i + 1] - s * acc1;
      if (J) {
        J[(size_t)(2 * i) * n + 2 * i] += 1.0;
        J[(size_t)(2 * i + 1) * n + 2 * i + 1] += 1.0;
      }
    }
    return true;
  }
};

// LU with partial pivoting: solves A x = b in place (A n x n row-major); false if singular
bool lu_solve(int n, std::vector<double>& A, std::vector<double>& b) {
  for (int c = 0; c < n; ++c) {
    int pr = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[(size_t)r * n + c]) > std::fabs(A[(size_t)pr * n + c])) pr = r;
    if (!(std::fabs(A[(size_t)pr * n + c]) > 0.0)) return false;
    if (pr != c) {
      for (int j = 0; j < n; ++j) std::swap(A[(size_t)c * n + j], A[(size_t)pr * n + j]);
      std::swap(b[c], b[pr]);
    }
    const double inv = 1.0 / A[(size_t)c * n + c];
    for (int r = c + 1; r < n; ++r) {
      const double f = A[(size_t)r * n + c] * inv;
      if (f == 0.0) continue;
      for (int j = c; j < n; ++j) A[(size_t)r * n + j] -= f * A[(size_t)c * n + j];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double t = b[r];
    for (int j = r + 1; j < n; ++j) t -= A[(size_t)r * n + j] * b[j];
    b[r] = t / A[(size_t)r * n + r];
  }
  return true;
}

// Newton on Phi(m; t) = 0 from m (in/out): converged when |Phi| < tol_abs.
// A step is shortened so that no two markers move by more than half their
// distance relative to each other (factor a <= 1); it fails on a singular
// Jacobian or a residual that does not fall to (1 - a / 10) of the last one
// (unless already within 1e3 tol_abs).
bool cl_newton(const ClusterEq& E, std::vector<double>& m, double t, int maxit, double tol_abs, int& its) {
  const int k = E.k, n = 2 * k;
  std::vector<double> F(n), Fn(n), J((size_t)n * n), mn(n);
  if (!E.eval(m.data(), t, F.data(), nullptr)) return false;
  double nf = nrm(F.data(), n);
  for (int it = 0; it < maxit; ++it) {
    if (nf < tol_abs) return true;
    if (!E.eval(m.data(), t, F.data(), J.data())) return false;
    std::vector<double> dlt(F);
    for (double& x : dlt) x = -x;
    ++its;
    if (!lu_solve(n, J, dlt)) return false;
    double a = 1.0;
    for (int i = 0; i < k; ++i)
      for (int j = i + 1; j < k; ++j) {
        const double r0 = dlt[2 * i] - dlt[2 * j], r1 = dlt[2 * i + 1] - dlt[2 * j + 1];
        const double d0 = m[2 * i] - m[2 * j], d1 = m[2 * i + 1] - m[2 * j + 1];
        const double rr = r0 * r0 + r1 * r1, dd = d0 * d0 + d1 * d1;
        if (rr > 0.25 * dd) a = std::min(a, 0.5 * std::sqrt(dd / rr));
      }
    for (int c = 0; c < n; ++c) mn[c] = m[c] + a * dlt[c];
    if (!E.eval(mn.data(), t, Fn.data(), nullptr)) return false;
    const double nn = nrm(Fn.data(), n);
    if (!(nn <= (1.0 - 0.1 * a) * nf) && nf > 1e3 * tol_abs) return false;
    m = mn;
    nf = nn;
  }
  return nf < tol_abs;
}

// A two-marker cluster reduces to one unknown.  With W = w_g + w_a, the
// weighted centre c of (m_g, m_a) is explicit and d = m_g - m_a solves
// d = a + dt/2 W U(d) (g_a - g_g), a = v0_g - v0_a + dt/2 (ext_g - ext_a).
// U(d) y is orthogonal to d, so d . (d - a) = 0: d lies on the circle over a,
// d = |a| cos(psi) e(phi_a + psi), |psi| < pi/2, and the remaining equation is
//   F(psi) = -|a|^2 / 2 sin(2 psi) + dt/2 W e_perp . (g_a - g_g) = 0.
// At psi -> +-pi/2 (d -> 0) F tends to +-dt W lambda (a . (v0_g - v0_a)) / (|a|
// w) for equal weights, so a root exists whenever that is non-zero.  The root
// taken is the one nearest psi_ref (0, the uncoupled solution d = a, or the
// previous map's root): sign changes on a fixed grid of 64 points, then
// bisection.  Returns false if the grid shows no sign change.
struct PairEq {
  const ClusterEq& E;
  double c[2], a[2], na, pa;
  explicit PairEq(const ClusterEq& e) : E(e) {
    const double* v0 = E.v0;
    const double* x = E.ext;
    const double wg = E.w[0], wa = E.w[1], W = wg + wa;
    for (int q = 0; q < 2; ++q) {
      c[q] = (wg * v0[q] + wa * v0[2 + q]) / W + 0.5 * E.dt * (wg * x[q] + wa * x[2 + q]) / W;
      a[q] = v0[q] - v0[2 + q] + 0.5 * E.dt * (x[q] - x[2 + q]);
    }
    na = std::sqrt(a[0] * a[0] + a[1] * a[1]);
    pa = std::atan2(a[1], a[0]);
  }
  // midpoints of psi; false outside the box
  void mid(double psi, double* m) const {
    const double wg = E.w[0], wa = E.w[1], W = wg + wa;
    const double ph = pa + psi, rho = na * std::cos(psi);
    const double d0 = rho * std::cos(ph), d1 = rho * std::sin(ph);
    m[0] = c[0] + (wa / W) * d0;
    m[1] = c[1] + (wa / W) * d1;
    m[2] = c[0] - (wg / W) * d0;
    m[3] = c[1] - (wg / W) * d1;
  }
  bool F(double psi, double& f) const {
    double m[4], g[4], H[3];
    mid(psi, m);
    for (int i = 0; i < 2; ++i) {
      if (!grad_hess(E.ms, E.L, m[2 * i], m[2 * i + 1], &g[2 * i], H)) return false;
      g[2 * i] -= 2.0 * E.lam * (m[2 * i] - E.v0[2 * i]) / E.w[i];
      g[2 * i + 1] -= 2.0 * E.lam * (m[2 * i + 1] - E.v0[2 * i + 1]) / E.w[i];
    }
    const double ph = pa + psi;
    const double ep0 = -std::sin(ph), ep1 = std::cos(ph);
    f = -0.5 * na * na * std::sin(2.0 * psi) + 0.5 * E.dt * (E.w[0] + E.w[1]) * (ep0 * (g[2] - g[0]) + ep1 * (g[3] - g[1]));
    return true;
  }
};

constexpr int kPairGrid = 64;
inline double pair_grid(int j) { return -0.5 * M_PI + M_PI * (j + 0.5) / kPairGrid; }

bool pair_root(const ClusterEq& E, double psi_ref, double& psi_out, std::vector<double>& m) {
  PairEq P(E);
  if (!(P.na > 0.0)) return false;
  double f[kPairGrid];
  for (int j = 0; j < kPairGrid; ++j)
    if (!P.F(pair_grid(j), f[j])) return false;
  int best = -1;
  double bd = 1e300;
  for (int j = 0; j + 1 < kPairGrid; ++j) {
    if ((f[j] < 0.0) == (f[j + 1] < 0.0) && f[j] != 0.0) continue;
    const double dd = std::fabs(0.5 * (pair_grid(j) + pair_grid(j + 1)) - psi_ref);
    if (dd < bd) {
      bd = dd;
      best = j;
    }
  }
  if (best < 0) return false;
  double lo = pair_grid(best), hi = pair_grid(best + 1), flo = f[best];
  for (int it = 0; it < 200 && hi - lo > 1e-15; ++it) {
    const double md = 0.5 * (lo + hi);
    double fm;
    if (!P.F(md, fm)) return false;
    if ((fm < 0.0) == (flo < 0.0) && fm != 0.0) {
      lo = md;
      flo = fm;
    } else {
      hi = md;
    }
  }
  psi_out = 0.5 * (lo + hi);
  m.resize(4);
  P.mid(psi_out, m.data());
  return true;
}

// The cluster's root (midpoints m, interleaved): Newton from mprev if given,
// else (or if that fails) the homotopy from t = 0.  Returns false if no root
// was reached within the budget; its counts Newton iterations.
bool cluster_root(const ClusterEq& E, const double* mprev, std::vector<double>& m, int& its) {
  const int n = 2 * E.k;
  if (E.k == 2) {  // the one-unknown form; the reference angle is the previous root's
    double psi_ref = 0.0, psi = 0.0;
    its = 0;
    if (mprev) {
      PairEq P(E);
      const double d0 = mprev[0] - mprev[2], d1 = mprev[1] - mprev[3];
      double r = std::atan2(d1, d0) - P.pa;
      while (r > M_PI) r -= 2.0 * M_PI;
      while (r < -M_PI) r += 2.0 * M_PI;
      if (std::fabs(r) < 0.5 * M_PI) psi_ref = r;
    }
    return pair_root(E, psi_ref, psi, m);
  }
  double scale = 1.0;
  for (int c = 0; c < n; ++c) scale = std::max(scale, std::fabs(E.v0[c]) + 1.0);
  const double tol_final = 1e-13 * scale, tol_path = 1e-12 * scale;
  its = 0;
  if (mprev) {
    m.assign(mprev, mprev + n);
    if (cl_newton(E, m, 1.0, 10, tol_final, its)) return true;
  }
  // t = 0: the cluster drifts with ext alone
  m.resize(n);
  for (int c = 0; c < n; ++c) m[c] = E.v0[c] + 0.5 * E.dt * E.ext[c];
  const std::vector<double> m_t0(m);
  std::vector<double> mp(m), guess(n);
  double t = 0.0, tp = 0.0, dtau = 0.05;
  bool have_prev = false, path_ok = true;
  while (t < 1.0) {
    if (its > 60) {
      path_ok = false;
      break;
    }
    const double t1 = std::min(1.0, t + dtau);
    for (int c = 0; c < n; ++c) guess[c] = have_prev ? m[c] + (m[c] - mp[c]) * (t1 - t) / (t - tp) : m[c];
    if (cl_newton(E, guess, t1, 8, tol_path, its)) {
      mp = m;
      tp = t;
      m = guess;
      t = t1;
      have_prev = true;
      dtau = std::min(1.5 * dtau, 0.5);
    } else {
      dtau *= 0.5;
      if (dtau < 1e-7) {
        path_ok = false;
        break;
      }
    }
  }
  if (path_ok && cl_newton(E, m, 1.0, 20, tol_final, its)) return true;
  // The path folded back before t = 1.  In the stiff limit the cluster's own
  // dynamics lines its markers up along an eigenvector of Hess L (every pair
  // term vanishes there), so Newton starts from the t = 0 configuration
  // projected on the line through its centre along each eigenvector in turn.
  double W = 0.0, c0 = 0.0, c1 = 0.0;
  for (int i = 0; i < E.k; ++i) {
    W += E.w[i];
    c0 += E.w[i] * m_t0[2 * i];
    c1 += E.w[i] * m_t0[2 * i + 1];
  }
  c0 /= W;
  c1 /= W;
  double g[2], H[3];
  if (!grad_hess(E.ms, E.L, c0, c1, g, H)) return false;
  // eigenvectors of [[H0, H1], [H1, H2]], smaller eigenvalue first
  const double th = 0.5 * std::atan2(2.0 * H[1], H[0] - H[2]);
  const double ca = std::cos(th), sa = std::sin(th);
  const double l_a = H[0] * ca * ca + 2.0 * H[1] * ca * sa + H[2] * sa * sa;
  const double l_b = H[0] * sa * sa - 2.0 * H[1] * ca * sa + H[2] * ca * ca;
  double e[2][2] = {{ca, sa}, {-sa, ca}};
  if (l_b < l_a) std::swap(e[0], e[1]);
  for (int q = 0; q < 2; ++q) {
    if (its > 90) return false;
    for (int i = 0; i < E.k; ++i) {
      const double p = (m_t0[2 * i] - c0) * e[q][0] + (m_t0[2 * i + 1] - c1) * e[q][1];
      m[2 * i] = c0 + p * e[q][0];
      m[2 * i + 1] = c1 + p * e[q][1];
    }
    if (cl_newton(E, m, 1.0, 20, tol_final, its)) return true;
  }
  return false;
}

// The cluster rows of a map: for every cluster, ext = vdot minus the cluster's
// own terms at the iterate (v_mid, G_eff of the map), then the root; Ghat's
// cluster rows <- 2 m - v0 (rows of failed clusters keep Gv).  mroot holds each
// cluster's last root (2 * cluster size values per cluster, empty = none).
// nfail counts a cluster's failed solves since its last root: after two it
// keeps the plain rows for the rest of the step.
constexpr int kClusterTries = 2;
int cluster_rows(const orc_mesh* ms, const std::vector<int>& off, const std::vector<int>& idx, const double* v0,
                 const double* w, int64_t N, const Scratch& s, double lam, double dt, double* Ghat,
                 std::vector<std::vector<double>>& mroot, std::vector<int>& nfail) {
  int fails = 0;
  const int ncl = (int)off.size() - 1;
  for (int c = 0; c < ncl; ++c) {
    if (nfail[c] >= kClusterTries) continue;
    const int k = off[c + 1] - off[c];
    const int* id = &idx[off[c]];
    std::vector<double> v0c(2 * k), wc(k), ext(2 * k), m;
    for (int i = 0; i < k; ++i) {
      const int a = id[i];
      v0c[2 * i] = v0[a];
      v0c[2 * i + 1] = v0[N + a];
      wc[i] = w[a];
      double e0 = s.dotv[a], e1 = s.dotv[N + a];
      for (int j = 0; j < k; ++j) {
        if (j == i) continue;
        const int b = id[j];
        const double d0 = s.vmid[a] - s.vmid[b], d1 = s.vmid[N + a] - s.vmid[N + b];
        const double y0 = s.Geff[b] - s.Geff[a], y1 = s.Geff[N + b] - s.Geff[N + a];
        const double qd = d0 * d0 + d1 * d1;
        if (!(qd >= 1e-24)) continue;
        const double dy = (d0 * y0 + d1 * y1) / qd, inv = 1.0 / std::sqrt(qd);
        e0 -= w[b] * (y0 - d0 * dy) * inv;
        e1 -= w[b] * (y1 - d1 * dy) * inv;
      }
      ext[2 * i] = e0;
      ext[2 * i + 1] = e1;
    }
    ClusterEq E{ms, s.L.data(), lam, dt, k, v0c.data(), wc.data(), ext.data()};
    int its = 0;
    const bool ok = cluster_root(E, mroot[c].empty() ? nullptr : mroot[c].data(), m, its);
    if (!ok) {
      fails++;
      nfail[c] = mroot[c].empty() ? nfail[c] + 1 : 1;
      mroot[c].clear();
      continue;
    }
    nfail[c] = 0;
    mroot[c] = m;
    for (int i = 0; i < k; ++i) {
      Ghat[id[i]] = 2.0 * m[2 * i] - v0c[2 * i];
      Ghat[N + id[i]] = 2.0 * m[2 * i + 1] - v0c[2 * i + 1];
    }
  }
  return fails;
}

}  // namespace

extern "C" {

// Stiff clusters of the markers v (N x 2): the CSR (off[ncl + 1], idx) of
// find_links + build_clusters.  Returns ncl, or -1 if the arrays are too small.
int orc_clusters(const orc_mesh* ms, const double* v, const double* w, int64_t N, double theta, int kmax, int* off,
                 int off_cap, int* idx, int idx_cap) {
  double box[4];
  box_of(ms, box);
  std::vector<int> o, ix;
  build_clusters(find_links(box, v, w, N, theta), kmax, o, ix);
  if ((int)o.size() > off_cap || (int)ix.size() > idx_cap) return -1;
  std::copy(o.begin(), o.end(), off);
  std::copy(ix.begin(), ix.end(), idx);
  return (int)o.size() - 1;
}

// One cluster's root for given inputs (interleaved v0, ext; mprev may be null):
// m_out (2k), returns 1 on success, 0 if no root was reached; *its = Newton iterations.
int orc_cluster_root(const orc_mesh* ms, const double* L, double lam, double dt, int k, const double* v0,
                     const double* w, const double* ext, const double* mprev, double* m_out, int* its) {
  ClusterEq E{ms, L, lam, dt, k, v0, w, ext};
  std::vector<double> m;
  int it = 0;
  const bool ok = cluster_root(E, mprev, m, it);
  if (its) *its = it;
  if (ok) std::copy(m.begin(), m.end(), m_out);
  return ok ? 1 : 0;
}

// The cluster equations' residual at m (interleaved) and t; returns 0 if undefined.
int orc_cluster_residual(const orc_mesh* ms, const double* L, double lam, double dt, int k, const double* v0,
                         const double* w, const double* ext, const double* m, double t, double* F, double* J) {
  ClusterEq E{ms, L, lam, dt, k, v0, w, ext};
  return E.eval(m, t, F, J) ? 1 : 0;
}

// grad L and Hess L (H11, H12, H22) at (x, y); returns 0 outside the box
int orc_grad_hess(const orc_mesh* ms, const double* L, double x, double y, double* g, double* H) {
  return grad_hess(ms, L, x, y, g, H) ? 1 : 0;
}

}  // extern "C"

extern "C" {

// solver.jl:300-423 (step_anderson!).  v1 holds the initial guess on entry and
// the solution on exit.  iter_res (optional, length max_iter) receives ||r_k||.
void orc_step_anderson(const orc_mesh* ms, const orc_solver_params* Pp, double* v1, const double* v0, const double* w,
                       int64_t N, double S0, orc_step_stats* st, double* iter_res) {
  const orc_solver_params& P = *Pp;
  const int64_t n2 = 2 * N;
  const int m = P.m_anderson;
  Scratch s(N, ms->n_dofs);
  std::vector<double> Gv(n2), r(n2), rp(n2), Gp(n2), vold(n2), dF(n2 * m), dG(n2 * m), vbest(v1, v1 + n2);
  int history = 0, slot = 0, n_restart = 0;
  double nrm_r0 = 0.0, nrm_r = 0.0, nrm_best = std::numeric_limits<double>::infinity();
  double nrm_best_window = std::numeric_limits<double>::infinity();
  // stiff clusters (extension): found once per step on v0; Anderson mixes the
  // map with the clusters' rows solved exactly (Gh, rh), while the exit, best
  // iterate and stagnation tests use the plain map
  std::vector<int> coff, cidx;
  const bool clusters = P.cl_theta > 0.0 && P.collision_model == 0;
  if (clusters) {
    double box[4];
    box_of(ms, box);
    build_clusters(find_links(box, v0, w, N, P.cl_theta), std::max(2, P.cl_kmax), coff, cidx);
  }
  const int ncl = clusters ? (int)coff.size() - 1 : 0;
  std::vector<std::vector<double>> mroot(ncl);
  std::vector<int> nfail(ncl, 0);
  std::vector<double> Gh(ncl ? n2 : 0), rh(ncl ? n2 : 0);
  int cl_fail = 0;
  for (int k = 1; k <= P.max_iter; ++k) {
    std::memcpy(vold.data(), v1, sizeof(double) * n2);
    const double lam = picard_map(ms, Gv.data(), v1, v0, w, N, S0, P.dt, s, P);
    for (int64_t i = 0; i < n2; ++i) r[i] = Gv[i] - v1[i];
    nrm_r = nrm(r.data(), n2);
    if (ncl) {
      std::memcpy(Gh.data(), Gv.data(), sizeof(double) * n2);
      cl_fail += cluster_rows(ms, coff, cidx, v0, w, N, s, lam, P.dt, Gh.data(), mroot, nfail);
      for (int64_t i = 0; i < n2; ++i) rh[i] = Gh[i] - v1[i];
    }
    // the vectors Anderson mixes
    const std::vector<double>& Gm = ncl ? Gh : Gv;
    const std::vector<double>& rm = ncl ? rh : r;
    if (iter_res) iter_res[k - 1] = nrm_r;
    if (k == 1) nrm_r0 = nrm_r;
    if (nrm_r < nrm_best) {
      nrm_best = nrm_r;
      std::memcpy(vbest.data(), P.exit_picard_step ? Gm.data() : v1, sizeof(double) * n2);
    }
    const double eff_tol = std::max(P.tol * (nrm(v1, n2) + 1e-30), P.abs_floor);
    if (nrm_r < eff_tol) {
      // the exit Picard step; with stiff clusters the map Anderson mixes (its
      // cluster rows are the clusters' exact solutions at the final iterate)
      if (P.exit_picard_step) std::memcpy(v1, Gm.data(), sizeof(double) * n2);
      *st = {k, n_restart, nrm_r, nrm_r0, 0, 0, 0, 0, 0, 0, ncl, cl_fail, 0};
      return;
    }
    if (k > P.stag_window && k % P.stag_window == 0) {
      const double rel = (nrm_best_window - nrm_best) / (nrm_best_window + 1e-30);
      if (rel < P.stag_rel_tol) {
        std::memcpy(v1, vbest.data(), sizeof(double) * n2);
        *st = {k, n_restart, nrm_best, nrm_r0, 1, 0, 0, 0, 0, 0, ncl, cl_fail, 0};
        return;
      }
      nrm_best_window = nrm_best;
    }
    bool just_restarted = false;
    if (k > 1 && nrm_r > P.restart_factor * nrm_best) {
      history = 0;
      slot = 0;
      n_restart++;
      just_restarted = true;
    }
    const double beta = k > P.damp_decay_start ? P.damping * P.damp_decay_factor : P.damping;
    if (k == 1 || just_restarted || !P.use_anderson) {
      for (int64_t i = 0; i < n2; ++i) v1[i] = beta * Gm[i] + (1 - beta) * vold[i];
    } else {
      slot = slot % m + 1;
      history = std::min(history + 1, m);
      double* F = &dF[(size_t)(slot - 1) * n2];
      double* Gc = &dG[(size_t)(slot - 1) * n2];
      for (int64_t i = 0; i < n2; ++i) {
        F[i] = rm[i] - rp[i];
        Gc[i] = Gm[i] - Gp[i];
      }
      const int h = history;
      std::vector<double> ATA(h * h), ATr(h), gam(h);
      for (int a = 0; a < h; ++a) {
        for (int b = 0; b <= a; ++b) {
          double t = 0.0;
          const double* Fa = &dF[(size_t)a * n2];
          const double* Fb = &dF[(size_t)b * n2];
          for (int64_t i = 0; i < n2; ++i) t += Fa[i] * Fb[i];
          ATA[a * h + b] = ATA[b * h + a] = t;
        }
        double t = 0.0;
        const double* Fa = &dF[(size_t)a * n2];
        for (int64_t i = 0; i < n2; ++i) t += Fa[i] * rm[i];
        ATr[a] = t;
      }
      double dm = 0.0;
      for (int j = 0; j < h; ++j) dm += ATA[j * h + j];
      dm /= h;
      const double l2 = P.reg_factor * dm + 1e-30;
      for (int j = 0; j < h; ++j) ATA[j * h + j] += l2;
      small_solve(h, ATA, ATr, gam.data());
      for (int64_t i = 0; i < n2; ++i) {
        double x = Gm[i];
        for (int j = 0; j < h; ++j) x -= dG[(size_t)j * n2 + i] * gam[j];
        v1[i] = beta * x + (1 - beta) * vold[i];
      }
    }
    std::memcpy(rp.data(), rm.data(), sizeof(double) * n2);
    std::memcpy(Gp.data(), Gm.data(), sizeof(double) * n2);
  }
  std::memcpy(v1, vbest.data(), sizeof(double) * n2);
  *st = {P.max_iter, n_restart, nrm_best, nrm_r0, 2, 0, 0, 0, 0, 0, ncl, cl_fail, 0};
}

}  // extern "C"

namespace {

// solver.jl:507-545 (_anderson_inner!): Anderson on the fixed point u = Phi(u)
// with the outer solve's window, damping and regularisation and no stagnation
// logic.  Stops at ||Phi(u) - u|| < tol with u <- Phi(u), or after maxit maps
// with u the last update.  Returns the number of maps.
template <class MapF>
int anderson_inner(MapF&& phi, double* u, int64_t n2, double tol, int maxit, int m, double damping, double reg,
                   bool* converged = nullptr) {
  if (converged) *converged = false;
  std::vector<double> Fu(n2), r(n2), rp(n2), Fp(n2), uold(n2), dF((size_t)m * n2), dG((size_t)m * n2);
  int hist = 0, slot = 0;
  for (int k = 1; k <= maxit; ++k) {
    std::memcpy(uold.data(), u, sizeof(double) * n2);
    phi(Fu.data(), u);
    for (int64_t i = 0; i < n2; ++i) r[i] = Fu[i] - u[i];
    const double nr = nrm(r.data(), n2);
    if (nr < tol) {
      std::memcpy(u, Fu.data(), sizeof(double) * n2);
      if (converged) *converged = true;
      return k;
    }
    if (k == 1) {
      for (int64_t i = 0; i < n2; ++i) u[i] = damping * Fu[i] + (1 - damping) * uold[i];
    } else {
      slot = slot % m + 1;
      hist = std::min(hist + 1, m);
      double* F = &dF[(size_t)(slot - 1) * n2];
      double* Gc = &dG[(size_t)(slot - 1) * n2];
      for (int64_t i = 0; i < n2; ++i) {
        F[i] = r[i] - rp[i];
        Gc[i] = Fu[i] - Fp[i];
      }
      std::vector<double> ATA(hist * hist), ATr(hist), gam(hist);
      for (int a = 0; a < hist; ++a) {
        const double* Fa = &dF[(size_t)a * n2];
        for (int b = 0; b <= a; ++b) {
          const double* Fb = &dF[(size_t)b * n2];
          double t = 0.0;
          for (int64_t i = 0; i < n2; ++i) t += Fa[i] * Fb[i];
          ATA[a * hist + b] = ATA[b * hist + a] = t;
        }
        double t = 0.0;
        for (int64_t i = 0; i < n2; ++i) t += Fa[i] * r[i];
        ATr[a] = t;
      }
      double dm = 0.0;
      for (int j = 0; j < hist; ++j) dm += ATA[j * hist + j];
      const double l2 = reg * dm / hist + 1e-30;
      for (int j = 0; j < hist; ++j) ATA[j * hist + j] += l2;
      small_solve(hist, ATA, ATr, gam.data());
      for (int64_t i = 0; i < n2; ++i) {
        double x = Fu[i];
        for (int j = 0; j < hist; ++j) x -= dG[(size_t)j * n2 + i] * gam[j];
        u[i] = damping * x + (1 - damping) * uold[i];
      }
    }
    std::memcpy(rp.data(), r.data(), sizeof(double) * n2);
    std::memcpy(Fp.data(), Fu.data(), sizeof(double) * n2);
  }
  return maxit;
}

}  // namespace

extern "C" {

// solver.jl:569-634 (step_defect!): defect correction around a frozen metric.
// Outer iteration k: the full map G(v_k) with the metric A_k in the same pair
// pass, then Anderson on Phi_k(u) = G(v_k) - dt A_k (G_eff(u) - G_eff(v_k)) to
// max(eta ||G(v_k) - v_k||, abs_floor / 2).  The exit returns G(v) (or the best
// G(v) after stagnation / max_iter).
// Extension (dc_fallback): the rest of a defect-correction step that handed
// over after k outer maps, as step_anderson! from v1 with the iterations left.
static void defect_handover(const orc_mesh* ms, const orc_solver_params& P, double* v1, const double* v0,
                            const double* w, int64_t N, double S0, orc_step_stats* st, int k, double nrm0, int inner) {
  orc_solver_params Pa = P;
  Pa.solver = 0;
  Pa.max_iter = P.max_iter - k;
  orc_step_anderson(ms, &Pa, v1, v0, w, N, S0, st, nullptr);
  st->iters += k;
  st->r0 = nrm0;
  st->inner = inner;
  st->handover = k;
}

void orc_step_defect(const orc_mesh* ms, const orc_solver_params* Pp, double* v1, const double* v0, const double* w,
                     int64_t N, double S0, orc_step_stats* st, double* iter_res) {
  const orc_solver_params& P = *Pp;
  const int64_t n2 = 2 * N;
  Scratch s(N, ms->n_dofs), s_in(N, ms->n_dofs);
  std::vector<double> Gv(n2), r(n2), A(3 * N), Gk(n2), Ge(n2), vbest(v1, v1 + n2);
  double box[4];
  box_of(ms, box);
  double nrm0 = 0.0, nrm_best = std::numeric_limits<double>::infinity();
  double nrm_best_window = std::numeric_limits<double>::infinity(), nr_prev = 0.0;
  int inner = 0;
  // near pairs (gamma, alpha) of the current outer iterate, sorted by gamma, and
  // the midpoints at v_k
  std::vector<int64_t> nI, nJ;
  std::vector<double> vmk(n2);
  const double lo1 = box[0], hi1 = box[1], lo2 = box[2], hi2 = box[3];
  // w_a U(d) g with U(d) g = (g - d (d.g) / |d|^2) / |d|
  auto uterm = [](double wa, double d1, double d2, double g1, double g2, double* o) {
    const double q = d1 * d1 + d2 * d2;
    const double dg = (d1 * g1 + d2 * g2) / q;
    const double inv = 1.0 / std::sqrt(q);
    o[0] = wa * (g1 - d1 * dg) * inv;
    o[1] = wa * (g2 - d2 * dg) * inv;
  };
  auto frozen = [&](double* out, const double* u) {
    landau_geff(ms, Ge.data(), u, v0, w, N, S0, s_in, P.use_gonzalez, P.use_logsq);
    for (int64_t a = 0; a < N; ++a) {
      const double d1 = Ge[a] - Gk[a], d2 = Ge[N + a] - Gk[N + a];
      out[a] = Gv[a] - P.dt * (A[a] * d1 + A[N + a] * d2);
      out[N + a] = Gv[N + a] - P.dt * (A[N + a] * d1 + A[2 * N + a] * d2);
    }
    // near field: the pair's exact term at u replaces its frozen-metric model
    // w U(d_k) (G_k,alpha - G_gamma(u)); both agree at u = v_k, so the fixed
    // point is unchanged.
    for (size_t p = 0; p < nI.size(); ++p) {
      const int64_t g = nI[p], a = nJ[p];
      double ex[2], fr[2];
      uterm(w[a], s_in.vmid[g] - s_in.vmid[a], s_in.vmid[N + g] - s_in.vmid[N + a], Ge[a] - Ge[g],
            Ge[N + a] - Ge[N + g], ex);
      uterm(w[a], vmk[g] - vmk[a], vmk[N + g] - vmk[N + a], Gk[a] - Ge[g], Gk[N + a] - Ge[N + g], fr);
      out[g] += P.dt * (ex[0] - fr[0]);
      out[N + g] += P.dt * (ex[1] - fr[1]);
    }
  };
  // outer Anderson window (extension): x_k, f_k = T(x_k) - x_k
  const int mo = P.dc_outer_m;
  std::vector<double> xk(n2), U(n2), Up(n2), f(n2), fp(n2), dFo((size_t)std::max(mo, 1) * n2),
      dGo((size_t)std::max(mo, 1) * n2);
  int ohist = 0, oslot = 0;
  for (int k = 1; k <= P.max_iter; ++k) {
    // the hand-over needs at least one iteration left for step_anderson!
    const bool fallback = P.dc_fallback && k < P.max_iter;
    landau_geff(ms, s.Geff.data(), v1, v0, w, N, S0, s, P.use_gonzalez, P.use_logsq);
    orc_collision_metric(box, s.vmid.data(), w, s.Geff.data(), N, s.dotv.data(), A.data(), P.nthreads);
    for (int64_t i = 0; i < n2; ++i) {
      Gv[i] = v0[i] + P.dt * s.dotv[i];
      r[i] = Gv[i] - v1[i];
    }
    const double nr = nrm(r.data(), n2);
    if (iter_res) iter_res[k - 1] = nr;
    if (k == 1) nrm0 = nr;
    if (nr < nrm_best) {
      nrm_best = nr;
      std::memcpy(vbest.data(), Gv.data(), sizeof(double) * n2);
    }
    if (nr < std::max(P.tol * (nrm(v1, n2) + 1e-30), P.abs_floor)) {
      std::memcpy(v1, Gv.data(), sizeof(double) * n2);
      *st = {k, 0, nr, nrm0, 0, 0, 0, 0, 0, inner, 0, 0, 0};
      return;
    }
    // hand-over extension: an outer iteration that does not reduce the residual
    // by 0.4 continues the step with step_anderson! from the best iterate's image
    if (fallback && k >= 2 && nr > 0.4 * nr_prev) {
      std::memcpy(v1, vbest.data(), sizeof(double) * n2);
      defect_handover(ms, P, v1, v0, w, N, S0, st, k, nrm0, inner);
      return;
    }
    nr_prev = nr;
    if (k > P.dc_stag_window && k % P.dc_stag_window == 0) {
      if ((nrm_best_window - nrm_best) / (nrm_best_window + 1e-30) < P.stag_rel_tol) {
        std::memcpy(v1, vbest.data(), sizeof(double) * n2);
        if (fallback) {
          defect_handover(ms, P, v1, v0, w, N, S0, st, k, nrm0, inner);
          return;
        }
        *st = {k, 0, nrm_best, nrm0, 1, 0, 0, 0, 0, inner, 0, 0, 0};
        return;
      }
      nrm_best_window = nrm_best;
    }
    std::memcpy(Gk.data(), s.Geff.data(), sizeof(double) * n2);
    std::memcpy(vmk.data(), s.vmid.data(), sizeof(double) * n2);
    if (k == 1 && P.dc_near_theta > 0.0) {  // the step's near pairs: those of its first outer iterate
      const double th2 = P.dc_near_theta * P.dc_near_theta;
      for (int64_t g = 0; g < N; ++g) {
        if (!in_box(vmk[g], vmk[N + g], lo1, hi1, lo2, hi2)) continue;
        for (int64_t a = 0; a < N; ++a) {
          if (a == g || !in_box(vmk[a], vmk[N + a], lo1, hi1, lo2, hi2)) continue;
          const double d1 = vmk[g] - vmk[a], d2 = vmk[N + g] - vmk[N + a];
          const double q = d1 * d1 + d2 * d2;
          if (q >= 1e-24 && th2 * q < w[a] * w[a]) {
            nI.push_back(g);
            nJ.push_back(a);
          }
        }
      }
    }
    std::memcpy(xk.data(), v1, sizeof(double) * n2);
    bool inner_ok = false;
    inner += anderson_inner(frozen, v1, n2, std::max(P.dc_eta * nr, 0.5 * P.abs_floor), P.dc_max_inner,
                            P.m_anderson, P.damping, P.reg_factor, &inner_ok);
    if (!inner_ok && fallback) {  // the Anderson solve continues from x_k
      std::memcpy(v1, xk.data(), sizeof(double) * n2);
      defect_handover(ms, P, v1, v0, w, N, S0, st, k, nrm0, inner);
      return;
    }
    if (mo > 0) {  // Anderson on x -> T(x) (the inner solution), beta = 1
      std::memcpy(U.data(), v1, sizeof(double) * n2);
      for (int64_t i = 0; i < n2; ++i) f[i] = U[i] - xk[i];
      if (k > 1 || ohist > 0) {
        oslot = oslot % mo + 1;
        ohist = std::min(ohist + 1, mo);
        double* F = &dFo[(size_t)(oslot - 1) * n2];
        double* Gc = &dGo[(size_t)(oslot - 1) * n2];
        for (int64_t i = 0; i < n2; ++i) {
          F[i] = f[i] - fp[i];
          Gc[i] = U[i] - Up[i];
        }
        const int h = ohist;
        std::vector<double> ATA(h * h), ATr(h), gam(h);
        for (int a2 = 0; a2 < h; ++a2) {
          const double* Fa = &dFo[(size_t)a2 * n2];
          for (int b2 = 0; b2 <= a2; ++b2) {
            const double* Fb = &dFo[(size_t)b2 * n2];
            double t = 0.0;
            for (int64_t i = 0; i < n2; ++i) t += Fa[i] * Fb[i];
            ATA[a2 * h + b2] = ATA[b2 * h + a2] = t;
          }
          double t = 0.0;
          for (int64_t i = 0; i < n2; ++i) t += Fa[i] * f[i];
          ATr[a2] = t;
        }
        double dm = 0.0;
        for (int j = 0; j < h; ++j) dm += ATA[j * h + j];
        const double l2 = P.reg_factor * dm / h + 1e-30;
        for (int j = 0; j < h; ++j) ATA[j * h + j] += l2;
        small_solve(h, ATA, ATr, gam.data());
        for (int64_t i = 0; i < n2; ++i) {
          double x = U[i];
          for (int j = 0; j < h; ++j) x -= dGo[(size_t)j * n2 + i] * gam[j];
          v1[i] = x;
        }
      }
      std::memcpy(Up.data(), U.data(), sizeof(double) * n2);
      std::memcpy(fp.data(), f.data(), sizeof(double) * n2);
    }
  }
  std::memcpy(v1, vbest.data(), sizeof(double) * n2);
  *st = {P.max_iter, 0, nrm_best, nrm0, 2, 0, 0, 0, 0, inner, 0, 0, 0};
}

// One full time step of main.jl:226-307 (Landau, Anderson): predictor,
// step_anderson!, projection, entropy / energy / momentum of the new state.
}  // extern "C"

// ---------------------------------------------------------------------------
// Matrix-free GMRES and Jacobian-free Newton-Krylov (newton_krylov.jl), generic
// over a residual closure on plain vectors.
namespace {

using Vec = std::vector<double>;
using Resid = std::function<void(Vec& F, const Vec& x)>;

double vdot(const Vec& a, const Vec& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// newton_krylov.jl:37-58 (_sym_givens): [c s; s -c] [a; b] = [rho; 0]
void sym_givens(double a, double b, double& c, double& s, double& rho) {
  auto sgn = [](double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); };
  if (b == 0.0) {
    c = sgn(a) + (a == 0.0 ? 1.0 : 0.0);
    s = 0.0;
    rho = std::fabs(a);
  } else if (a == 0.0) {
    c = 0.0;
    s = sgn(b);
    rho = std::fabs(b);
  } else if (std::fabs(b) > std::fabs(a)) {
    const double t = a / b;
    s = sgn(b) / std::sqrt(1.0 + t * t);
    c = s * t;
    rho = b / s;
  } else {
    const double t = b / a;
    c = sgn(a) / std::sqrt(1.0 + t * t);
    s = c * t;
    rho = a / c;
  }
}

// newton_krylov.jl:60-131 (gmres!): from x0 = 0, unrestarted, modified
// Gram-Schmidt, Givens QR of the Hessenberg matrix; returns the product count.
int gmres(Vec& x, const std::function<void(Vec& y, const Vec& u)>& A, const Vec& b, int mem, double rtol,
          double atol, int itmax, double* rnorm_out) {
  const size_t n = b.size();
  mem = std::min(itmax, mem);
  std::fill(x.begin(), x.end(), 0.0);
  const double beta = std::sqrt(vdot(b, b));
  if (beta == 0.0) {
    *rnorm_out = 0.0;
    return 0;
  }
  const double eps_tol = atol + rtol * beta;
  const double btol = std::pow(2.220446049250313e-16, 0.75);
  std::vector<Vec> V(mem + 1, Vec(n));
  std::vector<double> R((size_t)mem * mem, 0.0), c(mem), sn(mem), z(mem + 1, 0.0);
  Vec w(n);
  z[0] = beta;
  for (size_t i = 0; i < n; ++i) V[0][i] = b[i] / beta;
  int k = 0;
  double rnorm = beta;
  auto Rm = [&](int i, int j) -> double& { return R[(size_t)i * mem + j]; };
  while (k < mem) {
    ++k;
    A(w, V[k - 1]);
    for (int i = 0; i < k; ++i) {
      const double h = vdot(V[i], w);
      Rm(i, k - 1) = h;
      for (size_t j = 0; j < n; ++j) w[j] -= h * V[i][j];
    }
    const double hnext = std::sqrt(vdot(w, w));
    for (int i = 0; i < k - 1; ++i) {
      const double tmp = c[i] * Rm(i, k - 1) + sn[i] * Rm(i + 1, k - 1);
      Rm(i + 1, k - 1) = sn[i] * Rm(i, k - 1) - c[i] * Rm(i + 1, k - 1);
      Rm(i, k - 1) = tmp;
    }
    sym_givens(Rm(k - 1, k - 1), hnext, c[k - 1], sn[k - 1], Rm(k - 1, k - 1));
    z[k] = sn[k - 1] * z[k - 1];
    z[k - 1] = c[k - 1] * z[k - 1];
    rnorm = std::fabs(z[k]);
    if (rnorm <= eps_tol || hnext <= btol) break;
    if (k < mem)
      for (size_t j = 0; j < n; ++j) V[k][j] = w[j] / hnext;
  }
  for (int i = k - 1; i >= 0; --i) {
    double acc = z[i];
    for (int j = i + 1; j < k; ++j) acc -= Rm(i, j) * z[j];
    z[i] = std::fabs(Rm(i, i)) <= btol ? 0.0 : acc / Rm(i, i);
  }
  for (int i = 0; i < k; ++i)
    for (size_t j = 0; j < n; ++j) x[j] += z[i] * V[i][j];
  *rnorm_out = rnorm;
  return k;
}

// newton_krylov.jl:189-303 (newton_krylov!): status 0 converged, 1 stalled,
// 2 budget; F holds F(x) on return.
int newton_krylov(Vec& x, const Resid& Fn, Vec& F, int krylov_max, double tol, int max_evals, double fd_h,
                  double eta_max, int* n_evals_out, double* rnorm_out, int* n_newton_out, double* rnorm0_out) {
  const double eta_gamma = 0.9, armijo = 1e-4;
  const int max_backtrack = 8;
  const size_t n = x.size();
  Vec Ft(n), xt(n), xp(n), Fp(n), d(n), xb(x), Fb(n);
  Fn(F, x);
  int n_evals = 1;
  double rnorm = std::sqrt(vdot(F, F));
  const double rnorm0 = rnorm;
  double rnorm_best = rnorm;
  Fb = F;
  auto jvp = [&](Vec& y, const Vec& u) {
    const double h = fd_h / std::sqrt(vdot(u, u));
    for (size_t i = 0; i < n; ++i) xp[i] = x[i] + h * u[i];
    Fn(Fp, xp);
    ++n_evals;
    for (size_t i = 0; i < n; ++i) y[i] = (Fp[i] - F[i]) / h;
  };
  double eta = eta_max, rnorm_prev = rnorm;
  int n_newton = 0, status = 2;
  while (true) {
    if (rnorm < tol) {
      status = 0;
      break;
    }
    const int budget = max_evals - n_evals;
    if (budget < 2) break;
    if (n_newton > 0) {
      double ew = eta_gamma * (rnorm / rnorm_prev) * (rnorm / rnorm_prev);
      const double safe = eta_gamma * eta * eta;
      if (safe > 0.1) ew = std::max(ew, safe);
      eta = std::min(std::max(ew, 0.5 * tol / rnorm), eta_max);
    }
    double lin_r;
    gmres(d, jvp, F, krylov_max, eta, 0.0, std::min(krylov_max, budget - 1), &lin_r);
    ++n_newton;
    double lam = 1.0, rnorm_t = std::numeric_limits<double>::infinity();
    bool accepted = false;
    for (int bt = 0; bt <= max_backtrack; ++bt) {
      for (size_t i = 0; i < n; ++i) xt[i] = x[i] - lam * d[i];
      Fn(Ft, xt);
      ++n_evals;
      rnorm_t = std::sqrt(vdot(Ft, Ft));
      if (rnorm_t <= (1 - armijo * lam * (1 - eta)) * rnorm) {
        accepted = true;
        break;
      }
      if (n_evals >= max_evals) break;
      lam /= 2;
    }
    if (rnorm_t < rnorm_best) {
      rnorm_best = rnorm_t;
      xb = xt;
      Fb = Ft;
    }
    if (!accepted) {
      status = n_evals >= max_evals ? 2 : 1;
      break;
    }
    x = xt;
    F = Ft;
    rnorm_prev = rnorm;
    rnorm = rnorm_t;
  }
  if (status != 0 && rnorm_best < rnorm) {
    x = xb;
    F = Fb;
    rnorm = rnorm_best;
  }
  *n_evals_out = n_evals;
  *rnorm_out = rnorm;
  *n_newton_out = n_newton;
  *rnorm0_out = rnorm0;
  return status;
}

}  // namespace

extern "C" {

// gmres! on the dense operator y = A u (A row-major n x n); returns the product count.
int orc_gmres_dense(int n, const double* A, const double* b, double rtol, double atol, int itmax, int mem, double* x,
                    double* rnorm) {
  Vec bv(b, b + n), xv(n);
  auto op = [&](Vec& y, const Vec& u) {
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += A[(size_t)i * n + j] * u[j];
      y[i] = s;
    }
  };
  const int k = gmres(xv, op, bv, mem, rtol, atol, itmax, rnorm);
  std::memcpy(x, xv.data(), sizeof(double) * n);
  return k;
}

// newton_krylov! on F(x) = x - x0 - dt f((x0 + x) / 2), f(y) = tanh(B y) - y^3 / 2
// (the implicit-midpoint-like problem of the reference's own solver tests).
// out[0..3] = n_evals, rnorm, n_newton, rnorm0; returns the status.
int orc_newton_midpoint(int n, const double* B, const double* x0, double dt, double tol, int max_evals, double fd_h,
                        int krylov_max, double* x, double* F_out, double* out) {
  auto Fn = [&](Vec& F, const Vec& xv) {
    Vec y(n);
    for (int i = 0; i < n; ++i) y[i] = 0.5 * (x0[i] + xv[i]);
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += B[(size_t)i * n + j] * y[j];
      F[i] = xv[i] - x0[i] - dt * (std::tanh(s) - 0.5 * y[i] * y[i] * y[i]);
    }
  };
  Vec xv(x, x + n), F(n);
  int ne, nn;
  double rn, r0;
  const int st = newton_krylov(xv, Fn, F, krylov_max, tol, max_evals, fd_h, 0.9, &ne, &rn, &nn, &r0);
  std::memcpy(x, xv.data(), sizeof(double) * n);
  std::memcpy(F_out, F.data(), sizeof(double) * n);
  out[0] = ne;
  out[1] = rn;
  out[2] = nn;
  out[3] = r0;
  return st;
}

// solver.jl:437-486 (step_newton!): F(v) = v - G(v) with G the Picard map;
// v1 holds the initial guess on entry and the solution (G(v), with
// exit_picard_step) on exit.
void orc_step_newton(const orc_mesh* ms, const orc_solver_params* Pp, double* v1, const double* v0, const double* w,
                     int64_t N, double S0, orc_step_stats* st) {
  const orc_solver_params& P = *Pp;
  const int64_t n2 = 2 * N;
  Scratch s(N, ms->n_dofs);
  Vec Gv(n2);
  auto Fn = [&](Vec& F, const Vec& x) {
    picard_map(ms, Gv.data(), x.data(), v0, w, N, S0, P.dt, s, P);
    for (int64_t i = 0; i < n2; ++i) F[i] = x[i] - Gv[i];
  };
  Vec x(v1, v1 + n2), F(n2);
  const double nv = std::sqrt(vdot(x, x));
  const double eff_tol = std::max(P.tol * (nv + 1e-30), P.abs_floor);
  const double fd_h = P.nk_fd_rel > 0 ? P.nk_fd_rel * std::max(nv, 1.0) : (P.nk_fd_h > 0 ? P.nk_fd_h : 1e-6);
  int ne, nn;
  double rn, r0;
  const int status = newton_krylov(x, Fn, F, P.nk_krylov_max, eff_tol, P.max_iter, fd_h, P.nk_eta_max, &ne, &rn, &nn,
                                   &r0);
  if (P.exit_picard_step)
    for (int64_t i = 0; i < n2; ++i) x[i] -= F[i];
  std::memcpy(v1, x.data(), sizeof(double) * n2);
  st->iters = ne;
  st->n_restart = nn;
  st->residual = rn;
  st->r0 = r0;
  st->exit_reason = status;
  st->inner = 0;
  st->n_clusters = 0;
  st->cl_fail = 0;
}

// v (in/out), fc (in: projection of v; out: projection of the new v),
// S0 = entropy of fc on entry.
void orc_time_step(const orc_mesh* ms, const orc_solver_params* P, double* v, const double* w, int64_t N, double* fc,
                   double S0, orc_step_stats* st) {
  const int64_t n2 = 2 * N;
  std::vector<double> v1(n2);
  orc_predictor(ms, P, v, w, N, fc, v1.data(), nullptr);
  if (P->solver == 1)
    orc_step_defect(ms, P, v1.data(), v, w, N, S0, st, nullptr);
  else if (P->solver == 2)
    orc_step_newton(ms, P, v1.data(), v, w, N, S0, st);
  else
    orc_step_anderson(ms, P, v1.data(), v, w, N, S0, st, nullptr);
  std::memcpy(v, v1.data(), sizeof(double) * n2);
  orc_l2_project(ms, v, w, N, fc);
  st->entropy = orc_entropy(ms, fc, P->use_logsq);
  double pe[3];
  orc_momentum_energy(v, w, N, pe);
  st->p1 = pe[0];
  st->p2 = pe[1];
  st->energy = pe[2];
}

}  // extern "C"
