// nk.hpp -- the device Newton-Krylov engine (nk.hip): matrix-free GMRES and
// inexact Jacobian-free Newton-Krylov (newton_krylov.jl) on device vectors of
// length n, driven from the host.
#pragma once
#include <functional>
#include <vector>

#include "collop_launch.hpp"

namespace collop {

constexpr int kNkMaxKrylov = 64;  // largest GMRES basis (nk_krylov_max)

struct NkResult {
  int status;     // 0 converged, 1 stalled (line search failed), 2 budget
  int n_evals;    // residual evaluations
  int n_newton;   // Newton steps
  double rnorm, rnorm0;
};

// Device buffers (all length n unless noted) and the stream the engine runs on.
struct NkDevice {
  // y <- A u (device pointers); nonzero = failure
  using Op = std::function<int(double* y, const double* u)>;
  // F <- F(x), *norm <- ||F|| (device pointers); nonzero = failure
  using Resid = std::function<int(const double* x, double* F, double* norm)>;

  int n = 0;
  hipStream_t st = nullptr;
  double* V = nullptr;   // Krylov basis, (mem + 1) columns of ldv
  size_t ldv = 0;
  double* w = nullptr;   // Arnoldi work vector
  double* part = nullptr;  // dot partials (nk_part_size())
  double* hdev = nullptr;  // kNkMaxKrylov + 1 device scalars
  double *Ft = nullptr, *xt = nullptr, *xp = nullptr, *Fp = nullptr, *d = nullptr, *xb = nullptr, *Fb = nullptr;

  // gmres! from x = 0; *k_out products, *rnorm_out the GMRES residual norm
  int gmres(double* x, const Op& A, const double* b, int mem, double rtol, double atol, int itmax, int* k_out,
            double* rnorm_out);
  // newton_krylov! from x (overwritten); F holds F(x) on return
  int newton(double* x, const Resid& Fn, double* F, int krylov_max, double tol, int max_evals, double fd_h,
             double eta_max, NkResult* res);
  int norm(const double* a, double* out);

 private:
  int dot_dev(const double* a, const double* b, int idx, bool take_sqrt);
  int fetch(int count, double* out);
};

size_t nk_part_size();
void launch_neg(double* y, const double* x, int n, hipStream_t st);
void launch_dense_mv(double* y, const double* A, const double* u, int n, hipStream_t st);
void launch_axpby(double* y, const double* a, double alpha, const double* b, int n, hipStream_t st);

}  // namespace collop
