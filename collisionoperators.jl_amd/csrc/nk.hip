// nk.hip -- matrix-free GMRES and inexact Jacobian-free Newton-Krylov on device
// vectors (newton_krylov.jl; wired to the Picard map as step_newton!,
// solver.jl:437-486).
//
// The Krylov basis, the iterates and the residuals live in HBM (at N = 65536
// and a 30-vector basis, 32 MB); the host keeps the (mem x mem) Givens-rotated
// Hessenberg matrix and the scalars.  Each Arnoldi step is one operator product
// (for Newton, one Picard map) followed by modified Gram-Schmidt on the device:
// the coefficients h_i stay in device memory between the dot and the axpy
// launches, and the host reads the finished column once per step.  Dots are
// deterministic: block partials in a fixed tree, then one workgroup over the
// partials in a fixed order.
#include "nk.hpp"

#include <cmath>
#include <limits>

namespace collop {
namespace {

constexpr int kNkThreads = 256;
constexpr int kNkMaxBlocks = 256;

int nk_blocks(int n) {
  const int b = (n + 2 * kNkThreads - 1) / (2 * kNkThreads);
  return b < 1 ? 1 : (b > kNkMaxBlocks ? kNkMaxBlocks : b);
}

__global__ __launch_bounds__(kNkThreads) void k_dot_part(const double* __restrict__ a, const double* __restrict__ b,
                                                         int n, double* __restrict__ part) {
  __shared__ double red[kNkThreads / 64];
  double s = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += a[i] * b[i];
  s = block_sum<kNkThreads>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[idx] = sum of the partials (sqrt of it when take_sqrt)
__global__ __launch_bounds__(kNkThreads) void k_dot_final(const double* __restrict__ part, int nb,
                                                          double* __restrict__ out, int idx, int take_sqrt) {
  __shared__ double red[kNkThreads / 64];
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[b];
  s = block_sum<kNkThreads>(s, red);
  if (threadIdx.x == 0) out[idx] = take_sqrt ? sqrt(s) : s;
}

// y -= h[idx] x: the coefficient is read from device memory
__global__ void k_axpy_dev(double* __restrict__ y, const double* __restrict__ x, int n, const double* __restrict__ h,
                           int idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] -= h[idx] * x[i];
}

// y = x / d
__global__ void k_div(double* __restrict__ y, const double* __restrict__ x, int n, double d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] / d;
}

// y = a + alpha b
__global__ void k_axpby(double* __restrict__ y, const double* __restrict__ a, double alpha,
                        const double* __restrict__ b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a[i] + alpha * b[i];
}

// y = (Fp - F) / h
__global__ void k_fd(double* __restrict__ y, const double* __restrict__ Fp, const double* __restrict__ F, double h,
                     int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (Fp[i] - F[i]) / h;
}

struct Coef {
  double z[kNkMaxKrylov];
};

// x = sum_j z_j V_j, accumulated in j order from zero (x = V z)
__global__ void k_lincomb(double* __restrict__ x, const double* __restrict__ V, size_t ld, int k, Coef c, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < k; ++j) s += c.z[j] * V[(size_t)j * ld + i];
  x[i] = s;
}

// y = -x
__global__ void k_neg(double* __restrict__ y, const double* __restrict__ x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = -x[i];
}

// y = A u, A row-major n x n, one row per thread in column order
__global__ void k_dense_mv(double* __restrict__ y, const double* __restrict__ A, const double* __restrict__ u, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int j = 0; j < n; ++j) s += A[(size_t)i * n + j] * u[j];
  y[i] = s;
}

inline dim3 ew(int n) { return dim3((n + 255) / 256); }

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

}  // namespace

int NkDevice::dot_dev(const double* a, const double* b, int idx, bool take_sqrt) {
  const int nb = nk_blocks(n);
  hipLaunchKernelGGL(k_dot_part, dim3(nb), dim3(kNkThreads), 0, st, a, b, n, part);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kNkThreads), 0, st, part, nb, hdev, idx, take_sqrt ? 1 : 0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int NkDevice::fetch(int count, double* out) {
  if (hipMemcpyAsync(out, hdev, (size_t)count * 8, hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
  return hipStreamSynchronize(st) == hipSuccess ? 0 : -1;
}

int NkDevice::norm(const double* a, double* out) {
  if (dot_dev(a, a, 0, true)) return -1;
  return fetch(1, out);
}

// newton_krylov.jl:60-131 (gmres!)
int NkDevice::gmres(double* x, const Op& A, const double* b, int mem, double rtol, double atol, int itmax, int* k_out,
                    double* rnorm_out) {
  mem = std::min(itmax, mem);
  if (mem > kNkMaxKrylov) return -1;
  if (hipMemsetAsync(x, 0, (size_t)n * 8, st) != hipSuccess) return -1;
  double beta;
  if (norm(b, &beta)) return -1;
  if (beta == 0.0) {
    *k_out = 0;
    *rnorm_out = 0.0;
    return 0;
  }
  const double eps_tol = atol + rtol * beta;
  const double btol = std::pow(2.220446049250313e-16, 0.75);
  std::vector<double> R((size_t)std::max(mem, 1) * std::max(mem, 1), 0.0), c(mem), s(mem), z(mem + 1, 0.0);
  auto Rm = [&](int i, int j) -> double& { return R[(size_t)i * mem + j]; };
  z[0] = beta;
  hipLaunchKernelGGL(k_div, ew(n), dim3(256), 0, st, V, b, n, beta);
  int k = 0;
  double rnorm = beta;
  double hcol[kNkMaxKrylov + 1];
  while (k < mem) {
    ++k;
    const double* vk = V + (size_t)(k - 1) * ldv;
    if (A(w, vk)) return -1;
    for (int i = 0; i < k; ++i) {  // modified Gram-Schmidt, h_i kept on the device
      const double* vi = V + (size_t)i * ldv;
      if (dot_dev(vi, w, i, false)) return -1;
      hipLaunchKernelGGL(k_axpy_dev, ew(n), dim3(256), 0, st, w, vi, n, hdev, i);
    }
    if (dot_dev(w, w, k, true)) return -1;
    if (fetch(k + 1, hcol)) return -1;
    for (int i = 0; i < k; ++i) Rm(i, k - 1) = hcol[i];
    const double hnext = hcol[k];
    for (int i = 0; i < k - 1; ++i) {
      const double tmp = c[i] * Rm(i, k - 1) + s[i] * Rm(i + 1, k - 1);
      Rm(i + 1, k - 1) = s[i] * Rm(i, k - 1) - c[i] * Rm(i + 1, k - 1);
      Rm(i, k - 1) = tmp;
    }
    sym_givens(Rm(k - 1, k - 1), hnext, c[k - 1], s[k - 1], Rm(k - 1, k - 1));
    z[k] = s[k - 1] * z[k - 1];
    z[k - 1] = c[k - 1] * z[k - 1];
    rnorm = std::fabs(z[k]);
    if (rnorm <= eps_tol || hnext <= btol) break;
    if (k < mem) hipLaunchKernelGGL(k_div, ew(n), dim3(256), 0, st, V + (size_t)k * ldv, w, n, hnext);
  }
  for (int i = k - 1; i >= 0; --i) {
    double acc = z[i];
    for (int j = i + 1; j < k; ++j) acc -= Rm(i, j) * z[j];
    z[i] = std::fabs(Rm(i, i)) <= btol ? 0.0 : acc / Rm(i, i);
  }
  Coef cf{};
  for (int i = 0; i < k; ++i) cf.z[i] = z[i];
  hipLaunchKernelGGL(k_lincomb, ew(n), dim3(256), 0, st, x, V, ldv, k, cf, n);
  if (hipGetLastError() != hipSuccess) return -1;
  *k_out = k;
  *rnorm_out = rnorm;
  return 0;
}

// newton_krylov.jl:189-303 (newton_krylov!)
int NkDevice::newton(double* x, const Resid& Fn, double* F, int krylov_max, double tol, int max_evals, double fd_h,
                     double eta_max, NkResult* res) {
  const double eta_gamma = 0.9, armijo = 1e-4;
  const int max_backtrack = 8;
  const size_t nb = (size_t)n * 8;
  double rnorm;
  if (Fn(x, F, &rnorm)) return -1;
  int n_evals = 1;
  const double rnorm0 = rnorm;
  double rnorm_best = rnorm;
  if (hipMemcpyAsync(xb, x, nb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
  if (hipMemcpyAsync(Fb, F, nb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
  bool fail = false;
  // J u by forward difference about x (F holds F(x)); the probe moves by fd_h
  Op jvp = [&](double* y, const double* u) -> int {
    double nu;
    if (norm(u, &nu)) return -1;
    const double h = fd_h / nu;
    hipLaunchKernelGGL(k_axpby, ew(n), dim3(256), 0, st, xp, x, h, u, n);
    double nf;
    if (Fn(xp, Fp, &nf)) return -1;
    ++n_evals;
    hipLaunchKernelGGL(k_fd, ew(n), dim3(256), 0, st, y, Fp, F, h, n);
    return hipGetLastError() == hipSuccess ? 0 : -1;
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
      double ewc = eta_gamma * (rnorm / rnorm_prev) * (rnorm / rnorm_prev);
      const double safe = eta_gamma * eta * eta;
      if (safe > 0.1) ewc = std::max(ewc, safe);
      eta = std::min(std::max(ewc, 0.5 * tol / rnorm), eta_max);
    }
    int k_lin;
    double r_lin;
    if (gmres(d, jvp, F, krylov_max, eta, 0.0, std::min(krylov_max, budget - 1), &k_lin, &r_lin)) {
      fail = true;
      break;
    }
    ++n_newton;
    double lam = 1.0, rnorm_t = std::numeric_limits<double>::infinity();
    bool accepted = false;
    for (int bt = 0; bt <= max_backtrack; ++bt) {  // Armijo backtracking on ||F||
      hipLaunchKernelGGL(k_axpby, ew(n), dim3(256), 0, st, xt, x, -lam, d, n);
      if (Fn(xt, Ft, &rnorm_t)) return -1;
      ++n_evals;
      if (rnorm_t <= (1 - armijo * lam * (1 - eta)) * rnorm) {
        accepted = true;
        break;
      }
      if (n_evals >= max_evals) break;
      lam /= 2;
    }
    if (rnorm_t < rnorm_best) {
      rnorm_best = rnorm_t;
      if (hipMemcpyAsync(xb, xt, nb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
      if (hipMemcpyAsync(Fb, Ft, nb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    }
    if (!accepted) {
      status = n_evals >= max_evals ? 2 : 1;
      break;
    }
    if (hipMemcpyAsync(x, xt, nb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    if (hipMemcpyAsync(F, Ft, nb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    rnorm_prev = rnorm;
    rnorm = rnorm_t;
  }
  if (fail) return -1;
  if (status != 0 && rnorm_best < rnorm) {
    if (hipMemcpyAsync(x, xb, nb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    if (hipMemcpyAsync(F, Fb, nb, hipMemcpyDeviceToDevice, st) != hipSuccess) return -1;
    rnorm = rnorm_best;
  }
  res->status = status;
  res->n_evals = n_evals;
  res->n_newton = n_newton;
  res->rnorm = rnorm;
  res->rnorm0 = rnorm0;
  return 0;
}

void launch_neg(double* y, const double* x, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_neg, ew(n), dim3(256), 0, st, y, x, n);
}

void launch_dense_mv(double* y, const double* A, const double* u, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_dense_mv, ew(n), dim3(256), 0, st, y, A, u, n);
}

void launch_axpby(double* y, const double* a, double alpha, const double* b, int n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_axpby, ew(n), dim3(256), 0, st, y, a, alpha, b, n);
}

size_t nk_part_size() { return kNkMaxBlocks; }

}  // namespace collop
