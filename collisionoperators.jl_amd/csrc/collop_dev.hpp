// collop_dev.hpp -- types and device helpers shared by the gfx950 kernels.
//
// Data layout in HBM (one context, N markers, P_DEG = 2):
//   markers     v[2N]  : v1[0..N) then v2[0..N)   (Julia column-major N x 2)
//   weights     w[N]
//   pair source Src64[N] / Src32[N]: one 64 B / 32 B record per source marker
//               {v1, v2, G1, G2, w_masked, in_domain, -, -}; read wave-uniformly
//               (scalar loads) by the pair kernels
//   slice partials part[S][2N]: per-alpha-slice pair sums, reduced in a fixed
//               order (deterministic)
//   mesh tables: breakpoints, 3x3 extraction per cell, 0-based basis starts,
//               banded Cholesky factors of the 1D mass matrices, basis values at
//               the Gauss points of every cell
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace collop {

constexpr int kMaxWindow = 8;     // largest Anderson window m
constexpr int kScatterThreads = 256;
constexpr int kMaxBp = 512;       // breakpoints per direction staged in LDS

struct MeshView {
  const double* bp1;
  const double* bp2;
  const double* C1;      // 9 * ne1, C[i + 3 j + 9 cell]
  const double* C2;
  const int* bs1;        // 0-based first basis index per cell
  const int* bs2;
  const double* chol1;   // 3 * nd1: {1 / L_ii, L_i,i-1, L_i,i-2}
  const double* chol2;
  const double* minv1;   // nd1 x nd1 dense inverse of the 1D mass matrix (symmetric)
  const double* minv2;
  const double* phiq1;   // ne1 * 3 * nq: phi_j(x_a) on cell e at [(3 e + j) nq + a]
  const double* phiq2;
  const double* qw;      // nq Gauss weights on [0,1]
  const int* erange1;    // 2 * nd1: first / last cell in the support of basis i
  const int* erange2;
  int nb1, nb2, ne1, ne2, nd1, nd2, n_dofs, nq;
  int uniform1, uniform2;
  double lo1, hi1, lo2, hi2;
  double inv_h1, inv_h2;   // 1 / cell width when uniform (cell guess only)
};

struct alignas(64) Src64 {
  double v1, v2, g1, g2, w, in_domain, pad0, pad1;
};

struct alignas(32) Src32 {
  float v1, v2, g1, g2, w, in_domain, pad0, pad1;
};

__device__ __forceinline__ bool in_box(double a, double b, double lo1, double hi1, double lo2, double hi2) {
  return !(a <= lo1 || a >= hi1 || b <= lo2 || b >= hi2);
}

// Largest i with bp[i] <= x for bp[0] < x < bp[nb-1] (searchsortedlast,
// MantisWrappers.jl:171; projection_gpu.jl:20-32).  `bp` may live in LDS.
__device__ __forceinline__ int cell_of(const double* bp, int nb, int uniform, double lo, double inv_h, double x) {
  if (uniform) {
    int i = (int)((x - lo) * inv_h);
    i = i < 0 ? 0 : (i > nb - 2 ? nb - 2 : i);
    while (i > 0 && bp[i] > x) --i;
    while (i < nb - 2 && bp[i + 1] <= x) ++i;
    return i;
  }
  int lo_i = 0, hi_i = nb - 1;
  while (lo_i + 1 < hi_i) {
    int mid = (lo_i + hi_i) >> 1;
    if (bp[mid] <= x) lo_i = mid; else hi_i = mid;
  }
  return lo_i;
}

// Quadratic Bernstein values / derivatives at xi (projection_gpu.jl:34-35)
// mapped through the cell's extraction operator (phi_j = sum_i C[i,j] B_i).
__device__ __forceinline__ void basis3(const double* __restrict__ C, int cell, double xi, double phi[3]) {
  const double om = 1.0 - xi;
  const double B0 = om * om, B1 = 2.0 * xi * om, B2 = xi * xi;
  const double* c = C + 9 * cell;
#pragma unroll
  for (int j = 0; j < 3; ++j) phi[j] = c[3 * j] * B0 + c[3 * j + 1] * B1 + c[3 * j + 2] * B2;
}

__device__ __forceinline__ void basis3_grad(const double* __restrict__ C, int cell, double xi, double phi[3],
                                            double dphi[3]) {
  const double om = 1.0 - xi;
  const double B0 = om * om, B1 = 2.0 * xi * om, B2 = xi * xi;
  const double dB0 = -2.0 * om, dB1 = 2.0 - 4.0 * xi, dB2 = 2.0 * xi;
  const double* c = C + 9 * cell;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    phi[j] = c[3 * j] * B0 + c[3 * j + 1] * B1 + c[3 * j + 2] * B2;
    dphi[j] = c[3 * j] * dB0 + c[3 * j + 1] * dB1 + c[3 * j + 2] * dB2;
  }
}

// a b + c d rounded as fma(a, b, c d).  Sums of two products that more than one
// kernel forms use it, so that the compiler's choice of which product to fuse
// cannot differ between those kernels.
__device__ __forceinline__ double dot2(double a, double b, double c, double d) { return __builtin_fma(a, b, c * d); }

// Deterministic block sum (fixed shuffle tree, then fixed order over waves).
// All threads of the block must call it; the result is valid in thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double x, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[wid] = x;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) t += sh[i];
  }
  return t;
}

// Deterministic block sums of NV values at once: one shuffle tree per value,
// one pair of barriers for all of them; the sums are valid in thread 0 and are
// bitwise those of NV separate block_sum calls.  sh holds (NT / 64) * NV doubles.
template <int NT, int NV>
__device__ __forceinline__ void block_sum_n(double (&x)[NV], double* sh) {
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x[v] += __shfl_xor(x[v], off, 64);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) sh[wid * NV + v] = x[v];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < NT / 64; ++i) t += sh[i * NV + v];
      x[v] = t;
    }
  }
}

// ---- marker location against LDS breakpoints (fem.hip, ensemble.hip) ----
__device__ __forceinline__ void stage_bp(const MeshView& m, double* sb1, double* sb2) {
  for (int i = threadIdx.x; i < m.nb1; i += blockDim.x) sb1[i] = m.bp1[i];
  for (int i = threadIdx.x; i < m.nb2; i += blockDim.x) sb2[i] = m.bp2[i];
}

struct Located {
  int i, j;
  double xi1, xi2, h1, h2;
};

// locate_particle (MantisWrappers.jl:168-179) against LDS breakpoints
__device__ __forceinline__ bool locate(const MeshView& m, const double* sb1, const double* sb2, double x, double y,
                                       Located& L) {
  if (!in_box(x, y, m.lo1, m.hi1, m.lo2, m.hi2)) return false;
  L.i = cell_of(sb1, m.nb1, m.uniform1, m.lo1, m.inv_h1, x);
  L.j = cell_of(sb2, m.nb2, m.uniform2, m.lo2, m.inv_h2, y);
  L.h1 = sb1[L.i + 1] - sb1[L.i];
  L.h2 = sb2[L.j + 1] - sb2[L.j];
  L.xi1 = (x - sb1[L.i]) / L.h1;
  L.xi2 = (y - sb2[L.j]) / L.h2;
  return true;
}

// ---- banded Cholesky solve along one direction of the coefficient array held
// in LDS: chol = {1/L_ii, L_i,i-1, L_i,i-2}; element i of the line is X[i * istride].
__device__ __forceinline__ void band_solve_line(double* X, int n, int istride, const double* __restrict__ chol) {
  double y2 = 0.0, y1 = 0.0;  // y_{i-2}, y_{i-1}
  for (int i = 0; i < n; ++i) {
    const double l1 = chol[3 * i + 1], l2 = chol[3 * i + 2];
    double s = X[i * istride];
    s -= l1 * y1;
    s -= l2 * y2;
    const double yi = s * chol[3 * i];
    X[i * istride] = yi;
    y2 = y1;
    y1 = yi;
  }
  double x1 = 0.0, x2 = 0.0;  // x_{i+1}, x_{i+2}
  for (int i = n - 1; i >= 0; --i) {
    const double u1 = i + 1 < n ? chol[3 * (i + 1) + 1] : 0.0;
    const double u2 = i + 2 < n ? chol[3 * (i + 2) + 2] : 0.0;
    double s = X[i * istride];
    s -= u1 * x1;
    s -= u2 * x2;
    const double xi = s * chol[3 * i];
    X[i * istride] = xi;
    x2 = x1;
    x1 = xi;
  }
}

// ---- pair terms (pair.hip, ensemble.hip) ----------------------------------
template <int NEWTON>
__device__ __forceinline__ double rsqrt_f64(double q) {
  double y = __builtin_amdgcn_rsq(q);
  const double h = 0.5 * q;
#pragma unroll
  for (int i = 0; i < NEWTON; ++i) {
    const double e = __builtin_fma(-h, y * y, 0.5);
    y = __builtin_fma(y, e, y);
  }
  return y;
}

// One ordered pair (gamma in registers, alpha uniform).  Returns the weight
// factor t = w_a / |d|^3 (0 for masked pairs) through `t_out` for the metric.
template <int NEWTON>
__device__ __forceinline__ void pair_term_f64(double vg1, double vg2, double Gg1, double Gg2, double av1, double av2,
                                              double ag1, double ag2, double aw, double& acc1, double& acc2,
                                              double& d1o, double& d2o, double& t_out) {
  const double d1 = vg1 - av1;
  const double d2 = vg2 - av2;
  const double q = __builtin_fma(d2, d2, d1 * d1);
  const double y = rsqrt_f64<NEWTON>(__builtin_fmax(q, 1e-24));
  const double g1 = ag1 - Gg1;
  const double g2 = ag2 - Gg2;
  const double c = __builtin_fma(g1, d2, -(g2 * d1));
  double t = (aw * y) * (y * y);
  t = q < 1e-24 ? 0.0 : t;  // coincident-pair guard (functions.jl:224)
  const double k = t * c;
  acc1 = __builtin_fma(k, d2, acc1);
  acc2 = __builtin_fma(-k, d1, acc2);
  d1o = d1;
  d2o = d2;
  t_out = t;
}

// FP32 pair term that also returns d and the weight factor t = w / |d|^3 (0 for
// masked pairs), for the metric.
__device__ __forceinline__ void pair_term_f32_t(float vg1, float vg2, float Gg1, float Gg2, float av1, float av2,
                                                float ag1, float ag2, float aw, float& b1, float& b2, float& d1o,
                                                float& d2o, float& t_out) {
  const float d1 = vg1 - av1;
  const float d2 = vg2 - av2;
  const float q = __builtin_fmaf(d2, d2, d1 * d1);
  const float y = __builtin_amdgcn_rsqf(__builtin_fmaxf(q, 1e-12f));
  const float g1 = ag1 - Gg1;
  const float g2 = ag2 - Gg2;
  const float c = __builtin_fmaf(g1, d2, -(g2 * d1));
  float t = (aw * y) * (y * y);
  t = q < 1e-12f ? 0.0f : t;
  const float kk = t * c;
  b1 = __builtin_fmaf(kk, d2, b1);
  b2 = __builtin_fmaf(-kk, d1, b2);
  d1o = d1;
  d2o = d2;
  t_out = t;
}

// FP32 pair term with the FP32 coincidence guard of collision_gpu.jl:187.
__device__ __forceinline__ void pair_term_f32(float vg1, float vg2, float Gg1, float Gg2, float av1, float av2,
                                              float ag1, float ag2, float aw, float& b1, float& b2) {
  const float d1 = vg1 - av1;
  const float d2 = vg2 - av2;
  const float q = __builtin_fmaf(d2, d2, d1 * d1);
  const float y = __builtin_amdgcn_rsqf(__builtin_fmaxf(q, 1e-12f));
  const float g1 = ag1 - Gg1;
  const float g2 = ag2 - Gg2;
  const float c = __builtin_fmaf(g1, d2, -(g2 * d1));
  float t = (aw * y) * (y * y);
  t = q < 1e-12f ? 0.0f : t;
  const float kk = t * c;
  b1 = __builtin_fmaf(kk, d2, b1);
  b2 = __builtin_fmaf(-kk, d1, b2);
}

// FP32 pair sums are compensated blockwise: BLK terms are summed in FP32, then
// folded into a (hi, lo) pair with Knuth's TwoSum, so the error of the
// per-marker sum does not grow with N.
__device__ __forceinline__ void two_sum_fold(float& hi, float& lo, float x) {
  const float s = hi + x;
  const float bp = s - hi;
  const float e = (hi - (s - bp)) + (x - bp);
  hi = s;
  lo += e;
}

}  // namespace collop
