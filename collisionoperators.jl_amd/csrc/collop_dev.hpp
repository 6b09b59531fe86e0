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

struct MeshView {
  const double* bp1;
  const double* bp2;
  const double* C1;      // 9 * ne1, C[i + 3 j + 9 cell]
  const double* C2;
  const int* bs1;        // 0-based first basis index per cell
  const int* bs2;
  const double* chol1;   // 3 * nd1: {1 / L_ii, L_i,i-1, L_i,i-2}
  const double* chol2;
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

}  // namespace collop
