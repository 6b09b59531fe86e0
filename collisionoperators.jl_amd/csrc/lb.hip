// lb.hip -- the Lenard-Bernstein collision operator on gfx950 (functions.jl:66-115;
// the LB branches of picard_map!, solver.jl:185-222, and of the predictor,
// main.jl:233-238).
//
//   v_dot_a = -nu (g_a + A + B v_a),
//
// g the clamped log-density gradient of the projection at the marker, (A1, A2,
// B) the drift multipliers: the 3x3 system [n 0 U1; 0 n U2; U1 U2 Q] (A1, A2,
// B) = -(Sg1, Sg2, P) with n = sum w, U = sum w v, Q = sum w |v|^2, Sg = sum w g,
// P = sum w v.g, which makes the collision conserve momentum and energy.
//
// Everything is O(N) and HBM/latency bound: one gather that evaluates g and
// forms twelve block sums in the same pass, one single-workgroup reduction and
// solve, one element-wise velocity kernel.  Sums are per block in a fixed
// shuffle tree, then over blocks in a fixed order (deterministic).  The
// Anderson map uses fused forms of the same bodies (lb_dev.hpp): the gather
// launch also runs the entropy quadrature's blocks (fem.hip), and the epilogue
// reduces the block sums and forms v_dot itself (step.hip).
#include "lb_dev.hpp"

namespace collop {
namespace {

__global__ __launch_bounds__(kLbThreads) void k_lb_gather(MeshView m, const double* __restrict__ x,
                                                          const double* __restrict__ coeff,
                                                          const double* __restrict__ w,
                                                          const double* __restrict__ dv, int N, double fs_floor,
                                                          double g_max, double* __restrict__ g,
                                                          double* __restrict__ part) {
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kLbSums * kLbThreads / 64];
  lb_gather_body(m, x, coeff, w, dv, N, fs_floor, g_max, g, part, sb1, sb2, red, blockIdx.x, gridDim.x);
}

// One workgroup: the reduction, lambda and the drift multipliers, out = (lambda,
// A1, A2, B, S1).
__global__ __launch_bounds__(kLbThreads) void k_lb_solve(const double* __restrict__ part, int nparts,
                                                         const double* __restrict__ qpart, int nq,
                                                         const double* __restrict__ S0, double* __restrict__ out) {
  __shared__ double red[(kLbSums + 1) * kLbThreads / 64];
  lb_coef_body(part, nparts, qpart, nq, S0, red, out);
}

__global__ void k_lb_velocity(const double* __restrict__ x, const double* __restrict__ g,
                              const double* __restrict__ dv, const double* __restrict__ w,
                              const double* __restrict__ coef, double nu, int N, double* __restrict__ dot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2 * N) dot[i] = lb_dot_value(x, g, dv, w, coef, nu, N, i);
}

}  // namespace

int lb_blocks(int N) {
  const int b = (N + kLbThreads * 2 - 1) / (kLbThreads * 2);
  return b < 1 ? 1 : (b > 512 ? 512 : b);
}

void launch_lb_rhs(const MeshView& m, const double* x, const double* coeff, const double* w, const double* dv, int N,
                   const double* qpart, int nq, const double* S0, double nu, double* g, double* part, double* coef,
                   double* dot, hipStream_t st) {
  const int nb = lb_blocks(N);
  hipLaunchKernelGGL(k_lb_gather, dim3(nb), dim3(kLbThreads), 0, st, m, x, coeff, w, dv, N, 1e-30, 100.0, g, part);
  hipLaunchKernelGGL(k_lb_solve, dim3(1), dim3(kLbThreads), 0, st, part, nb, qpart, nq, S0, coef);
  if (N > 0)
    hipLaunchKernelGGL(k_lb_velocity, dim3((2 * N + 255) / 256), dim3(256), 0, st, x, g, dv, w, coef, nu, N, dot);
}

}  // namespace collop
