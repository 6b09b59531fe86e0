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
// shuffle tree, then over blocks in a fixed order (deterministic).
#include "collop_launch.hpp"

namespace collop {
namespace {

constexpr int kLbThreads = 256;

// Block sums of one block of markers, into part[blockIdx.x][kLbSums]:
//   0 n, 1 U1, 2 U2, 3 Q, 4 Sg1, 5 Sg2, 6 P          (at x, with g)
//   7 sum dv1, 8 sum dv2, 9 sum x.dv, 10 -sum w dv.g, 11 sum |dv|^2   (dv given)
// coeff given: g = clamp(grad f / max(|f|, fs_floor), +-g_max) of f = sum coeff
// phi at x, zero outside the box (functions.jl:37-64), written to g; coeff
// null: g is read.
__global__ __launch_bounds__(kLbThreads) void k_lb_gather(MeshView m, const double* __restrict__ x,
                                                          const double* __restrict__ coeff,
                                                          const double* __restrict__ w,
                                                          const double* __restrict__ dv, int N, double fs_floor,
                                                          double g_max, double* __restrict__ g,
                                                          double* __restrict__ part) {
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kLbSums * kLbThreads / 64];
  if (coeff) stage_bp(m, sb1, sb2);
  __syncthreads();
  double s[kLbSums];
#pragma unroll
  for (int j = 0; j < kLbSums; ++j) s[j] = 0.0;
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < N; a += gridDim.x * blockDim.x) {
    const double x1 = x[a], x2 = x[N + a];
    double o1 = 0.0, o2 = 0.0;
    if (coeff) {
      Located L;
      if (locate(m, sb1, sb2, x1, x2, L)) {
        double f1[3], d1[3], f2[3], d2[3];
        basis3_grad(m.C1, L.i, L.xi1, f1, d1);
        basis3_grad(m.C2, L.j, L.xi2, f2, d2);
        const double ih1 = 1.0 / L.h1, ih2 = 1.0 / L.h2;
        const int s1 = m.bs1[L.i], s2 = m.bs2[L.j];
        double acc1 = 0.0, acc2 = 0.0, fv = 0.0;
#pragma unroll
        for (int j2 = 0; j2 < 3; ++j2)
#pragma unroll
          for (int j1 = 0; j1 < 3; ++j1) {  // (j2, j1) order as functions.jl:48-55
            const double c = coeff[(s1 + j1) + (s2 + j2) * m.nd1];
            acc1 += c * (d1[j1] * f2[j2]) * ih1;
            acc2 += c * (f1[j1] * d2[j2]) * ih2;
            fv += c * (f1[j1] * f2[j2]);
          }
        const double invf = 1.0 / fmax(fabs(fv), fs_floor);
        o1 = fmin(fmax(acc1 * invf, -g_max), g_max);
        o2 = fmin(fmax(acc2 * invf, -g_max), g_max);
      }
      g[a] = o1;
      g[N + a] = o2;
    } else {
      o1 = g[a];
      o2 = g[N + a];
    }
    const double wa = w[a];
    s[0] += wa;
    s[1] += wa * x1;
    s[2] += wa * x2;
    s[3] += wa * (x1 * x1 + x2 * x2);
    s[4] += wa * o1;
    s[5] += wa * o2;
    s[6] += wa * (x1 * o1 + x2 * o2);
    if (dv) {
      const double e1 = dv[a], e2 = dv[N + a];
      s[7] += e1;
      s[8] += e2;
      s[9] += x1 * e1 + x2 * e2;
      s[10] -= wa * (e1 * o1 + e2 * o2);
      s[11] += e1 * e1 + e2 * e2;
    }
  }
  block_sum_n<kLbThreads, kLbSums>(s, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < kLbSums; ++j) part[(size_t)blockIdx.x * kLbSums + j] = s[j];
  }
}

// One workgroup: the block sums in a fixed order; the Gonzalez lambda when the
// entropy partials qpart (nq > 0) of S(f_in) are given (solver.jl:196-205):
// lambda = (S1 - S0 - sum -w dv.g) / |dv|^2, 0 for |dv|^2 <= 1e-30, and the
// corrected G_eff = g - lambda dv / w enters Sg and P through sum dv and sum
// x.dv; then the 3x3 solve by LU with partial pivoting (Julia's `\` on a square
// matrix, functions.jl:102-104).  out = (lambda, A1, A2, B, S1).
__global__ __launch_bounds__(kLbThreads) void k_lb_solve(const double* __restrict__ part, int nparts,
                                                         const double* __restrict__ qpart, int nq,
                                                         const double* __restrict__ S0, double* __restrict__ out) {
  __shared__ double red[(kLbSums + 1) * kLbThreads / 64];
  double s[kLbSums + 1];
#pragma unroll
  for (int j = 0; j <= kLbSums; ++j) s[j] = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
    for (int j = 0; j < kLbSums; ++j) s[j] += part[(size_t)b * kLbSums + j];
  for (int b = threadIdx.x; b < nq; b += blockDim.x) s[kLbSums] += qpart[b];
  block_sum_n<kLbThreads, kLbSums + 1>(s, red);
  if (threadIdx.x != 0) return;
  const double S1 = s[kLbSums];
  const double lam = (nq > 0 && s[11] > 1e-30) ? (S1 - S0[0] - s[10]) / s[11] : 0.0;
  double M[3][3] = {{s[0], 0.0, s[1]}, {0.0, s[0], s[2]}, {s[1], s[2], s[3]}};
  double r[3] = {-(s[4] - lam * s[7]), -(s[5] - lam * s[8]), -(s[6] - lam * s[9])};
  for (int k = 0; k < 3; ++k) {
    int pr = k;
    for (int i = k + 1; i < 3; ++i)
      if (fabs(M[i][k]) > fabs(M[pr][k])) pr = i;
    if (pr != k) {
      for (int j = 0; j < 3; ++j) {
        const double t = M[k][j];
        M[k][j] = M[pr][j];
        M[pr][j] = t;
      }
      const double t = r[k];
      r[k] = r[pr];
      r[pr] = t;
    }
    for (int i = k + 1; i < 3; ++i) {
      const double f = M[i][k] / M[k][k];
      for (int j = k; j < 3; ++j) M[i][j] -= f * M[k][j];
      r[i] -= f * r[k];
    }
  }
  double sol[3];
  for (int i = 2; i >= 0; --i) {
    double t = r[i];
    for (int j = i + 1; j < 3; ++j) t -= M[i][j] * sol[j];
    sol[i] = t / M[i][i];
  }
  out[0] = lam;
  out[1] = sol[0];
  out[2] = sol[1];
  out[3] = sol[2];
  out[4] = S1;
}

// dot_a = -nu (G_eff_a + A + B x_a), G_eff = g - lambda dv / w (dv given)
// (compute_LB_velocity!, functions.jl:108-115).
__global__ void k_lb_velocity(const double* __restrict__ x, const double* __restrict__ g,
                              const double* __restrict__ dv, const double* __restrict__ w,
                              const double* __restrict__ coef, double nu, int N, double* __restrict__ dot) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= N) return;
  const double lam = coef[0], A1 = coef[1], A2 = coef[2], B = coef[3];
  double g1 = g[a], g2 = g[N + a];
  if (dv) {
    const double iw = 1.0 / w[a];
    g1 -= lam * dv[a] * iw;
    g2 -= lam * dv[N + a] * iw;
  }
  dot[a] = -nu * (g1 + A1 + B * x[a]);
  dot[N + a] = -nu * (g2 + A2 + B * x[N + a]);
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
    hipLaunchKernelGGL(k_lb_velocity, dim3((N + 255) / 256), dim3(256), 0, st, x, g, dv, w, coef, nu, N, dot);
}

}  // namespace collop
