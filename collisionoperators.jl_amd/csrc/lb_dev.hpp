// lb_dev.hpp -- device bodies of the Lenard-Bernstein operator (functions.jl:66-115),
// shared by its own kernels (lb.hip), the gather fused with the entropy
// quadrature (fem.hip) and the Anderson epilogue that forms v_dot itself
// (step.hip).  Every caller runs workgroups of kLbThreads threads, so the
// reductions below are bitwise the same in each of them.
#pragma once
#include "collop_launch.hpp"

namespace collop {

constexpr int kLbThreads = 256;

// Block sums of markers lb*256 + tid, stride nb*256, into part[lb][kLbSums]:
//   0 n, 1 U1, 2 U2, 3 Q, 4 Sg1, 5 Sg2, 6 P          (at x, with g)
//   7 sum dv1, 8 sum dv2, 9 sum x.dv, 10 -sum w dv.g, 11 sum |dv|^2   (dv given)
// coeff given: g = clamp(grad f / max(|f|, fs_floor), +-g_max) of f = sum coeff
// phi at x, zero outside the box (functions.jl:37-64), written to g; coeff
// null: g is read.  sb1/sb2 hold kMaxBp doubles, red kLbSums * kLbThreads / 64.
__device__ __forceinline__ void lb_gather_body(const MeshView& m, const double* __restrict__ x,
                                               const double* __restrict__ coeff, const double* __restrict__ w,
                                               const double* __restrict__ dv, int N, double fs_floor, double g_max,
                                               double* __restrict__ g, double* __restrict__ part, double* sb1,
                                               double* sb2, double* red, int lb, int nb) {
  if (coeff) stage_bp(m, sb1, sb2);
  __syncthreads();
  double s[kLbSums];
#pragma unroll
  for (int j = 0; j < kLbSums; ++j) s[j] = 0.0;
  for (int a = lb * kLbThreads + threadIdx.x; a < N; a += nb * kLbThreads) {
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
    s[3] += wa * dot2(x1, x1, x2, x2);
    s[4] += wa * o1;
    s[5] += wa * o2;
    s[6] += wa * dot2(x1, o1, x2, o2);
    if (dv) {
      const double e1 = dv[a], e2 = dv[N + a];
      s[7] += e1;
      s[8] += e2;
      s[9] += dot2(x1, e1, x2, e2);
      s[10] -= wa * dot2(e1, o1, e2, o2);
      s[11] += dot2(e1, e1, e2, e2);
    }
  }
  block_sum_n<kLbThreads, kLbSums>(s, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < kLbSums; ++j) part[(size_t)lb * kLbSums + j] = s[j];
  }
}

// The gather's block sums in a fixed order, the Gonzalez lambda when the
// entropy partials qpart (nq > 0) of S(f_in) are given (solver.jl:196-205):
// lambda = (S1 - S0 - sum -w dv.g) / |dv|^2, 0 for |dv|^2 <= 1e-30; the
// corrected G_eff = g - lambda dv / w enters Sg and P through sum dv and sum
// x.dv; then the 3x3 solve [n 0 U1; 0 n U2; U1 U2 Q] (A1, A2, B) = -(Sg1, Sg2, P)
// by LU with partial pivoting (Julia's `\` on a square matrix,
// functions.jl:102-104).  All threads call it; thread 0 writes out = (lambda,
// A1, A2, B, S1).  red holds (kLbSums + 1) * kLbThreads / 64 doubles.
__device__ __forceinline__ void lb_coef_body(const double* __restrict__ part, int nparts,
                                             const double* __restrict__ qpart, int nq, const double* __restrict__ S0,
                                             double* red, double* out) {
  double s[kLbSums + 1];
#pragma unroll
  for (int j = 0; j <= kLbSums; ++j) s[j] = 0.0;
  for (int b = threadIdx.x; b < nparts; b += kLbThreads)
#pragma unroll
    for (int j = 0; j < kLbSums; ++j) s[j] += part[(size_t)b * kLbSums + j];
  for (int b = threadIdx.x; b < nq; b += kLbThreads) s[kLbSums] += qpart[b];
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

// Component i < 2N of dot = -nu (G_eff + A + B x), G_eff = g - lambda dv / w
// (dv given) (compute_LB_velocity!, functions.jl:108-115); coef = (lambda, A1,
// A2, B).
__device__ __forceinline__ double lb_dot_value(const double* __restrict__ x, const double* __restrict__ g,
                                               const double* __restrict__ dv, const double* __restrict__ w,
                                               const double* coef, double nu, int N, int i) {
  const int a = i < N ? i : i - N;
  double gi = g[i];
  if (dv) {
    const double iw = 1.0 / w[a];
    gi -= coef[0] * dv[i] * iw;
  }
  return -nu * (gi + (i < N ? coef[1] : coef[2]) + coef[3] * x[i]);
}

}  // namespace collop
