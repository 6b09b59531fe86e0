// fem.hip -- marker <-> P_DEG=2 tensor B-spline kernels on gfx950.
//
//   scatter    rhs_k = sum_a w_a phi_k(v_a)            (kernel (b); functions.jl:121-131)
//   solve      c = M^{-1} rhs with M = M1 (x) M2       (MantisWrappers.jl:107-120)
//   quadrature r_i = sum_e sum_q phi_i (1 + log f) w_q |e|, S = -sum f log f w_q |e|
//                                                     (functions.jl:137-179)
//   gathers    G_a = grad L(v_a), g_a = grad f / f     (kernel (c); functions.jl:37-64,182-202)
//
// These are HBM/latency-bound O(N) kernels.  The scatter privatises the rhs in
// LDS (9 KB at 32x32 cells, 35 KB at 64x64) so the 9 adds per marker hit LDS,
// not the L2 atomics the reference relies on (projection_gpu.jl:67).  The adds
// are 64-bit fixed-point integers: contribution c becomes round(c * 2^62 / W)
// with W = sum |w|, which cannot overflow (each dof receives at most W) and
// keeps ~2^-62 W resolution; integer addition is associative, so the rhs is
// bitwise independent of thread order and launch shape (the reference's FP64
// atomics are not, projection_gpu.jl:7-9).  Each workgroup writes its private
// rhs to a partial row; one workgroup sums the rows exactly and converts back.
// The breakpoints are staged in LDS for the cell search.
// The mass solve exploits the tensor structure: M c = b is C = M1^{-1} B M2^{-1}
// for the nd1 x nd2 coefficient array B, i.e. nd2 banded (half-bandwidth 2)
// solves along direction 1 followed by nd1 along direction 2, all in LDS in a
// single workgroup.
#include "collop_launch.hpp"

namespace collop {
namespace {

constexpr int kSolveThreads = 1024;

// ---------------------------------------------------------------------------
// Scatter.  grid = (blocks, fields).  Dynamic LDS: n_dofs doubles.
__global__ __launch_bounds__(kScatterThreads) void k_scatter(MeshView m, const double* __restrict__ v0,
                                                             const double* __restrict__ vin,
                                                             const double* __restrict__ w, int N, int fields,
                                                             double scale, double* __restrict__ vmid,
                                                             double* __restrict__ dv,
                                                             long long* __restrict__ part,
                                                             double* __restrict__ mom) {
  extern __shared__ unsigned long long lds_u[];
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kScatterThreads / 64];
  const int field = blockIdx.y;
  unsigned long long* rhs = lds_u;
  for (int k = threadIdx.x; k < m.n_dofs; k += blockDim.x) rhs[k] = 0ull;
  stage_bp(m, sb1, sb2);
  __syncthreads();
  double p1 = 0.0, p2 = 0.0, e2 = 0.0;
  const int stride = gridDim.x * blockDim.x;
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < N; a += stride) {
    double x, y;
    if (fields == 2) {
      const double a1 = v0[a], a2 = v0[N + a], b1 = vin[a], b2 = vin[N + a];
      const double m1 = 0.5 * (a1 + b1), m2 = 0.5 * (a2 + b2);
      if (field == 0) {
        vmid[a] = m1;
        vmid[N + a] = m2;
        dv[a] = b1 - a1;
        dv[N + a] = b2 - a2;
        x = m1;
        y = m2;
      } else {
        x = b1;
        y = b2;
      }
    } else {
      x = vin[a];
      y = vin[N + a];
    }
    const double wa = w[a];
    if (mom && field == 0) {
      p1 += wa * x;
      p2 += wa * y;
      e2 += wa * (x * x + y * y);
    }
    Located L;
    if (!locate(m, sb1, sb2, x, y, L)) continue;
    double f1[3], f2[3];
    basis3(m.C1, L.i, L.xi1, f1);
    basis3(m.C2, L.j, L.xi2, f2);
    const int s1 = m.bs1[L.i], s2 = m.bs2[L.j];
    const double ws = wa * scale;
#pragma unroll
    for (int j2 = 0; j2 < 3; ++j2)
#pragma unroll
      for (int j1 = 0; j1 < 3; ++j1) {
        const long long q = __double2ll_rn(ws * f1[j1] * f2[j2]);
        atomicAdd(&rhs[(s1 + j1) + (s2 + j2) * m.nd1], (unsigned long long)q);
      }
  }
  __syncthreads();
  long long* out = part + ((size_t)field * gridDim.x + blockIdx.x) * m.n_dofs;
  for (int k = threadIdx.x; k < m.n_dofs; k += blockDim.x) out[k] = (long long)rhs[k];
  if (mom && field == 0) {
    const double s1 = block_sum<kScatterThreads>(p1, red);
    const double s2 = block_sum<kScatterThreads>(p2, red);
    const double s3 = block_sum<kScatterThreads>(e2, red);
    if (threadIdx.x == 0) {
      mom[3 * blockIdx.x + 0] = s1;
      mom[3 * blockIdx.x + 1] = s2;
      mom[3 * blockIdx.x + 2] = s3;
    }
  }
}

// Exact sum of the fixed-point partial rows of one field, back to FP64.
__device__ __forceinline__ double rhs_of(const long long* __restrict__ P, int nblocks, int nd, int k,
                                         double inv_scale) {
  unsigned long long s = 0ull;  // two's-complement wrap-around is exact
  for (int b = 0; b < nblocks; ++b) s += (unsigned long long)P[(size_t)b * nd + k];
  return (double)(long long)s * inv_scale;
}

// One workgroup per field.  X lives in dynamic LDS (n_dofs doubles).
__global__ __launch_bounds__(kSolveThreads) void k_reduce_solve(MeshView m, const long long* __restrict__ part,
                                                                int nblocks, double inv_scale,
                                                                const double* __restrict__ direct,
                                                                double* __restrict__ out) {
  extern __shared__ double X[];
  const int field = blockIdx.x;
  const int nd = m.n_dofs;
  if (part) {
    const long long* P = part + (size_t)field * nblocks * nd;
    for (int k = threadIdx.x; k < nd; k += blockDim.x) X[k] = rhs_of(P, nblocks, nd, k, inv_scale);
  } else {
    for (int k = threadIdx.x; k < nd; k += blockDim.x) X[k] = direct[(size_t)field * nd + k];
  }
  __syncthreads();
  // direction 1: one line per i2 (contiguous along i1)
  for (int l = threadIdx.x; l < m.nd2; l += blockDim.x) band_solve_line(X + (size_t)l * m.nd1, m.nd1, 1, m.chol1);
  __syncthreads();
  // direction 2: one line per i1 (stride nd1)
  for (int l = threadIdx.x; l < m.nd1; l += blockDim.x) band_solve_line(X + l, m.nd2, m.nd1, m.chol2);
  __syncthreads();
  for (int k = threadIdx.x; k < nd; k += blockDim.x) out[(size_t)field * nd + k] = X[k];
}

// ---------------------------------------------------------------------------
// Quadrature pass 1: one thread per (cell, Gauss point).  grid.y = field.
constexpr int kQuadThreads = 256;

__global__ __launch_bounds__(kQuadThreads) void k_quad1(MeshView m, const double* __restrict__ c0,
                                                        const double* __restrict__ c1, int use_logsq,
                                                        double* __restrict__ iota, double* __restrict__ spart) {
  __shared__ double red[kQuadThreads / 64];
  const int field = blockIdx.y;
  const double* c = field == 0 ? c0 : c1;
  const int npt = m.nq * m.nq;
  const int nel = m.ne1 * m.ne2;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  double s_loc = 0.0;
  if (tid < nel * npt) {
    const int e = tid / npt, pt = tid - e * npt;
    const int e1 = e % m.ne1, e2 = e / m.ne1;
    const int a = pt % m.nq, b = pt / m.nq;
    const int s1 = m.bs1[e1], s2 = m.bs2[e2];
    double f = 0.0;
#pragma unroll
    for (int j2 = 0; j2 < 3; ++j2) {
      const double p2 = m.phiq2[(3 * e2 + j2) * m.nq + b];
      double row = 0.0;
#pragma unroll
      for (int j1 = 0; j1 < 3; ++j1) row += c[(s1 + j1) + (s2 + j2) * m.nd1] * m.phiq1[(3 * e1 + j1) * m.nq + a];
      f += row * p2;
    }
    const double jac = (m.bp1[e1 + 1] - m.bp1[e1]) * (m.bp2[e2 + 1] - m.bp2[e2]);
    const double wq = m.qw[a] * m.qw[b] * jac;
    double lg = 0.0;
    bool ok;
    if (use_logsq) {
      ok = f * f > 1e-60;
      if (ok) lg = 0.5 * log(f * f);
    } else {
      ok = f > 1e-30;
      if (ok) lg = log(f);
    }
    if (iota && field == 0) iota[(size_t)e * npt + pt] = ok ? (1.0 + lg) * wq : 0.0;
    s_loc = ok ? -f * lg * wq : 0.0;
  }
  if (spart) {
    const double s = block_sum<kQuadThreads>(s_loc, red);
    if (threadIdx.x == 0) spart[(size_t)field * gridDim.x + blockIdx.x] = s;
  }
}

// Quadrature pass 2: one thread per dof, gathering over the cells of its support.
__global__ void k_quad2(MeshView m, const double* __restrict__ iota, double* __restrict__ r) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m.n_dofs) return;
  const int i1 = k % m.nd1, i2 = k / m.nd1;
  const int npt = m.nq * m.nq;
  double acc = 0.0;
  for (int e2 = m.erange2[2 * i2]; e2 <= m.erange2[2 * i2 + 1]; ++e2) {
    const int j2 = i2 - m.bs2[e2];
    const double* p2 = m.phiq2 + (3 * e2 + j2) * m.nq;
    for (int e1 = m.erange1[2 * i1]; e1 <= m.erange1[2 * i1 + 1]; ++e1) {
      const int j1 = i1 - m.bs1[e1];
      const double* p1 = m.phiq1 + (3 * e1 + j1) * m.nq;
      const double* io = iota + (size_t)(e1 + e2 * m.ne1) * npt;
      for (int b = 0; b < m.nq; ++b) {
        double row = 0.0;
        for (int a = 0; a < m.nq; ++a) row += p1[a] * io[a + m.nq * b];
        acc += row * p2[b];
      }
    }
  }
  r[k] = acc;
}

// rhs without the solve (collop_l2_rhs)
__global__ void k_rhs(const long long* __restrict__ part, int nblocks, int nd, double inv_scale,
                      double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nd) out[k] = rhs_of(part, nblocks, nd, k, inv_scale);
}

// out[i] = sum_j part[i * istride + j * jstride], j < n: one workgroup per
// output, fixed strided partition over threads, then the fixed block tree.
constexpr int kSumThreads = 256;

__global__ __launch_bounds__(kSumThreads) void k_sum_rows(const double* __restrict__ part, int n, int istride,
                                                         int jstride, double* __restrict__ out) {
  __shared__ double red[kSumThreads / 64];
  const int i = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += kSumThreads) s += part[(size_t)i * istride + (size_t)j * jstride];
  s = block_sum<kSumThreads>(s, red);
  if (threadIdx.x == 0) out[i] = s;
}

// ---------------------------------------------------------------------------
// Gathers.  mode 0: out = G (N x 2).  mode 1: out = dS = -w G, dpart[block] =
// (dv . dS, |dv|^2).  mode 2: out = clamp(grad f / max(|f|, fs_floor), +-g_max).
// mode 3: out = (f, |grad f|) at the marker (warm-start features).
constexpr int kGatherThreads = 256;

__global__ __launch_bounds__(kGatherThreads) void k_gather(MeshView m, const double* __restrict__ x,
                                                           const double* __restrict__ coeff,
                                                           const double* __restrict__ w,
                                                           const double* __restrict__ dv, int N, int mode,
                                                           double fs_floor, double g_max, double* __restrict__ out,
                                                           double* __restrict__ dpart) {
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kGatherThreads / 64];
  stage_bp(m, sb1, sb2);
  __syncthreads();
  double dot = 0.0, nrm = 0.0;
  const int stride = gridDim.x * blockDim.x;
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < N; a += stride) {
    const double v1 = x[a], v2 = x[N + a];
    double o1 = 0.0, o2 = 0.0;
    Located L;
    if (locate(m, sb1, sb2, v1, v2, L)) {
      double f1[3], d1[3], f2[3], d2[3];
      basis3_grad(m.C1, L.i, L.xi1, f1, d1);
      basis3_grad(m.C2, L.j, L.xi2, f2, d2);
      const double ih1 = 1.0 / L.h1, ih2 = 1.0 / L.h2;
      const int s1 = m.bs1[L.i], s2 = m.bs2[L.j];
      double acc1 = 0.0, acc2 = 0.0, fv = 0.0;
#pragma unroll
      for (int j2 = 0; j2 < 3; ++j2)
#pragma unroll
        for (int j1 = 0; j1 < 3; ++j1) {  // same (j2, j1) order as functions.jl:193-197
          const double c = coeff[(s1 + j1) + (s2 + j2) * m.nd1];
          acc1 += c * (d1[j1] * f2[j2]) * ih1;
          acc2 += c * (f1[j1] * d2[j2]) * ih2;
          if (mode >= 2) fv += c * (f1[j1] * f2[j2]);
        }
      if (mode == 3) {
        o1 = fv;
        o2 = sqrt(acc1 * acc1 + acc2 * acc2);
      } else if (mode == 2) {
        const double invf = 1.0 / fmax(fabs(fv), fs_floor);
        o1 = fmin(fmax(acc1 * invf, -g_max), g_max);
        o2 = fmin(fmax(acc2 * invf, -g_max), g_max);
      } else {
        o1 = acc1;
        o2 = acc2;
      }
    }
    if (mode == 1) {
      const double wa = w[a];
      o1 = -wa * o1;
      o2 = -wa * o2;
      const double e1 = dv[a], e2 = dv[N + a];
      dot += e1 * o1 + e2 * o2;
      nrm += e1 * e1 + e2 * e2;
    }
    out[a] = o1;
    out[N + a] = o2;
  }
  if (mode == 1) {
    const double s1 = block_sum<kGatherThreads>(dot, red);
    const double s2 = block_sum<kGatherThreads>(nrm, red);
    if (threadIdx.x == 0) {
      dpart[2 * blockIdx.x] = s1;
      dpart[2 * blockIdx.x + 1] = s2;
    }
  }
}

}  // namespace

int scatter_blocks(int N) {
  int b = (N + kScatterThreads * 4 - 1) / (kScatterThreads * 4);
  return b < 1 ? 1 : (b > 128 ? 128 : b);
}

double scatter_scale(double w_abs_sum) {
  return w_abs_sum > 0.0 ? 4611686018427387904.0 / (w_abs_sum * (1.0 + 1e-12)) : 1.0;  // 2^62 / W
}

void launch_scatter(const MeshView& m, const double* v0, const double* vin, const double* w, int N, int fields,
                    int grid_fields, double w_abs_sum, double* vmid, double* dv, long long* part, double* mom,
                    hipStream_t st) {
  const int nb = scatter_blocks(N);
  const size_t lds = (size_t)m.n_dofs * sizeof(unsigned long long);
  hipLaunchKernelGGL(k_scatter, dim3(nb, grid_fields), dim3(kScatterThreads), lds, st, m, v0, vin, w, N, fields,
                     scatter_scale(w_abs_sum), vmid, dv, part, mom);
}

void launch_reduce_solve(const MeshView& m, const long long* part, int nblocks, double w_abs_sum,
                         const double* direct, int fields, double* out, hipStream_t st) {
  const size_t lds = (size_t)m.n_dofs * sizeof(double);
  hipLaunchKernelGGL(k_reduce_solve, dim3(fields), dim3(kSolveThreads), lds, st, m, part, nblocks,
                     1.0 / scatter_scale(w_abs_sum), direct, out);
}

void launch_rhs(const MeshView& m, const long long* part, int nblocks, double w_abs_sum, double* out,
                hipStream_t st) {
  hipLaunchKernelGGL(k_rhs, dim3((m.n_dofs + 255) / 256), dim3(256), 0, st, part, nblocks, m.n_dofs,
                     1.0 / scatter_scale(w_abs_sum), out);
}

int quad_blocks(const MeshView& m) {
  const int tot = m.ne1 * m.ne2 * m.nq * m.nq;
  return (tot + kQuadThreads - 1) / kQuadThreads;
}

void launch_quad1(const MeshView& m, const double* c0, const double* c1, int fields, int use_logsq, double* iota,
                  double* spart, hipStream_t st) {
  hipLaunchKernelGGL(k_quad1, dim3(quad_blocks(m), fields), dim3(kQuadThreads), 0, st, m, c0, c1, use_logsq, iota,
                     spart);
}

void launch_quad2(const MeshView& m, const double* iota, double* r, hipStream_t st) {
  hipLaunchKernelGGL(k_quad2, dim3((m.n_dofs + 127) / 128), dim3(128), 0, st, m, iota, r);
}

void launch_sum_rows(const double* part, int n, int count, int istride, int jstride, double* out, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_sum_rows, dim3(count), dim3(kSumThreads), 0, st, part, n, istride, jstride, out);
}

int gather_blocks(int N) {
  int b = (N + kGatherThreads * 2 - 1) / (kGatherThreads * 2);
  return b < 1 ? 1 : (b > 512 ? 512 : b);
}

void launch_gather(const MeshView& m, const double* x, const double* coeff, const double* w, const double* dv,
                   int N, int mode, double fs_floor, double g_max, double* out, double* dpart, hipStream_t st) {
  hipLaunchKernelGGL(k_gather, dim3(gather_blocks(N)), dim3(kGatherThreads), 0, st, m, x, coeff, w, dv, N, mode,
                     fs_floor, g_max, out, dpart);
}

}  // namespace collop
