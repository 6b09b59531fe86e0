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
#include "lb_dev.hpp"

namespace collop {
namespace {


// ---------------------------------------------------------------------------
// Kernel bodies.  Each body serves the single-run kernel and the batched
// ensemble kernel (which maps its block to a (run, local block) pair first),
// so a run computes the same arithmetic, in the same order, in both.
// Marker arrays hold component 0 at [a] and component 1 at [ld + a].

// Scatter of one block: markers lb*NT + tid, stride nbk*NT; rhs in LDS.
__device__ __forceinline__ void scatter_body(const MeshView& m, unsigned long long* rhs, double* sb1, double* sb2,
                                             double* red, const double* __restrict__ v0,
                                             const double* __restrict__ vin, const double* __restrict__ w, int N,
                                             int ld, int fields, int field, double scale, double* __restrict__ vmid,
                                             double* __restrict__ dv, long long* __restrict__ out_row,
                                             double* __restrict__ mom_out, int lb, int nbk,
                                             const UpdateArgs* upd = nullptr, double* vin_out = nullptr) {
  for (int k = threadIdx.x; k < m.n_dofs; k += blockDim.x) rhs[k] = 0ull;
  stage_bp(m, sb1, sb2);
  __syncthreads();
  double p1 = 0.0, p2 = 0.0, e2 = 0.0;
  const int stride = nbk * blockDim.x;
  for (int a = lb * blockDim.x + threadIdx.x; a < N; a += stride) {
    double x, y;
    if (fields == 2) {
      const double a1 = v0[a], a2 = v0[ld + a];
      double b1 = vin[a], b2 = vin[ld + a];
      if (upd) {  // the pending Anderson update of the input, written once (field 0) to vin_out
        const size_t cs = (size_t)2 * N;
        double x1 = upd->Gv[a], x2 = upd->Gv[ld + a];
        for (int j = 0; j < upd->hist; ++j) {
          x1 -= upd->dG[j * cs + a] * upd->g[j];
          x2 -= upd->dG[j * cs + ld + a] * upd->g[j];
        }
        b1 = dot2(upd->beta, x1, 1.0 - upd->beta, b1);
        b2 = dot2(upd->beta, x2, 1.0 - upd->beta, b2);
        if (field == 0) {
          vin_out[a] = b1;
          vin_out[ld + a] = b2;
        }
      }
      const double m1 = 0.5 * (a1 + b1), m2 = 0.5 * (a2 + b2);
      if (field == 0) {
        vmid[a] = m1;
        vmid[ld + a] = m2;
        dv[a] = b1 - a1;
        dv[ld + a] = b2 - a2;
        x = m1;
        y = m2;
      } else {
        x = b1;
        y = b2;
      }
    } else {
      x = vin[a];
      y = vin[ld + a];
    }
    const double wa = w[a];
    if (mom_out && field == 0) {
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
  for (int k = threadIdx.x; k < m.n_dofs; k += blockDim.x) out_row[k] = (long long)rhs[k];
  if (mom_out && field == 0) {
    const double s1 = block_sum<kScatterThreads>(p1, red);
    const double s2 = block_sum<kScatterThreads>(p2, red);
    const double s3 = block_sum<kScatterThreads>(e2, red);
    if (threadIdx.x == 0) {
      mom_out[0] = s1;
      mom_out[1] = s2;
      mom_out[2] = s3;
    }
  }
}

// Exact sum of the fixed-point partial rows of one field, back to FP64.
__device__ __forceinline__ double rhs_of(const long long* __restrict__ P, int nblocks, int nd, int k,
                                         double inv_scale) {
  // two's-complement wrap-around is exact, in any order: four independent
  // partial sums keep 32 loads in flight per thread
  unsigned long long s0 = 0ull, s1 = 0ull, s2 = 0ull, s3 = 0ull;
  int b = 0;
  for (; b + 32 <= nblocks; b += 32) {
#pragma unroll
    for (int j = 0; j < 32; j += 4) {
      s0 += (unsigned long long)P[(size_t)(b + j) * nd + k];
      s1 += (unsigned long long)P[(size_t)(b + j + 1) * nd + k];
      s2 += (unsigned long long)P[(size_t)(b + j + 2) * nd + k];
      s3 += (unsigned long long)P[(size_t)(b + j + 3) * nd + k];
    }
  }
  for (; b < nblocks; ++b) s0 += (unsigned long long)P[(size_t)b * nd + k];
  return (double)(long long)(s0 + s1 + s2 + s3) * inv_scale;
}

// rhs (exact row sum, or `direct`) -> M^{-1} rhs -> out.  One workgroup of
// kSolveThreads.  M = M1 (x) M2, so the coefficient array (i1 fastest) is
// C = M1^{-1} B M2^{-1}: two small dense products with the precomputed
// symmetric 1D inverses, phase one held in registers across a barrier so B
// and M1^{-1} B share one LDS array of n_dofs doubles.  The inverses are read
// from LDS copies when `stage` (they fit beside X), else from global memory.
constexpr int kSolveThreads = 1024;
constexpr int kSolvePer = 8;  // n_dofs <= kSolvePer * kSolveThreads

// LDS doubles of a solve workgroup: X, and the two inverses when staged.
__host__ __device__ inline int solve_stage(const MeshView& m) {
  return (m.n_dofs + m.nd1 * m.nd1 + m.nd2 * m.nd2) * 8 <= 48 * 1024;
}
inline size_t solve_lds(const MeshView& m) {
  return (size_t)(m.n_dofs + (solve_stage(m) ? m.nd1 * m.nd1 + m.nd2 * m.nd2 : 0)) * sizeof(double);
}

// X (LDS, n_dofs) holds B on entry; out <- M1^{-1} B M2^{-1}.
__device__ __forceinline__ void solve_core(const MeshView& m, double* X, double* __restrict__ out) {
  const int nd = m.n_dofs, nd1 = m.nd1, nd2 = m.nd2;
  const double* I1 = m.minv1;
  const double* I2 = m.minv2;
  if (solve_stage(m)) {  // X is followed by room for both inverses
    double* L1 = X + nd;
    double* L2 = L1 + nd1 * nd1;
    for (int k = threadIdx.x; k < nd1 * nd1; k += blockDim.x) L1[k] = I1[k];
    for (int k = threadIdx.x; k < nd2 * nd2; k += blockDim.x) L2[k] = I2[k];
    I1 = L1;
    I2 = L2;
  }
  __syncthreads();
  double t[kSolvePer];
#pragma unroll
  for (int q = 0; q < kSolvePer; ++q) {
    const int k = threadIdx.x + q * kSolveThreads;
    t[q] = 0.0;
    if (k < nd) {
      const int i1 = k % nd1, i2 = k - (k % nd1);
      const double* row = I1 + (size_t)i1 * nd1;
      double s = 0.0;
      for (int j = 0; j < nd1; ++j) s = __builtin_fma(row[j], X[i2 + j], s);
      t[q] = s;
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kSolvePer; ++q) {
    const int k = threadIdx.x + q * kSolveThreads;
    if (k < nd) X[k] = t[q];
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nd; k += blockDim.x) {
    const int i1 = k % nd1, i2 = k / nd1;
    const double* col = I2 + (size_t)i2 * nd2;
    double s = 0.0;
    for (int j = 0; j < nd2; ++j) s = __builtin_fma(X[i1 + (size_t)nd1 * j], col[j], s);
    out[k] = s;
  }
}

__device__ __forceinline__ void solve_body(const MeshView& m, double* X, const long long* __restrict__ P, int nblocks,
                                           double inv_scale, const double* __restrict__ direct,
                                           double* __restrict__ out) {
  const int nd = m.n_dofs;
  if (P) {
    for (int k = threadIdx.x; k < nd; k += blockDim.x) X[k] = rhs_of(P, nblocks, nd, k, inv_scale);
  } else {
    for (int k = threadIdx.x; k < nd; k += blockDim.x) X[k] = direct[k];
  }
  solve_core(m, X, out);
}

// Quadrature pass 1 of one block (one thread per cell and Gauss point).
constexpr int kQuadThreads = 256;

// f_s = sum_k c_k phi_k at Gauss point (a, b) of cell (e1, e2): rows over j1
// inside, j2 outside.
__device__ __forceinline__ double field_at_qp(const MeshView& m, const double* __restrict__ c, int e1, int e2, int a,
                                              int b) {
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
  return f;
}

__device__ __forceinline__ void quad1_body(const MeshView& m, double* red, const double* __restrict__ c, int use_logsq,
                                           double* __restrict__ iota, double* __restrict__ spart_out, int lb) {
  const int npt = m.nq * m.nq;
  const int nel = m.ne1 * m.ne2;
  const int tid = lb * blockDim.x + threadIdx.x;
  double s_loc = 0.0;
  if (tid < nel * npt) {
    const int e = tid / npt, pt = tid - e * npt;
    const int e1 = e % m.ne1, e2 = e / m.ne1;
    const int a = pt % m.nq, b = pt / m.nq;
    const double f = field_at_qp(m, c, e1, e2, a, b);
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
    if (iota) iota[(size_t)e * npt + pt] = ok ? (1.0 + lg) * wq : 0.0;
    s_loc = ok ? -f * lg * wq : 0.0;
  }
  if (spart_out) {
    const double s = block_sum<kQuadThreads>(s_loc, red);
    if (threadIdx.x == 0) spart_out[0] = s;
  }
}

// The pass-1 integrand (1 + log f) w_q |e| at Gauss point (a, b) of cell (e1, e2)
// (0 where f is below the floor), with f = sum c phi.
__device__ __forceinline__ double iota_at(const MeshView& m, const double* __restrict__ c, int e1, int e2, int a,
                                          int b, int use_logsq) {
  const double f = field_at_qp(m, c, e1, e2, a, b);
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
  return ok ? (1.0 + lg) * wq : 0.0;
}

// Quadrature pass 2 for dof k with the pass-1 integrand evaluated in place
// from the coefficients c (same points, same order as quad2_wave).
__device__ __forceinline__ void quad_r_wave(const MeshView& m, const double* __restrict__ c, int use_logsq,
                                            double* __restrict__ r, int k, int lane) {
  if (k >= m.n_dofs) return;  // wave-uniform
  const int i1 = k % m.nd1, i2 = k / m.nd1;
  const int npt = m.nq * m.nq;
  const int e1lo = m.erange1[2 * i1], n1 = m.erange1[2 * i1 + 1] - e1lo + 1;
  const int e2lo = m.erange2[2 * i2], n2 = m.erange2[2 * i2 + 1] - e2lo + 1;
  const int total = n1 * n2 * npt;
  double acc = 0.0;
  for (int idx = lane; idx < total; idx += 64) {
    const int cell = idx / npt, pt = idx - cell * npt;
    const int e2 = e2lo + cell / n1, e1 = e1lo + cell % n1;
    const int a = pt % m.nq, b = pt / m.nq;
    const double p1 = m.phiq1[(3 * e1 + i1 - m.bs1[e1]) * m.nq + a];
    const double p2 = m.phiq2[(3 * e2 + i2 - m.bs2[e2]) * m.nq + b];
    acc += p1 * iota_at(m, c, e1, e2, a, b, use_logsq) * p2;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) r[k] = acc;
}

// Quadrature pass 2 for dof k, one wave: the lanes stride over the (cell,
// point) pairs of the dof's support (at most 3 x 3 cells x nq^2 points), then
// a fixed shuffle tree.  Lane 0 writes r[k].
__device__ __forceinline__ void quad2_wave(const MeshView& m, const double* __restrict__ iota, double* __restrict__ r,
                                           int k, int lane) {
  if (k >= m.n_dofs) return;  // wave-uniform
  const int i1 = k % m.nd1, i2 = k / m.nd1;
  const int npt = m.nq * m.nq;
  const int e1lo = m.erange1[2 * i1], n1 = m.erange1[2 * i1 + 1] - e1lo + 1;
  const int e2lo = m.erange2[2 * i2], n2 = m.erange2[2 * i2 + 1] - e2lo + 1;
  const int total = n1 * n2 * npt;
  double acc = 0.0;
  for (int idx = lane; idx < total; idx += 64) {
    const int cell = idx / npt, pt = idx - cell * npt;
    const int e2 = e2lo + cell / n1, e1 = e1lo + cell % n1;
    const int a = pt % m.nq, b = pt / m.nq;
    const double p1 = m.phiq1[(3 * e1 + i1 - m.bs1[e1]) * m.nq + a];
    const double p2 = m.phiq2[(3 * e2 + i2 - m.bs2[e2]) * m.nq + b];
    acc += p1 * iota[(size_t)(e1 + e2 * m.ne1) * npt + pt] * p2;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) r[k] = acc;
}

// sum_{j < n} base[j * jstride]: fixed strided partition over threads, then the
// fixed block tree.  Result in thread 0.
constexpr int kSumThreads = 256;

__device__ __forceinline__ double sum_body(const double* __restrict__ base, int n, int jstride, double* red) {
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += kSumThreads) s += base[(size_t)j * jstride];
  return block_sum<kSumThreads>(s, red);
}

// Gathers.  mode 0: out = G.  mode 1: out = dS = -w G, dpart_out = (dv . dS,
// |dv|^2) of the block.  mode 2: out = clamp(grad f / max(|f|, fs_floor),
// +-g_max).  mode 3: out = (f, |grad f|) at the marker (warm-start features).
constexpr int kGatherThreads = 256;

__device__ __forceinline__ void gather_body(const MeshView& m, double* sb1, double* sb2, double* red,
                                            const double* __restrict__ x, int ld, const double* __restrict__ coeff,
                                            const double* __restrict__ w, const double* __restrict__ dv, int N,
                                            int mode, double fs_floor, double g_max, double* __restrict__ out,
                                            double* __restrict__ dpart_out, int lb, int nbk) {
  stage_bp(m, sb1, sb2);
  __syncthreads();
  double dot = 0.0, nrm = 0.0;
  const int stride = nbk * blockDim.x;
  for (int a = lb * blockDim.x + threadIdx.x; a < N; a += stride) {
    const double v1 = x[a], v2 = x[ld + a];
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
      const double e1 = dv[a], e2 = dv[ld + a];
      dot += e1 * o1 + e2 * o2;
      nrm += e1 * e1 + e2 * e2;
    }
    out[a] = o1;
    out[ld + a] = o2;
  }
  if (mode == 1) {
    const double s1 = block_sum<kGatherThreads>(dot, red);
    const double s2 = block_sum<kGatherThreads>(nrm, red);
    if (threadIdx.x == 0) {
      dpart_out[0] = s1;
      dpart_out[1] = s2;
    }
  }
}

// ---------------------------------------------------------------------------
// Single-run kernels.

// grid = (blocks, fields).  Dynamic LDS: n_dofs u64.
__global__ __launch_bounds__(kScatterThreads) void k_scatter(MeshView m, const double* __restrict__ v0,
                                                             const double* __restrict__ vin,
                                                             const double* __restrict__ w, int N, int fields,
                                                             double scale, double* __restrict__ vmid,
                                                             double* __restrict__ dv,
                                                             long long* __restrict__ part,
                                                             double* __restrict__ mom, UpdateArgs U, int has_upd,
                                                             double* __restrict__ vin_out) {
  extern __shared__ unsigned long long lds_u[];
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kScatterThreads / 64];
  const int field = blockIdx.y;
  scatter_body(m, lds_u, sb1, sb2, red, v0, vin, w, N, N, fields, field, scale, vmid, dv,
               part + ((size_t)field * gridDim.x + blockIdx.x) * m.n_dofs, mom ? mom + 3 * blockIdx.x : nullptr,
               blockIdx.x, gridDim.x, has_upd ? &U : nullptr, vin_out);
}

// One workgroup per field.  X lives in dynamic LDS (solve_lds).
__global__ __launch_bounds__(kSolveThreads) void k_reduce_solve(MeshView m, const long long* __restrict__ part,
                                                                int nblocks, double inv_scale,
                                                                const double* __restrict__ direct,
                                                                double* __restrict__ out) {
  extern __shared__ double X[];
  const int field = blockIdx.x;
  const int nd = m.n_dofs;
  solve_body(m, X, part ? part + (size_t)field * nblocks * nd : nullptr, nblocks, inv_scale,
             direct ? direct + (size_t)field * nd : nullptr, out + (size_t)field * nd);
}

// The same solve split over many workgroups, in the same arithmetic order per
// coefficient (so results equal solve_body's bit for bit): k_solve_cols takes
// one column i2 of B per workgroup -- the exact row sum (or `direct`) of its
// nd1 entries, then T = M1^{-1} B on that column -- and k_solve_rows one row
// i1 of T -- out = T M2^{-1} on that row.  Grid (nd2 [+1], fields) and (nd1,
// fields) of kSplitThreads.  The optional extra workgroup of k_solve_cols
// (blockIdx.x == nd2, field 0) sums sum_n partials into *sum_out exactly as
// k_sum_rows does.
constexpr int kSplitThreads = kSumThreads;

__global__ __launch_bounds__(kSplitThreads) void k_solve_cols(MeshView m, const long long* __restrict__ part,
                                                              int nblocks, double inv_scale,
                                                              const double* __restrict__ direct,
                                                              double* __restrict__ T,
                                                              const double* __restrict__ sum_part, int sum_n,
                                                              double* __restrict__ sum_out) {
  __shared__ double B[kSplitThreads];
  __shared__ double red[kSumThreads / 64];
  const int nd = m.n_dofs, nd1 = m.nd1, field = blockIdx.y, i2 = blockIdx.x;
  if (i2 == m.nd2) {  // the fused partial sum
    const double t = sum_body(sum_part, sum_n, 1, red);
    if (threadIdx.x == 0) *sum_out = t;
    return;
  }
  const int k = i2 * nd1 + threadIdx.x;
  if (threadIdx.x < nd1)
    B[threadIdx.x] = part ? rhs_of(part + (size_t)field * nblocks * nd, nblocks, nd, k, inv_scale)
                          : direct[(size_t)field * nd + k];
  __syncthreads();
  if (threadIdx.x < nd1) {
    const double* row = m.minv1 + (size_t)threadIdx.x * nd1;
    double s = 0.0;
    for (int j = 0; j < nd1; ++j) s = __builtin_fma(row[j], B[j], s);
    T[(size_t)field * nd + k] = s;
  }
}

__global__ __launch_bounds__(kSplitThreads) void k_solve_rows(MeshView m, const double* __restrict__ T,
                                                              double* __restrict__ out) {
  __shared__ double Trow[kSplitThreads];
  const int nd = m.n_dofs, nd1 = m.nd1, nd2 = m.nd2, field = blockIdx.y, i1 = blockIdx.x;
  if (threadIdx.x < nd2) Trow[threadIdx.x] = T[(size_t)field * nd + i1 + (size_t)nd1 * threadIdx.x];
  __syncthreads();
  if (threadIdx.x < nd2) {
    const double* col = m.minv2 + (size_t)threadIdx.x * nd2;
    double s = 0.0;
    for (int j = 0; j < nd2; ++j) s = __builtin_fma(Trow[j], col[j], s);
    out[(size_t)field * nd + i1 + (size_t)nd1 * threadIdx.x] = s;
  }
}

__global__ __launch_bounds__(kQuadThreads) void k_quad1(MeshView m, const double* __restrict__ c0,
                                                        const double* __restrict__ c1, int use_logsq,
                                                        double* __restrict__ iota, double* __restrict__ spart) {
  __shared__ double red[kQuadThreads / 64];
  const int field = blockIdx.y;
  quad1_body(m, red, field == 0 ? c0 : c1, use_logsq, field == 0 ? iota : nullptr,
             spart ? spart + (size_t)field * gridDim.x + blockIdx.x : nullptr, blockIdx.x);
}

// grid: (n_dofs + 3) / 4 blocks of pass-2 waves, then, when c1 is given,
// quad_blocks(m) blocks of pass 1 over c1 for the partials of S(c1).
__global__ __launch_bounds__(256) void k_quad_r(MeshView m, const double* __restrict__ c0,
                                                const double* __restrict__ c1, int use_logsq, double* __restrict__ r,
                                                double* __restrict__ spart) {
  const int nr = (m.n_dofs + 3) / 4;
  if ((int)blockIdx.x < nr) {
    quad_r_wave(m, c0, use_logsq, r, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
    return;
  }
  __shared__ double red[kQuadThreads / 64];
  const int lb = blockIdx.x - nr;
  quad1_body(m, red, c1, use_logsq, nullptr, spart + lb, lb);
}

// The Lenard-Bernstein gather of the Anderson map (lb_gather_body over nb
// blocks), then, when cin is given, quad_blocks(m) blocks of quadrature pass 1
// over cin for the partials qpart of the Gonzalez entropy S(f_in).
__global__ __launch_bounds__(kLbThreads) void k_lb_gather_q(MeshView m, const double* __restrict__ x,
                                                            const double* __restrict__ coeff,
                                                            const double* __restrict__ cin, int use_logsq,
                                                            const double* __restrict__ w,
                                                            const double* __restrict__ dv, int N, int nb,
                                                            double* __restrict__ g, double* __restrict__ part,
                                                            double* __restrict__ qpart) {
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kLbSums * kLbThreads / 64];
  if ((int)blockIdx.x < nb) {
    lb_gather_body(m, x, coeff, w, dv, N, 1e-30, 100.0, g, part, sb1, sb2, red, blockIdx.x, nb);
    return;
  }
  const int lb = blockIdx.x - nb;
  quad1_body(m, red, cin, use_logsq, nullptr, qpart + lb, lb);
}

__global__ __launch_bounds__(256) void k_quad2(MeshView m, const double* __restrict__ iota, double* __restrict__ r) {
  quad2_wave(m, iota, r, blockIdx.x * 4 + (threadIdx.x >> 6), threadIdx.x & 63);
}

// rhs without the solve (collop_l2_rhs)
__global__ void k_rhs(const long long* __restrict__ part, int nblocks, int nd, double inv_scale,
                      double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nd) out[k] = rhs_of(part, nblocks, nd, k, inv_scale);
}

// out[i] = sum_j part[i * istride + j * jstride], j < n: one workgroup per output.
__global__ __launch_bounds__(kSumThreads) void k_sum_rows(const double* __restrict__ part, int n, int istride,
                                                         int jstride, double* __restrict__ out) {
  __shared__ double red[kSumThreads / 64];
  const int i = blockIdx.x;
  const double s = sum_body(part + (size_t)i * istride, n, jstride, red);
  if (threadIdx.x == 0) out[i] = s;
}

__global__ __launch_bounds__(kGatherThreads) void k_gather(MeshView m, const double* __restrict__ x,
                                                           const double* __restrict__ coeff,
                                                           const double* __restrict__ w,
                                                           const double* __restrict__ dv, int N, int mode,
                                                           double fs_floor, double g_max, double* __restrict__ out,
                                                           double* __restrict__ dpart) {
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kGatherThreads / 64];
  gather_body(m, sb1, sb2, red, x, N, coeff, w, dv, N, mode, fs_floor, g_max, out,
              mode == 1 ? dpart + 2 * blockIdx.x : nullptr, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Run-record diagnostics (functions.jl:300-343).

// Cell histogram of the marker weights, m_e = sum of w_a over the markers that
// locate_particle puts in cell e (functions.jl:326-330), as fixed-point partial
// rows like the scatter's: exact and independent of thread order.
// Dynamic LDS: ne1 * ne2 u64.
__global__ __launch_bounds__(kScatterThreads) void k_cell_mass(MeshView m, const double* __restrict__ v,
                                                               const double* __restrict__ w, int N, double scale,
                                                               long long* __restrict__ part) {
  extern __shared__ unsigned long long hist[];
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  const int nel = m.ne1 * m.ne2;
  for (int k = threadIdx.x; k < nel; k += blockDim.x) hist[k] = 0ull;
  stage_bp(m, sb1, sb2);
  __syncthreads();
  for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < N; a += gridDim.x * blockDim.x) {
    Located L;
    if (!locate(m, sb1, sb2, v[a], v[N + a], L)) continue;
    atomicAdd(&hist[L.i + L.j * m.ne1], (unsigned long long)__double2ll_rn(w[a] * scale));
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nel; k += blockDim.x) part[(size_t)blockIdx.x * nel + k] = (long long)hist[k];
}

// One thread per (cell, Gauss point): f = f_s there, then the block sums of
// max(-f, 0) w_q |e| (compute_negative_part_l1) and, when the cell masses are
// given, (f - m_e / |e|)^2 w_q |e| (compute_fs_minus_fp_l2) into spart[b] and
// spart[gridDim.x + b].
__global__ __launch_bounds__(kQuadThreads) void k_diag_quad(MeshView m, const double* __restrict__ c,
                                                            const double* __restrict__ cell_mass,
                                                            double* __restrict__ spart) {
  __shared__ double red[kQuadThreads / 64];
  const int npt = m.nq * m.nq;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  double neg = 0.0, sq = 0.0;
  if (tid < m.ne1 * m.ne2 * npt) {
    const int e = tid / npt, pt = tid - e * npt;
    const int e1 = e % m.ne1, e2 = e / m.ne1;
    const int a = pt % m.nq, b = pt / m.nq;
    const double f = field_at_qp(m, c, e1, e2, a, b);
    const double jac = (m.bp1[e1 + 1] - m.bp1[e1]) * (m.bp2[e2 + 1] - m.bp2[e2]);
    const double wq = m.qw[a] * m.qw[b];
    if (f < 0.0) neg = (-f) * wq * jac;
    if (cell_mass) {
      const double d = f - cell_mass[e] / jac;
      sq = d * d * wq * jac;
    }
  }
  const double s1 = block_sum<kQuadThreads>(neg, red);
  const double s2 = block_sum<kQuadThreads>(sq, red);
  if (threadIdx.x == 0) {
    spart[blockIdx.x] = s1;
    spart[gridDim.x + blockIdx.x] = s2;
  }
}

// ---------------------------------------------------------------------------
// Batched ensemble kernels.  Block tables map blockIdx.x to (run, local block);
// runs whose control block says inactive return before any barrier.
__device__ __forceinline__ bool ens_skip(const EnsView& E, int r) { return E.ctl && !E.ctl[r].active; }

// grid = (sum of scatter blocks, grid_fields).  field 1 reads vin.
__global__ __launch_bounds__(kScatterThreads) void k_scatter_b(MeshView m, EnsView E, const double* __restrict__ v0,
                                                               const double* __restrict__ vin,
                                                               const double* __restrict__ w, int fields,
                                                               double* __restrict__ vmid, double* __restrict__ dv,
                                                               long long* __restrict__ part,
                                                               double* __restrict__ mom) {
  const int2 t = E.sblk[blockIdx.x];
  if (ens_skip(E, t.x)) return;
  extern __shared__ unsigned long long lds_u[];
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kScatterThreads / 64];
  const EnsRun& R = E.runs[t.x];
  const int field = blockIdx.y;
  const int row = R.sblk0 + t.y;
  scatter_body(m, lds_u, sb1, sb2, red, v0 ? v0 + R.off : nullptr, vin + R.off, w + R.off, R.n, E.NT, fields, field,
               R.scale, vmid ? vmid + R.off : nullptr, dv ? dv + R.off : nullptr,
               part + ((size_t)field * E.n_sblk + row) * m.n_dofs, mom ? mom + 3 * row : nullptr, t.y, R.nsb);
}

// grid = (R, fields).  From scatter rows (part != nullptr) or from
// vec[(r * 4 + in_slot + field) * n_dofs]; out to vec[(r * 4 + out_slot + field) * n_dofs]
// (out_slot < 0: to fc[r * n_dofs]).
__global__ __launch_bounds__(kSolveThreads) void k_solve_b(MeshView m, EnsView E, const long long* __restrict__ part,
                                                           int in_slot, int out_slot, double* __restrict__ vec,
                                                           double* __restrict__ fc) {
  const int r = blockIdx.x, field = blockIdx.y;
  if (ens_skip(E, r)) return;
  extern __shared__ double X[];
  const EnsRun& R = E.runs[r];
  const int nd = m.n_dofs;
  double* out = out_slot >= 0 ? vec + ((size_t)r * 4 + out_slot + field) * nd : fc + (size_t)r * nd;
  solve_body(m, X, part ? part + ((size_t)field * E.n_sblk + R.sblk0) * nd : nullptr, R.nsb, R.inv_scale,
             part ? nullptr : vec + ((size_t)r * 4 + in_slot + field) * nd, out);
}

// grid = (quad blocks, fields, R).  Field y reads c + r * rstride + y * nd.
__global__ __launch_bounds__(kQuadThreads) void k_quad1_b(MeshView m, EnsView E, const double* __restrict__ c,
                                                          int rstride, int use_logsq, double* __restrict__ iota,
                                                          double* __restrict__ spart) {
  const int r = blockIdx.z, field = blockIdx.y;
  if (ens_skip(E, r)) return;
  __shared__ double red[kQuadThreads / 64];
  const size_t npts = (size_t)m.ne1 * m.ne2 * m.nq * m.nq;
  quad1_body(m, red, c + (size_t)r * rstride + (size_t)field * m.n_dofs, use_logsq,
             (iota && field == 0) ? iota + r * npts : nullptr,
             spart ? spart + ((size_t)r * 2 + field) * gridDim.x + blockIdx.x : nullptr, blockIdx.x);
}

// grid = (dof blocks of 4, R): r_out = vec[(r * 4 + 2) * n_dofs]
__global__ __launch_bounds__(256) void k_quad2_b(MeshView m, EnsView E, const double* __restrict__ iota,
                                                 double* __restrict__ vec) {
  const int r = blockIdx.y;
  if (ens_skip(E, r)) return;
  const size_t npts = (size_t)m.ne1 * m.ne2 * m.nq * m.nq;
  quad2_wave(m, iota + r * npts, vec + ((size_t)r * 4 + 2) * m.n_dofs, blockIdx.x * 4 + (threadIdx.x >> 6),
             threadIdx.x & 63);
}

// grid = (R, count): out[r * ostride + c] = sum_{j < cnt[r]} part[row0[r] + c * istride + j * jstride]
__global__ __launch_bounds__(kSumThreads) void k_sum_rows_b(EnsView E, const double* __restrict__ part,
                                                           const int* __restrict__ row0, const int* __restrict__ cnt,
                                                           int istride, int jstride, double* __restrict__ out,
                                                           int ostride) {
  const int r = blockIdx.x, c = blockIdx.y;
  if (ens_skip(E, r)) return;
  __shared__ double red[kSumThreads / 64];
  const double s = sum_body(part + (size_t)row0[r] + (size_t)c * istride, cnt[r], jstride, red);
  if (threadIdx.x == 0) out[(size_t)r * ostride + c] = s;
}

// grid = sum of gather blocks.  coeff = vec[(r * 4 + slot) * n_dofs].
__global__ __launch_bounds__(kGatherThreads) void k_gather_b(MeshView m, EnsView E, const double* __restrict__ x,
                                                             const double* __restrict__ vec, int slot,
                                                             const double* __restrict__ w,
                                                             const double* __restrict__ dv, int mode,
                                                             double* __restrict__ out, double* __restrict__ dpart) {
  const int2 t = E.gblk[blockIdx.x];
  if (ens_skip(E, t.x)) return;
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kGatherThreads / 64];
  const EnsRun& R = E.runs[t.x];
  gather_body(m, sb1, sb2, red, x + R.off, E.NT, vec + ((size_t)t.x * 4 + slot) * m.n_dofs, w ? w + R.off : nullptr,
              dv ? dv + R.off : nullptr, R.n, mode, 0.0, 0.0, out + R.off,
              mode == 1 ? dpart + 2 * (R.gblk0 + t.y) : nullptr, t.y, R.ngb);
}

}  // namespace


double scatter_scale(double w_abs_sum) {
  return w_abs_sum > 0.0 ? 4611686018427387904.0 / (w_abs_sum * (1.0 + 1e-12)) : 1.0;  // 2^62 / W
}

void launch_scatter(const MeshView& m, const double* v0, const double* vin, const double* w, int N, int fields,
                    int grid_fields, double w_abs_sum, double* vmid, double* dv, long long* part, double* mom,
                    hipStream_t st, const UpdateArgs* upd, double* vin_out) {
  const int nb = scatter_blocks(N);
  const size_t lds = (size_t)m.n_dofs * sizeof(unsigned long long);
  UpdateArgs U{};
  if (upd) U = *upd;
  hipLaunchKernelGGL(k_scatter, dim3(nb, grid_fields), dim3(kScatterThreads), lds, st, m, v0, vin, w, N, fields,
                     scatter_scale(w_abs_sum), vmid, dv, part, mom, U, upd ? 1 : 0, vin_out);
}

void launch_reduce_solve(const MeshView& m, const long long* part, int nblocks, double w_abs_sum,
                         const double* direct, int fields, double* out, hipStream_t st) {
  const size_t lds = solve_lds(m);
  hipLaunchKernelGGL(k_reduce_solve, dim3(fields), dim3(kSolveThreads), lds, st, m, part, nblocks,
                     1.0 / scatter_scale(w_abs_sum), direct, out);
}

void launch_solve_split(const MeshView& m, const long long* part, int nblocks, double w_abs_sum,
                        const double* direct, int fields, double* out, double* tmp, hipStream_t st,
                        const double* sum_part, int sum_n, double* sum_out) {
  const int extra = sum_out ? 1 : 0;
  hipLaunchKernelGGL(k_solve_cols, dim3(m.nd2 + extra, fields), dim3(kSplitThreads), 0, st, m, part, nblocks,
                     1.0 / scatter_scale(w_abs_sum), direct, tmp, sum_part, sum_n, sum_out);
  hipLaunchKernelGGL(k_solve_rows, dim3(m.nd1, fields), dim3(kSplitThreads), 0, st, m, tmp, out);
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

void launch_lb_gather_q(const MeshView& m, const double* x, const double* coeff, const double* cin, int use_logsq,
                        const double* w, const double* dv, int N, double* g, double* part, double* qpart,
                        hipStream_t st) {
  static_assert(kQuadThreads == kLbThreads, "one block size for both halves of k_lb_gather_q");
  const int nb = lb_blocks(N);
  hipLaunchKernelGGL(k_lb_gather_q, dim3(nb + (cin ? quad_blocks(m) : 0)), dim3(kLbThreads), 0, st, m, x, coeff, cin,
                     use_logsq, w, dv, N, nb, g, part, qpart);
}

void launch_quad1(const MeshView& m, const double* c0, const double* c1, int fields, int use_logsq, double* iota,
                  double* spart, hipStream_t st) {
  hipLaunchKernelGGL(k_quad1, dim3(quad_blocks(m), fields), dim3(kQuadThreads), 0, st, m, c0, c1, use_logsq, iota,
                     spart);
}

void launch_quad2(const MeshView& m, const double* iota, double* r, hipStream_t st) {
  hipLaunchKernelGGL(k_quad2, dim3((m.n_dofs + 3) / 4), dim3(256), 0, st, m, iota, r);
}

void launch_quad_r(const MeshView& m, const double* c0, const double* c1, int use_logsq, double* r, double* spart,
                   hipStream_t st) {
  const int nb = (m.n_dofs + 3) / 4 + (c1 ? quad_blocks(m) : 0);
  hipLaunchKernelGGL(k_quad_r, dim3(nb), dim3(256), 0, st, m, c0, c1, use_logsq, r, spart);
}

void launch_sum_rows(const double* part, int n, int count, int istride, int jstride, double* out, hipStream_t st) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_sum_rows, dim3(count), dim3(kSumThreads), 0, st, part, n, istride, jstride, out);
}

void launch_fs_diagnostics(const MeshView& m, const double* v, const double* w, int N, double w_abs_sum,
                           const double* c, long long* part, double* cell_mass, double* spart, double* out,
                           hipStream_t st) {
  const int nel = m.ne1 * m.ne2, qb = quad_blocks(m);
  if (v) {
    const int nb = scatter_blocks(N);
    hipLaunchKernelGGL(k_cell_mass, dim3(nb), dim3(kScatterThreads), (size_t)nel * sizeof(unsigned long long), st,
                       m, v, w, N, scatter_scale(w_abs_sum), part);
    hipLaunchKernelGGL(k_rhs, dim3((nel + 255) / 256), dim3(256), 0, st, part, nb, nel,
                       1.0 / scatter_scale(w_abs_sum), cell_mass);
  }
  hipLaunchKernelGGL(k_diag_quad, dim3(qb), dim3(kQuadThreads), 0, st, m, c, v ? cell_mass : nullptr, spart);
  hipLaunchKernelGGL(k_sum_rows, dim3(2), dim3(kSumThreads), 0, st, spart, qb, qb, 1, out);
}


void launch_gather(const MeshView& m, const double* x, const double* coeff, const double* w, const double* dv,
                   int N, int mode, double fs_floor, double g_max, double* out, double* dpart, hipStream_t st) {
  hipLaunchKernelGGL(k_gather, dim3(gather_blocks(N)), dim3(kGatherThreads), 0, st, m, x, coeff, w, dv, N, mode,
                     fs_floor, g_max, out, dpart);
}

}  // namespace collop

namespace collop {

void launch_scatter_b(const MeshView& m, const EnsView& E, const double* v0, const double* vin, const double* w,
                      int fields, int grid_fields, double* vmid, double* dv, long long* part, double* mom,
                      hipStream_t st) {
  const size_t lds = (size_t)m.n_dofs * sizeof(unsigned long long);
  hipLaunchKernelGGL(k_scatter_b, dim3(E.n_sblk, grid_fields), dim3(kScatterThreads), lds, st, m, E, v0, vin, w,
                     fields, vmid, dv, part, mom);
}

void launch_solve_b(const MeshView& m, const EnsView& E, const long long* part, int fields, int in_slot, int out_slot,
                    double* vec, double* fc, hipStream_t st) {
  const size_t lds = solve_lds(m);
  hipLaunchKernelGGL(k_solve_b, dim3(E.R, fields), dim3(kSolveThreads), lds, st, m, E, part, in_slot, out_slot, vec,
                     fc);
}

void launch_quad1_b(const MeshView& m, const EnsView& E, const double* c, int rstride, int fields, int use_logsq,
                    double* iota, double* spart, hipStream_t st) {
  hipLaunchKernelGGL(k_quad1_b, dim3(quad_blocks(m), fields, E.R), dim3(kQuadThreads), 0, st, m, E, c, rstride,
                     use_logsq, iota, spart);
}

void launch_quad2_b(const MeshView& m, const EnsView& E, const double* iota, double* vec, hipStream_t st) {
  hipLaunchKernelGGL(k_quad2_b, dim3((m.n_dofs + 3) / 4, E.R), dim3(256), 0, st, m, E, iota, vec);
}

void launch_sum_rows_b(const EnsView& E, const double* part, const int* row0, const int* cnt, int count, int istride,
                       int jstride, double* out, int ostride, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_rows_b, dim3(E.R, count), dim3(kSumThreads), 0, st, E, part, row0, cnt, istride, jstride,
                     out, ostride);
}

void launch_gather_b(const MeshView& m, const EnsView& E, const double* x, const double* vec, int slot,
                     const double* w, const double* dv, int mode, double* out, double* dpart, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_b, dim3(E.n_gblk), dim3(kGatherThreads), 0, st, m, E, x, vec, slot, w, dv, mode, out,
                     dpart);
}

}  // namespace collop
