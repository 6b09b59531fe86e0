// step.hip -- element-wise and reduction kernels of the device-resident
// implicit step (solver.jl:86-113, 176-227, 300-423; main.jl:239-245).
//
// All reductions are deterministic: per-block partials in a fixed shuffle tree,
// then a fixed-order sum over blocks.
#include "anderson.hpp"
#include "collop_launch.hpp"
#include "lb_dev.hpp"

namespace collop {
namespace {

constexpr int kEwThreads = 256;

// Address of element i of a flattened (2 x n) component pair stored with
// component stride ld (ld = n for a single run).
__device__ __forceinline__ size_t el(int i, int n, int ld) { return i < n ? (size_t)i : (size_t)ld + (i - n); }

// Gonzalez lambda (solver.jl:94-105) from the gather partials, recomputed
// redundantly (same fixed order) by every block, then G_eff = -(dS + lambda dv)/w
// (solver.jl:107-111) and the pair-source records of markers lb*256 + tid.
__device__ __forceinline__ void geff_body(const double* __restrict__ vmid, const double* __restrict__ dS,
                                          const double* __restrict__ dv, const double* __restrict__ w, int N, int ld,
                                          const double* __restrict__ dpart, int nparts, const double* __restrict__ S1,
                                          const double* __restrict__ S0, int use_gonzalez, double lo1, double hi1,
                                          double lo2, double hi2, Src64* __restrict__ s64, Src32* __restrict__ s32,
                                          double* __restrict__ lambda_out, double* lam_sh, double* red, int lb,
                                          double* __restrict__ geff_out = nullptr,
                                          const double* __restrict__ Gk = nullptr,
                                          const double* __restrict__ A = nullptr, double* __restrict__ D = nullptr) {
  // the dot products over the gather partials, in a fixed partition and tree
  double sums[2] = {0.0, 0.0};
  if (use_gonzalez)
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
      sums[0] += dpart[2 * b];
      sums[1] += dpart[2 * b + 1];
    }
  block_sum_n<kEwThreads, 2>(sums, red);
  if (threadIdx.x == 0) {
    const double dot = sums[0], nrm2 = sums[1];
    const double lam = (use_gonzalez && nrm2 > 1e-30) ? (S1[0] - S0[0] - dot) / nrm2 : 0.0;
    *lam_sh = lam;
    if (lb == 0 && lambda_out) lambda_out[0] = lam;
  }
  __syncthreads();
  const double lam = *lam_sh;
  const int i = lb * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double iw = 1.0 / w[i];
  const double g1 = -(dS[i] + lam * dv[i]) * iw;
  const double g2 = -(dS[ld + i] + lam * dv[ld + i]) * iw;
  if (geff_out) {
    geff_out[i] = g1;
    geff_out[ld + i] = g2;
  }
  if (D) {  // frozen-metric product A (G_eff - G_k) of the defect-correction inner map (solver.jl:598-607)
    const double d1 = g1 - Gk[i];
    const double d2 = g2 - Gk[ld + i];
    D[i] = dot2(A[i], d1, A[ld + i], d2);
    D[ld + i] = dot2(A[ld + i], d1, A[2 * (size_t)ld + i], d2);
  }
  const double a = vmid[i], b = vmid[ld + i];
  const bool in = in_box(a, b, lo1, hi1, lo2, hi2);
  const double wm = in ? w[i] : 0.0;
  if (s64) {
    Src64 s;
    s.v1 = a;
    s.v2 = b;
    s.g1 = g1;
    s.g2 = g2;
    s.w = wm;
    s.in_domain = in ? 1.0 : 0.0;
    s.pad0 = s.pad1 = 0.0;
    s64[i] = s;
  }
  if (s32) {
    Src32 s;
    s.v1 = (float)a;
    s.v2 = (float)b;
    s.g1 = (float)g1;
    s.g2 = (float)g2;
    s.w = (float)wm;
    s.in_domain = in ? 1.0f : 0.0f;
    s.pad0 = s.pad1 = 0.0f;
    s32[i] = s;
  }
}

// Map epilogue of block lb over the flattened components i < 2n: slice sum ->
// v_out = v0 + dt dot_v (solver.jl:225) -> r = Gv - v1 (solver.jl:339), the new
// Anderson column (solver.jl:391-394) and the block's partial sums of |r|^2,
// |v1|^2, |Gv - v0|^2, dF_col . dF_j and dF_j . r (solver.jl:398-399).  The
// partials `part` are compact [S][2n]; dF / dG columns are `cs` apart.
__device__ __forceinline__ void epilogue_body(const EpiArgs& A, int N, int ld, size_t cs, double* red, int lb,
                                              const double* s_given = nullptr) {
  const int n2 = 2 * N;
  const int i = lb * blockDim.x + threadIdx.x;
  double st[kStatWidth];
#pragma unroll
  for (int j = 0; j < kStatWidth; ++j) st[j] = 0.0;
  if (i < n2) {
    const size_t e = el(i, N, ld);
    double s = 0.0;
    if (s_given)
      s = *s_given;
    else
      s = slot_sum(A.part, A.S, (size_t)n2, (size_t)i);
    double gv = A.v0[e] + A.dt * s;
    const double x1 = A.v1[e];
    if (A.cl_row) {  // a stiff cluster's row: the cluster solve's value replaces the map's
      const int mk = i < N ? i : i - N;
      const int row = A.cl_row[mk];
      if (row >= 0) {
        const int cp = 2 * row + (i < N ? 0 : 1);
        const double rt = gv - x1;
        gv = A.cl_x[cp];
        const double rh = gv - x1;
        st[kStatTrue] = rt * rt - rh * rh;
      }
    }
    A.Gv[e] = gv;
    const double ri = gv - x1;
    A.r[e] = ri;
    st[0] = ri * ri;
    st[1] = x1 * x1;
    const double inc = gv - A.v0[e];
    st[2] = inc * inc;
    // col >= 0: write the new column and its Gram row; hist > 0: dF_j . r
    // (col < 0 with hist > 0 re-bases the window on a new map at the same iterate)
    double f = 0.0;
    if (A.col >= 0) {
      f = ri - A.rp[e];
      A.dF[A.col * cs + e] = f;
      A.dG[A.col * cs + e] = gv - A.Gp[e];
    }
#pragma unroll
    for (int j = 0; j < kMaxWindow; ++j) {
      if (j < A.hist) {
        const double xj = (j == A.col) ? f : A.dF[j * cs + e];
        if (A.col >= 0) st[3 + j] = xj * f;
        st[3 + kMaxWindow + j] = xj * ri;
      }
    }
  }
  block_sum_n<kEwThreads, kStatWidth>(st, red);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < kStatWidth; ++j) A.stat_part[j] = st[j];
  }
}

// v1 = beta (Gv - dG gamma) + (1 - beta) v1  (solver.jl:383-412)
struct Gamma {
  double g[kMaxWindow];
};

__device__ __forceinline__ double update_value(const double* __restrict__ Gv, const double* __restrict__ dG, size_t e,
                                               size_t cs, const double* g, int hist, double beta, double v1) {
  double x = Gv[e];
  for (int j = 0; j < hist; ++j) x -= dG[j * cs + e] * g[j];
  return dot2(beta, x, 1.0 - beta, v1);
}

// v_out = v + dt * (slice sum) for flattened element i < 2n  (main.jl:245)
__device__ __forceinline__ void euler_body(const double* __restrict__ part, int S, const double* __restrict__ v,
                                           double dt, int N, int ld, double* __restrict__ out,
                                           double* __restrict__ dotv, int i) {
  const int n2 = 2 * N;
  if (i >= n2) return;
  const size_t e = el(i, N, ld);
  const double s = slot_sum(part, S, (size_t)n2, (size_t)i);
  if (dotv) dotv[e] = s;
  out[e] = v[e] + dt * s;
}

// ---- near field of the frozen-metric map ---------------------------------------
// w_a U(d) g, U(d) g = (g - d (d.g) / |d|^2) / |d|  (functions.jl:221-231 as written)
__device__ __forceinline__ void uterm(double wa, double d1, double d2, double g1, double g2, double& o1, double& o2) {
  const double q = dot2(d1, d1, d2, d2);
  const double dg = dot2(d1, g1, d2, g2) / q;
  const double inv = 1.0 / sqrt(q);
  o1 = wa * (g1 - d1 * dg) * inv;
  o2 = wa * (g2 - d2 * dg) * inv;
}

// For near target tg[t] with partners al[off[t] .. off[t+1]): replace each
// pair's frozen-metric model w U(d_k) (G_k,a - G_g(u)) by its exact term
// w U(d(u)) (G_a(u) - G_g(u)) at the inner iterate: D_g -= sum (exact - frozen),
// so Phi = G(v_k) - dt D gains + dt (exact - frozen).  Pairs in fixed order.
__global__ void k_near(int nt, const int* __restrict__ tg, const int* __restrict__ off, const int* __restrict__ al,
                       int N, const double* __restrict__ w, const double* __restrict__ vm,
                       const double* __restrict__ Ge, const double* __restrict__ vmk, const double* __restrict__ Gk,
                       double* __restrict__ D) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int g = tg[t];
  double c1 = 0.0, c2 = 0.0;
  for (int p = off[t]; p < off[t + 1]; ++p) {
    const int a = al[p];
    double e1, e2, f1, f2;
    uterm(w[a], vm[g] - vm[a], vm[N + g] - vm[N + a], Ge[a] - Ge[g], Ge[N + a] - Ge[N + g], e1, e2);
    uterm(w[a], vmk[g] - vmk[a], vmk[N + g] - vmk[N + a], Gk[a] - Ge[g], Gk[N + a] - Ge[N + g], f1, f2);
    c1 += e1 - f1;
    c2 += e2 - f2;
  }
  D[g] -= c1;
  D[N + g] -= c2;
}

// The frozen-metric map's value at flattened element i (solver.jl:598-607):
// G_eff of its marker from the gather's dS and the Gonzalez lambda (geff_body's
// arithmetic), D = A (G_eff - G_k), minus the near-pair corrections of k_near
// (the partners' G_eff formed the same way).
__device__ __forceinline__ double geff_of(const FrozenArgs& F, double lam, int a, int N, double& g2) {
  const double iw = 1.0 / F.w[a];
  g2 = -(F.dS[N + a] + lam * F.dv[N + a]) * iw;
  return -(F.dS[a] + lam * F.dv[a]) * iw;
}

__device__ __forceinline__ double frozen_value(const FrozenArgs& F, double lam, int i, int N) {
  const int a = i < N ? i : i - N;
  double g2;
  const double g1 = geff_of(F, lam, a, N, g2);
  const double d1 = g1 - F.Gk[a];
  const double d2 = g2 - F.Gk[N + a];
  double D = i < N ? dot2(F.A[a], d1, F.A[N + a], d2) : dot2(F.A[N + a], d1, F.A[2 * (size_t)N + a], d2);
  if (F.nidx) {
    const int t = F.nidx[a];
    if (t >= 0) {
      double c1 = 0.0, c2 = 0.0;
      for (int p = F.off[t]; p < F.off[t + 1]; ++p) {
        const int q = F.al[p];
        double h2;
        const double h1 = geff_of(F, lam, q, N, h2);
        double e1, e2, f1, f2;
        uterm(F.w[q], F.vm[a] - F.vm[q], F.vm[N + a] - F.vm[N + q], h1 - g1, h2 - g2, e1, e2);
        uterm(F.w[q], F.vmk[a] - F.vmk[q], F.vmk[N + a] - F.vmk[N + q], F.Gk[q] - g1, F.Gk[N + q] - g2, f1, f2);
        c1 += e1 - f1;
        c2 += e2 - f2;
      }
      D -= i < N ? c1 : c2;
    }
  }
  return D;
}

// ---- single-run kernels ---------------------------------------------------------
__global__ __launch_bounds__(kEwThreads) void k_epilogue_frozen(EpiArgs A, FrozenArgs F, int N) {
  __shared__ double red[kStatWidth * kEwThreads / 64];
  __shared__ double lam_sh;
  // the Gonzalez lambda from the gather partials, in geff_body's fixed order
  double sums[2] = {0.0, 0.0};
  if (F.use_gonzalez)
    for (int b = threadIdx.x; b < F.nparts; b += blockDim.x) {
      sums[0] += F.dpart[2 * b];
      sums[1] += F.dpart[2 * b + 1];
    }
  block_sum_n<kEwThreads, 2>(sums, red);
  if (threadIdx.x == 0) {
    const double dot = sums[0], nrm2 = sums[1];
    lam_sh = (F.use_gonzalez && nrm2 > 1e-30) ? (F.S1[0] - F.S0[0] - dot) / nrm2 : 0.0;
  }
  __syncthreads();
  const double lam = lam_sh;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double s = i < 2 * N ? frozen_value(F, lam, i, N) : 0.0;
  EpiArgs a = A;
  a.stat_part = A.stat_part + (size_t)blockIdx.x * kStatWidth;
  epilogue_body(a, N, N, (size_t)2 * N, red, blockIdx.x, &s);
}

// The Lenard-Bernstein map's epilogue: the reduction, lambda and the drift
// multipliers, recomputed by every block in launch_lb_rhs's fixed order, then
// v_dot of element i and the Anderson epilogue of v0 + dt v_dot.
__global__ __launch_bounds__(kEwThreads) void k_epilogue_lb(EpiArgs A, LbEpiArgs L, int N) {
  static_assert(kEwThreads == kLbThreads, "lb_coef_body reduces over kLbThreads threads");
  __shared__ double red[kStatWidth * kEwThreads / 64];
  static_assert(kStatWidth >= kLbSums + 1, "red also holds the LB reduction");
  __shared__ double coef[5];
  lb_coef_body(L.part, L.nparts, L.qpart, L.nq, L.S0, red, coef);
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x < 5) L.coef[threadIdx.x] = coef[threadIdx.x];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double s = i < 2 * N ? lb_dot_value(L.x, L.g, L.dv, L.w, coef, L.nu, N, i) : 0.0;
  EpiArgs a = A;
  a.stat_part = A.stat_part + (size_t)blockIdx.x * kStatWidth;
  epilogue_body(a, N, N, (size_t)2 * N, red, blockIdx.x, &s);
}

// The Anderson statistics of the epilogue partials part[n][kStatWidth]: the sums
// of launch_sum_rows, written by thread 0 to the pinned host buffer out, then
// seq to *seq_out with system-scope release, so a host that reads seq also
// reads the statistics.  Four groups of 256 threads share the columns (group
// g takes columns g, g + 4, ...); within a group, thread t sums rows t, t + 256,
// ... and the four waves reduce by the shuffle tree of block_sum, then in wave
// order -- the partition and tree of launch_sum_rows' 256-thread blocks, so the
// sums are bitwise the same.
constexpr int kPubGroups = 4;
// Thread 0 of a G * 256 workgroup ends with out[c] = the sum over the n rows of
// part[n][kStatWidth]; group g sums the columns c = g mod G, each with the
// same partition and tree, so the sums do not depend on G.
template <int G>
__device__ __forceinline__ void stat_rows_sum(const double* __restrict__ part, int n, double* out) {
  constexpr int kPubGroups = G;
  constexpr int kPer = (kStatWidth + kPubGroups - 1) / kPubGroups;
  __shared__ double wsum[kStatWidth][4];
  const int g = threadIdx.x >> 8, t = threadIdx.x & 255;
  const int wave = t >> 6, lane = t & 63;
  double s[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) s[q] = 0.0;
  for (int j = t; j < n; j += 256)
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int c = g + q * kPubGroups;
      if (c < kStatWidth) s[q] += part[(size_t)j * kStatWidth + c];
    }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s[q] += __shfl_xor(s[q], off, 64);
    const int c = g + q * kPubGroups;
    if (lane == 0 && c < kStatWidth) wsum[c][wave] = s[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < kStatWidth; ++c) {
      double v = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) v += wsum[c][i];
      out[c] = v;
    }
  }
}

__global__ __launch_bounds__(kPubGroups * 256) void k_stat_publish(const double* __restrict__ part, int n,
                                                                   double* out, unsigned* seq_out, unsigned seq) {
  stat_rows_sum<kPubGroups>(part, n, out);
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(seq_out, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---- device-resident Anderson solve -------------------------------------------
// Control of map k: the statistics of its epilogue, the state machine's
// decision (the least squares for gamma included) and the plan of map k + 1,
// then pub = {live, kind, err} and seq = seq0 + k published to the host.  A
// launch whose map did not run (the solve ended before it) changes nothing.
// One group of 256 threads: the state machine of thread 0 then has the
// register file of a whole SIMD lane to itself.
__global__ __launch_bounds__(256) void k_anderson_ctl(const double* __restrict__ part, int n, AndCtl* __restrict__ c,
                                                      int k, unsigned* pub, unsigned seq0) {
  // the state machine runs on a copy in LDS (thread 0), written back after
  constexpr int kWords = (int)(sizeof(AndersonRun) / sizeof(int));
  static_assert(sizeof(AndersonRun) % sizeof(int) == 0, "AndersonRun is copied as ints");
  __shared__ double st[kStatWidth];
  __shared__ int run_map;
  __shared__ __align__(16) int run_lds[kWords];
  if (threadIdx.x == 0) run_map = c->gate[0] != 0 && c->maps == k - 1;
  __syncthreads();
  if (!run_map) return;
  const int* src = reinterpret_cast<const int*>(&c->run);
  for (int i = threadIdx.x; i < kWords; i += blockDim.x) run_lds[i] = src[i];
  stat_rows_sum<1>(part, n, st);
  __syncthreads();  // st (written by thread 0) and run_lds complete
  if (threadIdx.x == 0) {
    AndersonRun& run = *reinterpret_cast<AndersonRun*>(run_lds);
    AndersonAct act;
    const int e = run.after_map(st, act);
    c->maps = k;
    int live = 1, kind = 0;
    if (e != COLLOP_OK) {
      c->err = 1;
      live = 0;
    } else {
      c->act = act;
      if (act.exit != kContinue) {
        live = 0;
      } else {
        const AndersonPlan pl = run.plan();
        kind = pl.kind;
        c->gate[1] = pl.kind;
        c->gate[2] = pl.col;
        c->gate[3] = pl.hist;
      }
    }
    c->gate[0] = live;
    for (int j = 0; j < kStatWidth; ++j) c->st[j] = st[j];
    pub[1] = (unsigned)live;
    pub[2] = (unsigned)kind;
    pub[3] = (unsigned)c->err;
  }
  __syncthreads();
  int* dst = reinterpret_cast<int*>(&c->run);
  for (int i = threadIdx.x; i < kWords; i += blockDim.x) dst[i] = run_lds[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(pub, seq0 + (unsigned)k, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The action of map k (if its control ran): best <- (exit_picard ? Gv : v1) when
// the map improved the best residual; then the exit (v1 <- Gv or v1 at
// convergence, v1 <- best otherwise) or the update v1 <- beta (Gv - dG gamma) +
// (1 - beta) v1 followed by (rp, Gp) <- (r, Gv).
__global__ void k_anderson_apply(const AndCtl* __restrict__ c, int k, double* v1, const double* __restrict__ Gv,
                                 const double* __restrict__ dG, const double* __restrict__ r, double* __restrict__ rp,
                                 double* __restrict__ Gp, double* __restrict__ best, int n2, int exit_picard) {
  if (c->maps != k || c->err) return;
  const AndersonAct& a = c->act;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  const double x = v1[i], g = Gv[i];
  double b = 0.0;
  if (a.copy_best) {
    b = exit_picard ? g : x;
    best[i] = b;
  }
  if (a.exit == kConverged) {
    v1[i] = exit_picard ? g : x;
    return;
  }
  if (a.exit != kContinue) {
    v1[i] = a.copy_best ? b : best[i];
    return;
  }
  if (a.update) v1[i] = update_value(Gv, dG, i, n2, a.gamma, a.h, a.beta, x);
  if (a.rotate) {
    rp[i] = r[i];
    Gp[i] = g;
  }
}

__global__ void k_near_index(int nt, const int* __restrict__ tg, int* __restrict__ nidx) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nt) nidx[tg[t]] = t;
}

__global__ __launch_bounds__(kEwThreads) void k_geff_pack(const double* __restrict__ vmid,
                                                          const double* __restrict__ dS,
                                                          const double* __restrict__ dv,
                                                          const double* __restrict__ w, int N,
                                                          const double* __restrict__ dpart, int nparts,
                                                          const double* __restrict__ S1,
                                                          const double* __restrict__ S0, int use_gonzalez,
                                                          double lo1, double hi1, double lo2, double hi2,
                                                          Src64* __restrict__ s64, Src32* __restrict__ s32,
                                                          double* __restrict__ lambda_out,
                                                          double* __restrict__ geff_out,
                                                          const double* __restrict__ Gk,
                                                          const double* __restrict__ A, double* __restrict__ D) {
  __shared__ double lam_sh;
  __shared__ double red[2 * kEwThreads / 64];
  geff_body(vmid, dS, dv, w, N, N, dpart, nparts, S1, S0, use_gonzalez, lo1, hi1, lo2, hi2, s64, s32, lambda_out,
            &lam_sh, red, blockIdx.x, geff_out, Gk, A, D);
}

__global__ __launch_bounds__(kEwThreads) void k_epilogue(EpiArgs A, int N) {
  __shared__ double red[kStatWidth * kEwThreads / 64];
  EpiArgs a = A;
  if (A.gate) {  // device-resident solve: the map's plan
    if (A.gate[0] == 0) return;
    a.col = A.gate[2];
    a.hist = A.gate[3];
    if (A.gate[1]) a.S = A.S32;
  }
  a.stat_part = A.stat_part + (size_t)blockIdx.x * kStatWidth;
  epilogue_body(a, N, N, (size_t)2 * N, red, blockIdx.x);
}

// bdst given: bdst <- bsrc first (the best iterate, saved in the same launch;
// bsrc may be v1, read before it is updated).
__global__ void k_update(double* v1, const double* __restrict__ Gv, const double* __restrict__ dG, Gamma gam,
                         int hist, double beta, int n2, const double* bsrc, double* __restrict__ bdst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  const double x = v1[i];
  if (bdst) bdst[i] = bsrc == v1 ? x : bsrc[i];
  v1[i] = update_value(Gv, dG, i, n2, gam.g, hist, beta, x);
}

__global__ void k_euler(const double* __restrict__ part, int S, const double* __restrict__ v, double dt, int N,
                        double* __restrict__ out, double* __restrict__ dotv) {
  euler_body(part, S, v, dt, N, N, out, dotv, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- batched ensemble kernels ---------------------------------------------------
// scal is [R][8]: S0, S1, lambda, P1, P2, |v|^2 w, -, -
__global__ __launch_bounds__(kEwThreads) void k_geff_pack_b(EnsView E, const double* __restrict__ vmid,
                                                            const double* __restrict__ dS,
                                                            const double* __restrict__ dv,
                                                            const double* __restrict__ w,
                                                            const double* __restrict__ dpart, double* __restrict__ scal,
                                                            int use_gonzalez, double lo1, double hi1, double lo2,
                                                            double hi2, Src64* __restrict__ s64,
                                                            Src32* __restrict__ s32) {
  const int2 t = E.pblk[blockIdx.x];
  const RunCtl& c = E.ctl[t.x];
  if (!c.active) return;
  __shared__ double lam_sh;
  __shared__ double red[2 * kEwThreads / 64];
  const EnsRun& R = E.runs[t.x];
  double* sc = scal + 8 * t.x;
  geff_body(vmid + R.off, dS + R.off, dv + R.off, w + R.off, R.n, E.NT, dpart + 2 * R.gblk0, R.ngb, sc + 1, sc + 0,
            use_gonzalez, lo1, hi1, lo2, hi2, c.kind == 0 ? s64 + R.off : nullptr, c.kind == 1 ? s32 + R.off : nullptr,
            sc + 2, &lam_sh, red, t.y);
}


__global__ __launch_bounds__(kEwThreads) void k_epilogue_b(EnsView E, EnsBufs B, double dt) {
  const int2 t = E.eblk[blockIdx.x];
  const RunCtl& c = E.ctl[t.x];
  if (!c.active) return;
  __shared__ double red[kStatWidth * kEwThreads / 64];
  const EnsRun& R = E.runs[t.x];
  EpiArgs a;
  a.part = B.part + R.poff;
  a.S = R.S;
  a.v0 = B.v0 + R.off;
  a.v1 = B.v1 + R.off;
  a.dt = dt;
  a.Gv = B.GV[c.cur] + R.off;
  a.r = B.R[c.cur] + R.off;
  a.rp = B.R[1 - c.cur] + R.off;
  a.Gp = B.GV[1 - c.cur] + R.off;
  a.dF = B.dF + R.off;
  a.dG = B.dG + R.off;
  a.col = c.col;
  a.hist = c.hist;
  a.stat_part = B.statp + (size_t)(R.eblk0 + t.y) * kStatWidth;
  a.cl_row = nullptr;
  a.cl_x = nullptr;
  a.gate = nullptr;
  a.S32 = 0;
  epilogue_body(a, R.n, E.NT, (size_t)2 * E.NT, red, t.y);
}

// Per-run update of the iterate before the map (RunCtl::upd, copy_best).
__global__ __launch_bounds__(kEwThreads) void k_update_b(EnsView E, EnsBufs B) {
  const int2 t = E.eblk[blockIdx.x];
  const RunCtl& c = E.ctl[t.x];
  if (c.upd == 0 && !c.copy_best) return;
  const EnsRun& R = E.runs[t.x];
  const int i = t.y * kEwThreads + threadIdx.x;
  if (i >= 2 * R.n) return;
  const size_t e = R.off + el(i, R.n, E.NT);
  const double* GV = B.GV[1 - c.cur];
  const double x = B.v1[e];
  if (c.copy_best) B.best[e] = c.exit_picard ? GV[e] : x;
  double nx = x;
  if (c.upd == 1) {
    nx = update_value(GV, B.dG, e, (size_t)2 * E.NT, c.gamma, c.upd_hist, c.beta, x);
  } else if (c.upd == 2) {
    nx = GV[e];
  } else if (c.upd == 3) {
    nx = c.copy_best ? (c.exit_picard ? GV[e] : x) : B.best[e];
  }
  B.v1[e] = nx;
}

__global__ void k_euler_b(EnsView E, const double* __restrict__ part, const double* __restrict__ v, double dt,
                          double* __restrict__ out, double* __restrict__ dotv) {
  const int2 t = E.eblk[blockIdx.x];
  const EnsRun& R = E.runs[t.x];
  euler_body(part + R.poff, R.S, v + R.off, dt, R.n, E.NT, out + R.off, dotv ? dotv + R.off : nullptr,
             t.y * kEwThreads + threadIdx.x);
}

// dst <- src over every run's markers (both components)
__global__ void k_copy_b(EnsView E, const double* __restrict__ src, double* __restrict__ dst) {
  const int2 t = E.eblk[blockIdx.x];
  const EnsRun& R = E.runs[t.x];
  const int i = t.y * kEwThreads + threadIdx.x;
  if (i >= 2 * R.n) return;
  const size_t e = R.off + el(i, R.n, E.NT);
  dst[e] = src[e];
}

}  // namespace

void launch_geff_pack(const double* vmid, const double* dS, const double* dv, const double* w, int N,
                      const double* dpart, int nparts, const double* S1, const double* S0, int use_gonzalez,
                      double lo1, double hi1, double lo2, double hi2, Src64* s64, Src32* s32, double* lambda_out,
                      hipStream_t st, double* geff_out, const double* Gk, const double* A, double* D) {
  hipLaunchKernelGGL(k_geff_pack, dim3((N + kEwThreads - 1) / kEwThreads), dim3(kEwThreads), 0, st, vmid, dS, dv, w,
                     N, dpart, nparts, S1, S0, use_gonzalez, lo1, hi1, lo2, hi2, s64, s32, lambda_out, geff_out, Gk,
                     A, D);
}

int epi_blocks(int N) { return (2 * N + kEwThreads - 1) / kEwThreads; }

void launch_epilogue(const EpiArgs& a, int N, hipStream_t st) {
  hipLaunchKernelGGL(k_epilogue, dim3(epi_blocks(N)), dim3(kEwThreads), 0, st, a, N);
}

void launch_epilogue_frozen(const EpiArgs& a, const FrozenArgs& f, int N, hipStream_t st) {
  hipLaunchKernelGGL(k_epilogue_frozen, dim3(epi_blocks(N)), dim3(kEwThreads), 0, st, a, f, N);
}

void launch_epilogue_lb(const EpiArgs& a, const LbEpiArgs& l, int N, hipStream_t st) {
  hipLaunchKernelGGL(k_epilogue_lb, dim3(epi_blocks(N)), dim3(kEwThreads), 0, st, a, l, N);
}

void launch_anderson_ctl(const double* part, int n, AndCtl* c, int k, unsigned* pub, unsigned seq0, hipStream_t st) {
  hipLaunchKernelGGL(k_anderson_ctl, dim3(1), dim3(256), 0, st, part, n, c, k, pub, seq0);
}

void launch_anderson_apply(const AndCtl* c, int k, double* v1, const double* Gv, const double* dG, const double* r,
                           double* rp, double* Gp, double* best, int N, int exit_picard, hipStream_t st) {
  const int n2 = 2 * N;
  hipLaunchKernelGGL(k_anderson_apply, dim3((n2 + 255) / 256), dim3(256), 0, st, c, k, v1, Gv, dG, r, rp, Gp, best,
                     n2, exit_picard);
}

void launch_stat_publish(const double* part, int n, double* out, unsigned* seq_out, unsigned seq, hipStream_t st) {
  hipLaunchKernelGGL(k_stat_publish, dim3(1), dim3(kPubGroups * 256), 0, st, part, n, out, seq_out, seq);
}

void launch_near_index(int nt, const int* tg, int N, int* nidx, hipStream_t st) {
  (void)hipMemsetAsync(nidx, 0xff, (size_t)N * sizeof(int), st);
  if (nt > 0) hipLaunchKernelGGL(k_near_index, dim3((nt + 255) / 256), dim3(256), 0, st, nt, tg, nidx);
}

void launch_update(double* v1, const double* Gv, const double* dG, const double* gam, int hist, double beta, int N,
                   hipStream_t st, const double* best_src, double* best_dst) {
  Gamma g{};
  for (int j = 0; j < hist && j < kMaxWindow; ++j) g.g[j] = gam[j];
  const int n2 = 2 * N;
  hipLaunchKernelGGL(k_update, dim3((n2 + 255) / 256), dim3(256), 0, st, v1, Gv, dG, g, hist, beta, n2, best_src,
                     best_dst);
}

void launch_near(int nt, const int* tg, const int* off, const int* al, int N, const double* w, const double* vm,
                 const double* Ge, const double* vmk, const double* Gk, double* D, hipStream_t st) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_near, dim3((nt + 127) / 128), dim3(128), 0, st, nt, tg, off, al, N, w, vm, Ge, vmk, Gk, D);
}

void launch_euler(const double* part, int S, const double* v, double dt, int N, double* out, double* dotv,
                  hipStream_t st) {
  const int n2 = 2 * N;
  hipLaunchKernelGGL(k_euler, dim3((n2 + 255) / 256), dim3(256), 0, st, part, S, v, dt, N, out, dotv);
}

}  // namespace collop

namespace collop {

void launch_geff_pack_b(const EnsView& E, const double* vmid, const double* dS, const double* dv, const double* w,
                        const double* dpart, double* scal, int use_gonzalez, double lo1, double hi1, double lo2,
                        double hi2, Src64* s64, Src32* s32, hipStream_t st) {
  hipLaunchKernelGGL(k_geff_pack_b, dim3(E.n_pblk), dim3(kEwThreads), 0, st, E, vmid, dS, dv, w, dpart, scal,
                     use_gonzalez, lo1, hi1, lo2, hi2, s64, s32);
}

void launch_epilogue_b(const EnsView& E, const EnsBufs& B, double dt, hipStream_t st) {
  hipLaunchKernelGGL(k_epilogue_b, dim3(E.n_eblk), dim3(kEwThreads), 0, st, E, B, dt);
}

void launch_update_b(const EnsView& E, const EnsBufs& B, hipStream_t st) {
  hipLaunchKernelGGL(k_update_b, dim3(E.n_eblk), dim3(kEwThreads), 0, st, E, B);
}

void launch_euler_b(const EnsView& E, const double* part, const double* v, double dt, double* out, double* dotv,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_euler_b, dim3(E.n_eblk), dim3(kEwThreads), 0, st, E, part, v, dt, out, dotv);
}

void launch_copy_b(const EnsView& E, const double* src, double* dst, hipStream_t st) {
  hipLaunchKernelGGL(k_copy_b, dim3(E.n_eblk), dim3(kEwThreads), 0, st, E, src, dst);
}

}  // namespace collop
