// step.hip -- element-wise and reduction kernels of the device-resident
// implicit step (solver.jl:86-113, 176-227, 300-423; main.jl:239-245).
//
// All reductions are deterministic: per-block partials in a fixed shuffle tree,
// then a fixed-order sum over blocks.
#include "collop_launch.hpp"

namespace collop {
namespace {

constexpr int kEwThreads = 256;

// Gonzalez lambda (solver.jl:94-105) from the gather partials, recomputed
// redundantly (same fixed order) by every block, then G_eff = -(dS + lambda dv)/w
// (solver.jl:107-111) and the pair-source records.
__global__ __launch_bounds__(kEwThreads) void k_geff_pack(const double* __restrict__ vmid,
                                                          const double* __restrict__ dS,
                                                          const double* __restrict__ dv,
                                                          const double* __restrict__ w, int N,
                                                          const double* __restrict__ dpart, int nparts,
                                                          const double* __restrict__ S1,
                                                          const double* __restrict__ S0, int use_gonzalez,
                                                          double lo1, double hi1, double lo2, double hi2,
                                                          Src64* __restrict__ s64, Src32* __restrict__ s32,
                                                          double* __restrict__ lambda_out) {
  __shared__ double lam_sh;
  if (threadIdx.x == 0) {
    double lam = 0.0;
    if (use_gonzalez) {
      double dot = 0.0, nrm2 = 0.0;
      for (int b = 0; b < nparts; ++b) {
        dot += dpart[2 * b];
        nrm2 += dpart[2 * b + 1];
      }
      lam = nrm2 > 1e-30 ? (S1[0] - S0[0] - dot) / nrm2 : 0.0;
    }
    lam_sh = lam;
    if (blockIdx.x == 0 && lambda_out) lambda_out[0] = lam;
  }
  __syncthreads();
  const double lam = lam_sh;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double iw = 1.0 / w[i];
  const double g1 = -(dS[i] + lam * dv[i]) * iw;
  const double g2 = -(dS[N + i] + lam * dv[N + i]) * iw;
  const double a = vmid[i], b = vmid[N + i];
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

// Map epilogue: slice sum -> v_out = v0 + dt dot_v (solver.jl:225) -> r = Gv - v1
// (solver.jl:339), the new Anderson column (solver.jl:391-394) and the partial
// sums of |r|^2, |v1|^2, |Gv - v0|^2, dF_col . dF_j and dF_j . r
// (solver.jl:398-399).
__global__ __launch_bounds__(kEwThreads) void k_epilogue(EpiArgs A, int N) {
  __shared__ double red[kEwThreads / 64];
  const int n2 = 2 * N;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double st[kStatWidth];
#pragma unroll
  for (int j = 0; j < kStatWidth; ++j) st[j] = 0.0;
  if (i < n2) {
    double s = 0.0;
    for (int k = 0; k < A.S; ++k) s += A.part[(size_t)k * n2 + i];
    const double gv = A.v0[i] + A.dt * s;
    A.Gv[i] = gv;
    const double x1 = A.v1[i];
    const double ri = gv - x1;
    A.r[i] = ri;
    st[0] = ri * ri;
    st[1] = x1 * x1;
    const double inc = gv - A.v0[i];
    st[2] = inc * inc;
    // col >= 0: write the new column and its Gram row; hist > 0: dF_j . r
    // (col < 0 with hist > 0 re-bases the window on a new map at the same iterate)
    double f = 0.0;
    if (A.col >= 0) {
      f = ri - A.rp[i];
      A.dF[(size_t)A.col * n2 + i] = f;
      A.dG[(size_t)A.col * n2 + i] = gv - A.Gp[i];
    }
#pragma unroll
    for (int j = 0; j < kMaxWindow; ++j) {
      if (j < A.hist) {
        const double xj = (j == A.col) ? f : A.dF[(size_t)j * n2 + i];
        if (A.col >= 0) st[3 + j] = xj * f;
        st[3 + kMaxWindow + j] = xj * ri;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kStatWidth; ++j) {
    const bool used = j < 3 || (A.col >= 0 && j >= 3 && j < 3 + A.hist) ||
                      (j >= 3 + kMaxWindow && j < 3 + kMaxWindow + A.hist);
    if (!used) continue;  // uniform across the block
    const double t = block_sum<kEwThreads>(st[j], red);
    if (threadIdx.x == 0) A.stat_part[(size_t)blockIdx.x * kStatWidth + j] = t;
  }
}

// v1 = beta (Gv - dG gamma) + (1 - beta) v1  (solver.jl:383-412)
struct Gamma {
  double g[kMaxWindow];
};

__global__ void k_update(double* __restrict__ v1, const double* __restrict__ Gv, const double* __restrict__ dG,
                         Gamma gam, int hist, double beta, int n2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  double x = Gv[i];
  for (int j = 0; j < hist; ++j) x -= dG[(size_t)j * n2 + i] * gam.g[j];
  v1[i] = beta * x + (1.0 - beta) * v1[i];
}

// v1 = v + dt * dot_v  (main.jl:245); also keeps dot_v.
__global__ void k_euler(const double* __restrict__ part, int S, const double* __restrict__ v, double dt, int n2,
                        double* __restrict__ out, double* __restrict__ dotv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  double s = 0.0;
  for (int k = 0; k < S; ++k) s += part[(size_t)k * n2 + i];
  if (dotv) dotv[i] = s;
  out[i] = v[i] + dt * s;
}

}  // namespace

void launch_geff_pack(const double* vmid, const double* dS, const double* dv, const double* w, int N,
                      const double* dpart, int nparts, const double* S1, const double* S0, int use_gonzalez,
                      double lo1, double hi1, double lo2, double hi2, Src64* s64, Src32* s32, double* lambda_out,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_geff_pack, dim3((N + kEwThreads - 1) / kEwThreads), dim3(kEwThreads), 0, st, vmid, dS, dv, w,
                     N, dpart, nparts, S1, S0, use_gonzalez, lo1, hi1, lo2, hi2, s64, s32, lambda_out);
}

int epi_blocks(int N) { return (2 * N + kEwThreads - 1) / kEwThreads; }

void launch_epilogue(const EpiArgs& a, int N, hipStream_t st) {
  hipLaunchKernelGGL(k_epilogue, dim3(epi_blocks(N)), dim3(kEwThreads), 0, st, a, N);
}

void launch_update(double* v1, const double* Gv, const double* dG, const double* gam, int hist, double beta, int N,
                   hipStream_t st) {
  Gamma g{};
  for (int j = 0; j < hist && j < kMaxWindow; ++j) g.g[j] = gam[j];
  const int n2 = 2 * N;
  hipLaunchKernelGGL(k_update, dim3((n2 + 255) / 256), dim3(256), 0, st, v1, Gv, dG, g, hist, beta, n2);
}

void launch_euler(const double* part, int S, const double* v, double dt, int N, double* out, double* dotv,
                  hipStream_t st) {
  const int n2 = 2 * N;
  hipLaunchKernelGGL(k_euler, dim3((n2 + 255) / 256), dim3(256), 0, st, part, S, v, dt, n2, out, dotv);
}

}  // namespace collop
