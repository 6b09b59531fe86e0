// probe.hip -- issue cost of single VALU instruction kinds on this GPU, the
// denominator of the pair kernels' issue-bound estimates.  Each thread runs 16
// independent chains of one instruction kind; 8 workgroups of 4 waves per CU
// (8 waves per SIMD) hide the latency, so the time per instruction is the
// issue cost.
#include <hip/hip_runtime.h>

#include "collop_launch.hpp"

namespace collop {
namespace {

constexpr int kChains = 16;

template <int KIND>
__global__ __launch_bounds__(256) void k_issue_f32(float* __restrict__ out, float a, float b, int iters) {
  float x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = threadIdx.x * 1e-3f + c + 1.0f;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      if (KIND == 0) x[c] = __builtin_fmaf(x[c], a, b);
      if (KIND == 1) x[c] = x[c] - b;
      if (KIND == 2) x[c] = __builtin_amdgcn_rsqf(x[c]);
      if (KIND == 3)
        x[c] = __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x[c]), 0x134, 0xf, 0xf,
                                                                  false));
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <int KIND>
__global__ __launch_bounds__(256) void k_issue_f64(double* __restrict__ out, double a, double b, int iters) {
  double x[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) x[c] = threadIdx.x * 1e-3 + c + 1.0;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      if (KIND == 0) x[c] = __builtin_fma(x[c], a, b);
      if (KIND == 1) x[c] = __builtin_amdgcn_rsq(x[c]);
    }
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kChains; ++c) s += x[c];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

int issue_probe(int kind, int iters, int blocks, void* out, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float* ms) {
  if (hipEventRecord(e0, st) != hipSuccess) return -1;
  float* o32 = static_cast<float*>(out);
  double* o64 = static_cast<double*>(out);
  switch (kind) {
    case 0: hipLaunchKernelGGL(k_issue_f32<0>, dim3(blocks), dim3(256), 0, st, o32, 0.999f, 1e-3f, iters); break;
    case 1: hipLaunchKernelGGL(k_issue_f32<1>, dim3(blocks), dim3(256), 0, st, o32, 0.999f, 1e-3f, iters); break;
    case 2: hipLaunchKernelGGL(k_issue_f32<2>, dim3(blocks), dim3(256), 0, st, o32, 0.999f, 1e-3f, iters); break;
    case 3: hipLaunchKernelGGL(k_issue_f32<3>, dim3(blocks), dim3(256), 0, st, o32, 0.999f, 1e-3f, iters); break;
    case 4: hipLaunchKernelGGL(k_issue_f64<0>, dim3(blocks), dim3(256), 0, st, o64, 0.999, 1e-3, iters); break;
    case 5: hipLaunchKernelGGL(k_issue_f64<1>, dim3(blocks), dim3(256), 0, st, o64, 0.999, 1e-3, iters); break;
    default: return -2;
  }
  if (hipGetLastError() != hipSuccess) return -1;
  if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1;
  return hipEventElapsedTime(ms, e0, e1) == hipSuccess ? 0 : -1;
}

int issue_probe_chains() { return kChains; }

}  // namespace collop
