// mlp.hip -- the per-marker warm-start MLP (6 -> 64 -> 64 -> 2, GELU) on gfx950
// matrix cores: forward, backward, Adam.
//
// The reference has only the hook site of its neural warm start (main.jl:164-166,
// 252-254; warmstart_nn.jl is absent), so this network is the library's own:
//   z1 = W1 x + b1, h1 = gelu(z1); z2 = W2 h1 + b2, h2 = gelu(z2); y = W3 h2 + b3,
// gelu(z) = z/2 (1 + tanh(sqrt(2/pi)(z + 0.044715 z^3))).
// Every product is a v_mfma_f32_16x16x16_bf16 with bfloat16 operands (weights,
// activations, back-propagated signals, round-to-nearest-even) and FP32
// accumulation.  The activations are kept transposed -- hidden units along M,
// the 16 markers of a tile along N -- so the D tile of one layer is already the
// B operand of the next (lane l holds column l % 16, rows 4 (l / 16) + j).
// Weight gradients contract over the markers of a tile (K = 16) with operands
// re-read from LDS in the transposed order.  One wave per workgroup; all tiles
// of a workgroup accumulate gradients in registers, then one partial row per
// workgroup is summed in a fixed order (deterministic training).
//
// Parameter vector (float32, 4738): W1[64][6] b1[64] W2[64][64] b2[64] W3[2][64] b3[2].
// Sample records (float32, 8 per marker): x0..x5 (raw features), t0, t1 (raw target).
#include "collop_launch.hpp"

namespace collop {
namespace {

constexpr int kHid = 64, kIn = 6, kOut = 2;
constexpr int oW1 = 0, oB1 = 384, oW2 = 448, oB2 = 4544, oW3 = 4608, oB3 = 4736, kNP = kMlpParams;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) {
  const float c = 0.7978845608028654f;
  const float t = tanhf(c * (z + 0.044715f * z * z * z));
  return 0.5f * z * (1.0f + t);
}

__device__ __forceinline__ float gelu_grad(float z) {
  const float c = 0.7978845608028654f;
  const float t = tanhf(c * (z + 0.044715f * z * z * z));
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

// Operand loads (lane l: r = l % 16, q = l / 16).
//   A[i][k], i = i0 + r, k = k0 + 4q + j, from row-major M (M[i * ld + k]) or
//   transposed (M[k * ld + i]);  B[k][n], k = k0 + 4q + j, n = n0 + r, likewise.
__device__ __forceinline__ bf16x4 ld_rows(const float* M, int ld, int i, int k) {
  bf16x4 v;
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = to_bf16(M[i * ld + k + j]);
  return v;
}
__device__ __forceinline__ bf16x4 ld_cols(const float* M, int ld, int i, int k) {
  bf16x4 v;
#pragma unroll
  for (int j = 0; j < 4; ++j) v[j] = to_bf16(M[(k + j) * ld + i]);
  return v;
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct Lds {
  float W1[kHid * 16];    // [64][16], inputs padded with zeros
  float W2[kHid * kHid];  // [64][64]
  float W3[16 * kHid];    // [16][64], rows >= 2 zero
  float b1[kHid], b2[kHid], b3[16];
  float X[16 * 16];       // [k][marker], k >= 6 zero
  float Z1[kHid * 16], H1[kHid * 16], Z2[kHid * 16], H2[kHid * 16];
  float DY[16 * 16], DZ2[kHid * 16], DZ1[kHid * 16];
};

}  // namespace

namespace {

__global__ __launch_bounds__(64) void k_mlp(MlpArgs a) {
  __shared__ Lds s;
  const int l = threadIdx.x, r = l & 15, q = l >> 4;
  // weights to LDS (zero padding)
  for (int i = l; i < kHid * 16; i += 64) {
    const int o = i / 16, k = i % 16;
    s.W1[i] = k < kIn ? a.params[oW1 + o * kIn + k] : 0.0f;
  }
  for (int i = l; i < kHid * kHid; i += 64) s.W2[i] = a.params[oW2 + i];
  for (int i = l; i < 16 * kHid; i += 64) s.W3[i] = i < kOut * kHid ? a.params[oW3 + i] : 0.0f;
  s.b1[l] = a.params[oB1 + l];
  s.b2[l] = a.params[oB2 + l];
  if (l < 16) s.b3[l] = l < kOut ? a.params[oB3 + l] : 0.0f;
  __syncthreads();

  f32x4 gW1[4], gW2[16], gW3[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) gW1[i] = gW3[i] = f32x4{0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 16; ++i) gW2[i] = f32x4{0, 0, 0, 0};
  float gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f;
  double loss = 0.0;

  for (int t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const long long tile = a.tiles ? a.tiles[t] : t;
    const long long s0 = tile * 16;
    // ---- inputs X[k][m] (normalised), targets
    for (int i = l; i < 256; i += 64) {
      const int k = i / 16, m = i % 16;
      float v = 0.0f;
      if (k < kIn && s0 + m < a.nsamples) v = (a.rec[(s0 + m) * 8 + k] - a.mu[k]) * a.isd[k];
      s.X[i] = v;
    }
    __syncthreads();
    // ---- layer 1: Z1 = W1 X
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
      f32x4 d = mfma(ld_rows(s.W1, 16, hb * 16 + r, 4 * q), ld_cols(s.X, 16, r, 4 * q), f32x4{0, 0, 0, 0});
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = hb * 16 + 4 * q + j;
        const float z = d[j] + s.b1[row];
        s.Z1[row * 16 + r] = z;
        s.H1[row * 16 + r] = gelu(z);
      }
    }
    __syncthreads();
    // ---- layer 2: Z2 = W2 H1
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
      f32x4 d = {0, 0, 0, 0};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        d = mfma(ld_rows(s.W2, kHid, hb * 16 + r, kb * 16 + 4 * q), ld_cols(s.H1, 16, r, kb * 16 + 4 * q), d);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = hb * 16 + 4 * q + j;
        const float z = d[j] + s.b2[row];
        s.Z2[row * 16 + r] = z;
        s.H2[row * 16 + r] = gelu(z);
      }
    }
    __syncthreads();
    // ---- layer 3: y = W3 H2 (rows 0, 1 of the D tile, held by lanes q == 0)
    f32x4 d3 = {0, 0, 0, 0};
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
      d3 = mfma(ld_rows(s.W3, kHid, r, kb * 16 + 4 * q), ld_cols(s.H2, 16, r, kb * 16 + 4 * q), d3);
    const bool valid = s0 + r < a.nsamples;
    float y0 = d3[0] + s.b3[0], y1 = d3[1] + s.b3[1];
    if (a.mode == 0) {
      if (q == 0 && valid) {
        a.y[(s0 + r) * 2 + 0] = y0;
        a.y[(s0 + r) * 2 + 1] = y1;
      }
      __syncthreads();
      continue;
    }
    // ---- loss and dY (rows 0, 1), mean over the batch
    for (int i = l; i < 256; i += 64) s.DY[i] = 0.0f;
    __syncthreads();
    if (q == 0 && valid) {
      const float e0 = y0 - a.rec[(s0 + r) * 8 + 6] / a.oscale[0];
      const float e1 = y1 - a.rec[(s0 + r) * 8 + 7] / a.oscale[1];
      loss += (double)e0 * e0 + (double)e1 * e1;
      s.DY[0 * 16 + r] = a.inv_batch * e0;
      s.DY[1 * 16 + r] = a.inv_batch * e1;
    }
    __syncthreads();
    if (a.mode == 1) continue;
    // ---- dZ2 = (W3^T dY) * gelu'(Z2)
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
      f32x4 d = mfma(ld_cols(s.W3, kHid, hb * 16 + r, 4 * q), ld_cols(s.DY, 16, r, 4 * q), f32x4{0, 0, 0, 0});
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = hb * 16 + 4 * q + j;
        s.DZ2[row * 16 + r] = d[j] * gelu_grad(s.Z2[row * 16 + r]);
      }
    }
    __syncthreads();
    // ---- dZ1 = (W2^T dZ2) * gelu'(Z1)
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
      f32x4 d = {0, 0, 0, 0};
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
        d = mfma(ld_cols(s.W2, kHid, hb * 16 + r, kb * 16 + 4 * q), ld_cols(s.DZ2, 16, r, kb * 16 + 4 * q), d);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int row = hb * 16 + 4 * q + j;
        s.DZ1[row * 16 + r] = d[j] * gelu_grad(s.Z1[row * 16 + r]);
      }
    }
    __syncthreads();
    // ---- weight gradients: contract over the 16 markers (K)
#pragma unroll
    for (int ib = 0; ib < 4; ++ib)  // dW3[o][i] = sum_m dY[o][m] H2[i][m]
      gW3[ib] = mfma(ld_rows(s.DY, 16, r, 4 * q), ld_rows(s.H2, 16, ib * 16 + r, 4 * q), gW3[ib]);
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
#pragma unroll
      for (int ib = 0; ib < 4; ++ib)  // dW2[o][i] = sum_m dZ2[o][m] H1[i][m]
        gW2[ob * 4 + ib] = mfma(ld_rows(s.DZ2, 16, ob * 16 + r, 4 * q), ld_rows(s.H1, 16, ib * 16 + r, 4 * q),
                                gW2[ob * 4 + ib]);
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)  // dW1[o][i] = sum_m dZ1[o][m] X[i][m]
      gW1[ob] = mfma(ld_rows(s.DZ1, 16, ob * 16 + r, 4 * q), ld_rows(s.X, 16, r, 4 * q), gW1[ob]);
    // bias gradients: lane l owns hidden unit l
    {
      float a1 = 0.0f, a2 = 0.0f;
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        a1 += s.DZ1[l * 16 + m];
        a2 += s.DZ2[l * 16 + m];
      }
      gb1 += a1;
      gb2 += a2;
      if (l < kOut) {
        float a3 = 0.0f;
#pragma unroll
        for (int m = 0; m < 16; ++m) a3 += s.DY[l * 16 + m];
        gb3 += a3;
      }
    }
    __syncthreads();
  }
  if (a.mode >= 1) {
    // loss: fixed shuffle tree over the wave
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (l == 0) a.loss_part[blockIdx.x] = loss;
  }
  if (a.mode != 2) return;
  float* g = a.grad_part + (size_t)blockIdx.x * kNP;
  // D layout: element (row 4q + j, col r) of each 16 x 16 block
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int o = ob * 16 + 4 * q + j;
      if (r < kIn) g[oW1 + o * kIn + r] = gW1[ob][j];
    }
#pragma unroll
  for (int ob = 0; ob < 4; ++ob)
#pragma unroll
    for (int ib = 0; ib < 4; ++ib)
#pragma unroll
      for (int j = 0; j < 4; ++j) g[oW2 + (ob * 16 + 4 * q + j) * kHid + ib * 16 + r] = gW2[ob * 4 + ib][j];
#pragma unroll
  for (int ib = 0; ib < 4; ++ib)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int o = 4 * q + j;
      if (o < kOut) g[oW3 + o * kHid + ib * 16 + r] = gW3[ib][j];
    }
  g[oB1 + l] = gb1;
  g[oB2 + l] = gb2;
  if (l < kOut) g[oB3 + l] = gb3;
}

// grad[p] = sum over rows of part[rows][kNP] (fixed order), then one Adam step
// (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void k_adam(const float* __restrict__ part, int rows, float* __restrict__ params, float* __restrict__ m,
                       float* __restrict__ v, float* __restrict__ grad_out, float lr, float b1c, float b2c,
                       int apply) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNP) return;
  float g = 0.0f;
  for (int r = 0; r < rows; ++r) g += part[(size_t)r * kNP + p];
  if (grad_out) grad_out[p] = g;
  if (!apply) return;
  const float mm = 0.9f * m[p] + 0.1f * g;
  const float vv = 0.999f * v[p] + 0.001f * g * g;
  m[p] = mm;
  v[p] = vv;
  params[p] -= lr * (mm / b1c) / (sqrtf(vv / b2c) + 1e-8f);
}

// Warm-start features of every marker: record [N][8] = (v1, v2, G1, G2, f, |grad f|, -, -)
__global__ void k_features(const double* __restrict__ v, const double* __restrict__ G,
                           const double* __restrict__ fg, int N, float* __restrict__ rec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  float* o = rec + (size_t)i * 8;
  o[0] = (float)v[i];
  o[1] = (float)v[N + i];
  o[2] = (float)G[i];
  o[3] = (float)G[N + i];
  o[4] = (float)fg[i];
  o[5] = (float)fg[N + i];
  o[6] = 0.0f;
  o[7] = 0.0f;
}

// Training target: the converged midpoint velocity minus the Euler velocity
__global__ void k_targets(const double* __restrict__ v_old, const double* __restrict__ v_new,
                          const double* __restrict__ dotv, double dt, int N, float* __restrict__ rec) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  rec[(size_t)i * 8 + 6] = (float)((v_new[i] - v_old[i]) / dt - dotv[i]);
  rec[(size_t)i * 8 + 7] = (float)((v_new[N + i] - v_old[N + i]) / dt - dotv[N + i]);
}

// v1 = v + dt (dot_v_euler + y * out_scale)
__global__ void k_warm_apply(const double* __restrict__ v, const double* __restrict__ dotv,
                             const float* __restrict__ y, float os0, float os1, double dt, int N,
                             double* __restrict__ v1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  v1[i] = v[i] + dt * (dotv[i] + (double)(y[2 * (size_t)i] * os0));
  v1[N + i] = v[N + i] + dt * (dotv[N + i] + (double)(y[2 * (size_t)i + 1] * os1));
}

}  // namespace

void launch_features(const double* v, const double* G, const double* fg, int N, float* rec, hipStream_t st) {
  hipLaunchKernelGGL(k_features, dim3((N + 255) / 256), dim3(256), 0, st, v, G, fg, N, rec);
}

void launch_targets(const double* v_old, const double* v_new, const double* dotv, double dt, int N, float* rec,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_targets, dim3((N + 255) / 256), dim3(256), 0, st, v_old, v_new, dotv, dt, N, rec);
}

void launch_warm_apply(const double* v, const double* dotv, const float* y, float os0, float os1, double dt, int N,
                       double* v1, hipStream_t st) {
  hipLaunchKernelGGL(k_warm_apply, dim3((N + 255) / 256), dim3(256), 0, st, v, dotv, y, os0, os1, dt, N, v1);
}

void launch_mlp(const MlpArgs& a, int grid, hipStream_t st) {
  if (a.ntiles <= 0) return;
  hipLaunchKernelGGL(k_mlp, dim3(grid), dim3(64), 0, st, a);
}

void launch_adam(const float* part, int rows, float* params, float* m, float* v, float* grad_out, float lr, int step,
                 int apply, hipStream_t st) {
  const float b1c = 1.0f - powf(0.9f, (float)step), b2c = 1.0f - powf(0.999f, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((kNP + 255) / 256), dim3(256), 0, st, part, rows, params, m, v, grad_out, lr, b1c,
                     b2c, apply);
}

}  // namespace collop
