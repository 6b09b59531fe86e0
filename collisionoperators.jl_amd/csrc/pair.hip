// pair.hip -- the O(N^2) Landau pair sum on gfx950 (kernel (a)).
//
//   dot_v_g = sum_{a != g, a in box, |d|^2 >= 1e-24} w_a U(d) (G_a - G_g),
//   d = v_g - v_a,  U(d) = (I - d d^T / |d|^2) / |d|        (functions.jl:205-237)
//
// In 2D, U(d) g = c (d2, -d1) / |d|^3 with c = g1 d2 - g2 d1 (the 2D cross
// product), so one pair costs two subtractions for d, a dot for |d|^2, one
// rsqrt, two subtractions for g, one cross product and two multiply-adds.
// The pair term is odd under (g <-> a): U(-d) = U(d) and g changes sign, so
// the FP32 variant keeps each pair contribution exactly antisymmetric.
//
// Mapping to CDNA4: one target gamma (or two) per lane, held in VGPRs; the
// source alpha is wave-uniform, so its 64-byte record is fetched with scalar
// loads into SGPRs and broadcast to the 64 lanes for free -- no LDS traffic and
// no barriers.  The alpha range is cut into S slices; workgroup id -> (slice =
// id % S, target block = id / S), so with S a multiple of 8 every XCD (which
// receives workgroups round-robin) streams only its own slices of the source
// array through its private L2.  Slice partials are summed in a fixed order by
// a second pass, so results are deterministic run to run.  No MFMA: this is
// not a dense contraction.  FP64 1/sqrt is v_rsq_f64 refined by Newton steps
// y <- y (3/2 - q y^2 / 2) (textbook Newton-Raphson on 1/y^2 = q).
#include "collop_launch.hpp"

namespace collop {
namespace {

constexpr int kThreads = 256;

template <int K, int NEWTON, bool METRIC, bool NEAR = false>
__global__ __launch_bounds__(kThreads) void k_pair_f64(const Src64* __restrict__ src, int N, int S, int chunk,
                                                        double* __restrict__ part, NearOut near = NearOut{}) {
  if (near.live && *near.live == 0) return;
  const int bid = blockIdx.x;
  const int slice = bid % S;
  const int gblk = bid / S;
  const int a0 = slice * chunk;
  const int a1 = min(N, a0 + chunk);
  double vg1[K], vg2[K], Gg1[K], Gg2[K], acc1[K], acc2[K], m11[K], m12[K], m22[K];
  bool act[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int g = gblk * (kThreads * K) + k * kThreads + (int)threadIdx.x;
    act[k] = false;
    vg1[k] = vg2[k] = Gg1[k] = Gg2[k] = 0.0;
    if (g < N) {
      const Src64 s = src[g];
      vg1[k] = s.v1;
      vg2[k] = s.v2;
      Gg1[k] = s.g1;
      Gg2[k] = s.g2;
      act[k] = s.in_domain != 0.0;
    }
    acc1[k] = acc2[k] = m11[k] = m12[k] = m22[k] = 0.0;
  }
#pragma unroll 2
  for (int a = a0; a < a1; ++a) {
    const double av1 = src[a].v1, av2 = src[a].v2, ag1 = src[a].g1, ag2 = src[a].g2, aw = src[a].w;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      double d1, d2, t;
      pair_term_f64<NEWTON>(vg1[k], vg2[k], Gg1[k], Gg2[k], av1, av2, ag1, ag2, aw, acc1[k], acc2[k], d1, d2, t);
      if (METRIC) {
        // A += w U(d) = w (I |d|^2 - d d^T) / |d|^3  (functions.jl:278-281)
        const double td2 = t * d2;
        m11[k] = __builtin_fma(td2, d2, m11[k]);
        m12[k] = __builtin_fma(-td2, d1, m12[k]);
        const double td1 = t * d1;
        m22[k] = __builtin_fma(td1, d1, m22[k]);
      }
      if (NEAR) {  // near pair: w_a / |d| > theta
        const double q = __builtin_fma(d2, d2, d1 * d1);
        if (q < near.qmax && act[k] && q >= 1e-24 && near.th2 * q < aw * aw) {
          const int g = gblk * (kThreads * K) + k * kThreads + (int)threadIdx.x;
          const int idx = atomicAdd(near.count, 1);
          if (idx < near.cap) near.list[idx] = make_int2(g, a);
        }
      }
    }
  }
  const size_t stride = (size_t)(METRIC ? 5 : 2) * (size_t)N;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int g = gblk * (kThreads * K) + k * kThreads + (int)threadIdx.x;
    if (g < N) {
      double* p = part + (size_t)slice * stride;
      p[g] = act[k] ? acc1[k] : 0.0;
      p[N + g] = act[k] ? acc2[k] : 0.0;
      if (METRIC) {
        p[2 * (size_t)N + g] = act[k] ? m11[k] : 0.0;
        p[3 * (size_t)N + g] = act[k] ? m12[k] : 0.0;
        p[4 * (size_t)N + g] = act[k] ? m22[k] : 0.0;
      }
    }
  }
}

template <int K, int BLK, bool METRIC, bool NEAR = false>
__global__ __launch_bounds__(kThreads) void k_pair_f32c(const Src32* __restrict__ src, int N, int S, int chunk,
                                                         double* __restrict__ part, NearOut near = NearOut{}) {
  if (near.live && *near.live == 0) return;
  const int bid = blockIdx.x;
  const int slice = bid % S;
  const int gblk = bid / S;
  const int a0 = slice * chunk;
  const int a1 = min(N, a0 + chunk);
  float vg1[K], vg2[K], Gg1[K], Gg2[K], b1[K], b2[K], h1[K], h2[K], l1[K], l2[K];
  // metric A = sum w U(d): block sums folded into FP32 (hi, lo) like the velocity sums
  float c11[K], c12[K], c22[K], q11[K], q12[K], q22[K], e11[K], e12[K], e22[K];
  bool act[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int g = gblk * (kThreads * K) + k * kThreads + (int)threadIdx.x;
    act[k] = false;
    vg1[k] = vg2[k] = Gg1[k] = Gg2[k] = 0.0f;
    if (g < N) {
      const Src32 s = src[g];
      vg1[k] = s.v1;
      vg2[k] = s.v2;
      Gg1[k] = s.g1;
      Gg2[k] = s.g2;
      act[k] = s.in_domain != 0.0f;
    }
    b1[k] = b2[k] = h1[k] = h2[k] = l1[k] = l2[k] = 0.0f;
    c11[k] = c12[k] = c22[k] = q11[k] = q12[k] = q22[k] = e11[k] = e12[k] = e22[k] = 0.0f;
  }
  for (int ab = a0; ab < a1; ab += BLK) {
    const int ae = min(a1, ab + BLK);
#pragma unroll 4
    for (int a = ab; a < ae; ++a) {
      const float av1 = src[a].v1, av2 = src[a].v2, ag1 = src[a].g1, ag2 = src[a].g2, aw = src[a].w;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (METRIC) {
          float d1, d2, t;
          pair_term_f32_t(vg1[k], vg2[k], Gg1[k], Gg2[k], av1, av2, ag1, ag2, aw, b1[k], b2[k], d1, d2, t);
          // A += w U(d) = w (I |d|^2 - d d^T) / |d|^3  (functions.jl:278-281)
          const float td2 = t * d2, td1 = t * d1;
          c11[k] = __builtin_fmaf(td2, d2, c11[k]);
          c12[k] = __builtin_fmaf(-td2, d1, c12[k]);
          c22[k] = __builtin_fmaf(td1, d1, c22[k]);
          if (NEAR) {  // near pair: w_a / |d| > theta
            const float q = __builtin_fmaf(d2, d2, d1 * d1);
            if (q < (float)near.qmax && act[k] && q >= 1e-12f && (float)near.th2 * q < aw * aw) {
              const int g = gblk * (kThreads * K) + k * kThreads + (int)threadIdx.x;
              const int idx = atomicAdd(near.count, 1);
              if (idx < near.cap) near.list[idx] = make_int2(g, a);
            }
          }
        } else {
          pair_term_f32(vg1[k], vg2[k], Gg1[k], Gg2[k], av1, av2, ag1, ag2, aw, b1[k], b2[k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      two_sum_fold(h1[k], l1[k], b1[k]);
      two_sum_fold(h2[k], l2[k], b2[k]);
      b1[k] = b2[k] = 0.0f;
      if (METRIC) {
        two_sum_fold(q11[k], e11[k], c11[k]);
        two_sum_fold(q12[k], e12[k], c12[k]);
        two_sum_fold(q22[k], e22[k], c22[k]);
        c11[k] = c12[k] = c22[k] = 0.0f;
      }
    }
  }
  const size_t stride = (size_t)(METRIC ? 5 : 2) * (size_t)N;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int g = gblk * (kThreads * K) + k * kThreads + (int)threadIdx.x;
    if (g < N) {
      double* p = part + (size_t)slice * stride;
      p[g] = act[k] ? (double)h1[k] + (double)l1[k] : 0.0;
      p[N + g] = act[k] ? (double)h2[k] + (double)l2[k] : 0.0;
      if (METRIC) {
        p[2 * (size_t)N + g] = act[k] ? (double)q11[k] + (double)e11[k] : 0.0;
        p[3 * (size_t)N + g] = act[k] ? (double)q12[k] + (double)e12[k] : 0.0;
        p[4 * (size_t)N + g] = act[k] ? (double)q22[k] + (double)e22[k] : 0.0;
      }
    }
  }
}

__device__ __forceinline__ void pack_body(const double* __restrict__ v, const double* __restrict__ G,
                                          const double* __restrict__ w, int N, int ld, double lo1, double hi1,
                                          double lo2, double hi2, Src64* __restrict__ s64, Src32* __restrict__ s32,
                                          int i) {
  if (i >= N) return;
  const double a = v[i], b = v[ld + i];
  const bool in = in_box(a, b, lo1, hi1, lo2, hi2);
  const double wm = in ? w[i] : 0.0;
  if (s64) {
    Src64 s;
    s.v1 = a;
    s.v2 = b;
    s.g1 = G[i];
    s.g2 = G[ld + i];
    s.w = wm;
    s.in_domain = in ? 1.0 : 0.0;
    s.pad0 = s.pad1 = 0.0;
    s64[i] = s;
  }
  if (s32) {
    Src32 s;
    s.v1 = (float)a;
    s.v2 = (float)b;
    s.g1 = (float)G[i];
    s.g2 = (float)G[ld + i];
    s.w = (float)wm;
    s.in_domain = in ? 1.0f : 0.0f;
    s.pad0 = s.pad1 = 0.0f;
    s32[i] = s;
  }
}

__global__ void k_pack(const double* __restrict__ v, const double* __restrict__ G, const double* __restrict__ w,
                       int N, double lo1, double hi1, double lo2, double hi2, Src64* __restrict__ s64,
                       Src32* __restrict__ s32) {
  pack_body(v, G, w, N, N, lo1, hi1, lo2, hi2, s64, s32, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_pack_b(EnsView E, const double* __restrict__ v, const double* __restrict__ G,
                         const double* __restrict__ w, double lo1, double hi1, double lo2, double hi2,
                         Src64* __restrict__ s64, Src32* __restrict__ s32) {
  const int2 t = E.pblk[blockIdx.x];
  const EnsRun& R = E.runs[t.x];
  pack_body(v + R.off, G + R.off, w + R.off, R.n, E.NT, lo1, hi1, lo2, hi2, s64 ? s64 + R.off : nullptr,
            s32 ? s32 + R.off : nullptr, t.y * blockDim.x + threadIdx.x);
}

// Batched pair sum: one persistent launch walks the tile list (longest tiles
// first) with a grid stride.  A tile is BT consecutive targets of one run
// against one source slice of that run; each lane owns one target and runs the
// same per-target loop as k_pair_f64<1, 1> / k_pair_f32c<1, 32>, so a run's
// partials equal the single-run kernel's at the same slice count.
template <int BT>
__global__ __launch_bounds__(BT) void k_pair_b(EnsView E, const int4* __restrict__ tiles, int ntiles,
                                               const Src64* __restrict__ s64, const Src32* __restrict__ s32,
                                               int force_f64, double* __restrict__ part) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int4 tl = tiles[t];
    const int r = tl.x;
    const bool f64 = force_f64 || E.ctl[r].kind == 0;
    if (!force_f64 && !E.ctl[r].active) continue;
    const EnsRun& R = E.runs[r];
    const int n = R.n, S = R.S;
    const int chunk = (n + S - 1) / S;
    const int a0 = R.off + tl.z * chunk;
    const int a1 = R.off + min(n, tl.z * chunk + chunk);
    const int loc = tl.y + (int)threadIdx.x;
    double* pr = part + R.poff + (size_t)tl.z * 2 * n;
    if (f64) {
      double vg1 = 0.0, vg2 = 0.0, Gg1 = 0.0, Gg2 = 0.0, acc1 = 0.0, acc2 = 0.0;
      bool act = false;
      if (loc < n) {
        const Src64 q = s64[R.off + loc];
        vg1 = q.v1;
        vg2 = q.v2;
        Gg1 = q.g1;
        Gg2 = q.g2;
        act = q.in_domain != 0.0;
      }
#pragma unroll 2
      for (int a = a0; a < a1; ++a) {
        const double av1 = s64[a].v1, av2 = s64[a].v2, ag1 = s64[a].g1, ag2 = s64[a].g2, aw = s64[a].w;
        double d1, d2, tt;
        pair_term_f64<1>(vg1, vg2, Gg1, Gg2, av1, av2, ag1, ag2, aw, acc1, acc2, d1, d2, tt);
      }
      if (loc < n) {
        pr[loc] = act ? acc1 : 0.0;
        pr[n + loc] = act ? acc2 : 0.0;
      }
    } else {
      float vg1 = 0.0f, vg2 = 0.0f, Gg1 = 0.0f, Gg2 = 0.0f, b1 = 0.0f, b2 = 0.0f, h1 = 0.0f, h2 = 0.0f, l1 = 0.0f,
            l2 = 0.0f;
      bool act = false;
      if (loc < n) {
        const Src32 q = s32[R.off + loc];
        vg1 = q.v1;
        vg2 = q.v2;
        Gg1 = q.g1;
        Gg2 = q.g2;
        act = q.in_domain != 0.0f;
      }
      for (int ab = a0; ab < a1; ab += 32) {
        const int ae = min(a1, ab + 32);
#pragma unroll 4
        for (int a = ab; a < ae; ++a) {
          const float av1 = s32[a].v1, av2 = s32[a].v2, ag1 = s32[a].g1, ag2 = s32[a].g2, aw = s32[a].w;
          pair_term_f32(vg1, vg2, Gg1, Gg2, av1, av2, ag1, ag2, aw, b1, b2);
        }
        two_sum_fold(h1, l1, b1);
        two_sum_fold(h2, l2, b2);
        b1 = b2 = 0.0f;
      }
      if (loc < n) {
        pr[loc] = act ? (double)h1 + (double)l1 : 0.0;
        pr[n + loc] = act ? (double)h2 + (double)l2 : 0.0;
      }
    }
  }
}

__global__ void k_slice_sum(const double* __restrict__ part, int S, int64_t len, double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  out[i] = slot_sum(part, S, (size_t)len, (size_t)i);
}

int grid_for(int N, const PairTile& t) {
  const int per_block = kThreads * t.gamma_per_lane;
  return ((N + per_block - 1) / per_block) * t.slices;
}

}  // namespace

void launch_pack_src(const double* v, const double* G, const double* w, int N, double lo1, double hi1, double lo2,
                     double hi2, Src64* s64, Src32* s32, hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_pack, dim3((N + 255) / 256), dim3(256), 0, st, v, G, w, N, lo1, hi1, lo2, hi2, s64, s32);
}

void launch_pair_f64(const Src64* src, int N, const PairTile& t, double* part, hipStream_t st, const int* live,
                     bool pow2w) {
  if (N <= 0) return;
  if (t.sym) return launch_pair_sym_f64(src, N, t.sym, false, nullptr, part, st, live, pow2w);
  const int chunk = (N + t.slices - 1) / t.slices;
  const dim3 grid(grid_for(N, t)), block(kThreads);
  NearOut g{};
  g.live = live;
  if (t.gamma_per_lane == 2) {
    if (t.newton == 1)
      hipLaunchKernelGGL((k_pair_f64<2, 1, false>), grid, block, 0, st, src, N, t.slices, chunk, part, g);
    else
      hipLaunchKernelGGL((k_pair_f64<2, 2, false>), grid, block, 0, st, src, N, t.slices, chunk, part, g);
  } else {
    if (t.newton == 1)
      hipLaunchKernelGGL((k_pair_f64<1, 1, false>), grid, block, 0, st, src, N, t.slices, chunk, part, g);
    else
      hipLaunchKernelGGL((k_pair_f64<1, 2, false>), grid, block, 0, st, src, N, t.slices, chunk, part, g);
  }
}

void launch_pair_metric_f64(const Src64* src, int N, const PairTile& t, double* part, hipStream_t st,
                            const NearOut* near) {
  if (N <= 0) return;
  if (t.sym) return launch_pair_sym_f64(src, N, t.sym, true, near, part, st);
  const int chunk = (N + t.slices - 1) / t.slices;
  const dim3 grid(grid_for(N, t)), block(kThreads);
  if (near) {  // one target per lane, one Newton step
    hipLaunchKernelGGL((k_pair_f64<1, 1, true, true>), dim3(grid_for(N, PairTile{1, t.slices, 1})), block, 0, st, src,
                       N, t.slices, chunk, part, *near);
    return;
  }
  if (t.gamma_per_lane == 2) {
    if (t.newton == 1)
      hipLaunchKernelGGL((k_pair_f64<2, 1, true>), grid, block, 0, st, src, N, t.slices, chunk, part);
    else
      hipLaunchKernelGGL((k_pair_f64<2, 2, true>), grid, block, 0, st, src, N, t.slices, chunk, part);
  } else {
    if (t.newton == 1)
      hipLaunchKernelGGL((k_pair_f64<1, 1, true>), grid, block, 0, st, src, N, t.slices, chunk, part);
    else
      hipLaunchKernelGGL((k_pair_f64<1, 2, true>), grid, block, 0, st, src, N, t.slices, chunk, part);
  }
}

void launch_pair_f32c(const Src32* src, int N, const PairTile& t, double* part, hipStream_t st, const int* live,
                      bool pow2w) {
  if (N <= 0) return;
  if (t.sym) return launch_pair_sym_f32c(src, N, t.sym, false, nullptr, part, st, live, pow2w);
  const int chunk = (N + t.slices - 1) / t.slices;
  const dim3 grid(grid_for(N, t)), block(kThreads);
  NearOut g{};
  g.live = live;
  if (t.gamma_per_lane == 2)
    hipLaunchKernelGGL((k_pair_f32c<2, 32, false>), grid, block, 0, st, src, N, t.slices, chunk, part, g);
  else
    hipLaunchKernelGGL((k_pair_f32c<1, 32, false>), grid, block, 0, st, src, N, t.slices, chunk, part, g);
}

void launch_pair_metric_f32c(const Src32* src, int N, const PairTile& t, double* part, hipStream_t st,
                             const NearOut* near, bool pow2w) {
  if (N <= 0) return;
  if (t.sym) return launch_pair_sym_f32c(src, N, t.sym, true, near, part, st, nullptr, pow2w && !near);
  const int chunk = (N + t.slices - 1) / t.slices;
  const dim3 grid(grid_for(N, t)), block(kThreads);
  if (near) {
    hipLaunchKernelGGL((k_pair_f32c<1, 32, true, true>), dim3(grid_for(N, PairTile{1, t.slices, 0})), block, 0, st,
                       src, N, t.slices, chunk, part, *near);
    return;
  }
  if (t.gamma_per_lane == 2)
    hipLaunchKernelGGL((k_pair_f32c<2, 32, true>), grid, block, 0, st, src, N, t.slices, chunk, part);
  else
    hipLaunchKernelGGL((k_pair_f32c<1, 32, true>), grid, block, 0, st, src, N, t.slices, chunk, part);
}

void launch_slice_sum(const double* part, int S, int64_t len, double* out, hipStream_t st) {
  if (len <= 0) return;
  hipLaunchKernelGGL(k_slice_sum, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, part, S, len, out);
}

void launch_pack_b(const EnsView& E, const double* v, const double* G, const double* w, double lo1, double hi1,
                   double lo2, double hi2, Src64* s64, Src32* s32, hipStream_t st) {
  hipLaunchKernelGGL(k_pack_b, dim3(E.n_pblk), dim3(256), 0, st, E, v, G, w, lo1, hi1, lo2, hi2, s64, s32);
}

void launch_pair_b(const EnsView& E, const EnsTiles& T, const Src64* s64, const Src32* s32, int force_f64,
                   double* part, hipStream_t st) {
  if (T.ntiles <= 0) return;
  if (T.sym) return launch_pair_sym_b(E, T, s64, s32, force_f64, part, st);
  const dim3 grid(T.grid);
  switch (T.bt) {
    case 64:
      hipLaunchKernelGGL(k_pair_b<64>, grid, dim3(64), 0, st, E, T.tiles, T.ntiles, s64, s32, force_f64, part);
      break;
    case 128:
      hipLaunchKernelGGL(k_pair_b<128>, grid, dim3(128), 0, st, E, T.tiles, T.ntiles, s64, s32, force_f64, part);
      break;
    case 512:
      hipLaunchKernelGGL(k_pair_b<512>, grid, dim3(512), 0, st, E, T.tiles, T.ntiles, s64, s32, force_f64, part);
      break;
    default:
      hipLaunchKernelGGL(k_pair_b<256>, grid, dim3(256), 0, st, E, T.tiles, T.ntiles, s64, s32, force_f64, part);
      break;
  }
}

}  // namespace collop
