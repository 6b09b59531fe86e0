// pair_sym.hip -- the Landau pair sum with every unordered pair evaluated once.
//
// The pair term is odd under exchange of target and source: with d = v_g - v_a
// and g = G_a - G_g, target g receives  w_a c (d2, -d1)/|d|^3  and target a
// receives  -w_g c (d2, -d1)/|d|^3  (c = g1 d2 - g2 d1 is even: both d and g
// change sign), and the metric term U(d) = U(-d) is even.  One evaluation of
// c/|d|^3 (and of U(d)/|d|^3 for the metric) therefore serves both ordered
// pairs; the ordered sum of functions.jl:205-237 / collision_gpu.jl:49-107
// evaluates each of them separately.  Masking is the same: w_masked = 0 for a
// source outside the box, output 0 for a target outside it, and no term for
// |d|^2 < 1e-24 (functions.jl:219-224).
//
// Partition.  Markers are cut into super-rows of E = 128 R markers; workgroup t
// (R waves) handles one block (a, b) of super-rows, a <= b:
//  * a < b: wave r owns 128 targets of super-row a, two per lane, in VGPRs.
//    Super-row b is staged in LDS as 2R tiles of 64 sources.  In phase p wave r
//    meets tile (2r + p) mod 2R, so each tile is met by exactly one wave per
//    phase and each (target, source) pair of the block exactly once.  Within a
//    phase lane i pairs its targets with source (i + k) mod 64 at step k; the
//    sources' accumulators travel with them, rotated by one lane per step with
//    DPP (wave_rol:1), and return to LDS when the phase ends.
//  * a == b: the ordered sum over the super-row (every pair from both sides),
//    with the source record broadcast through scalar loads as in k_pair_f64.
// Partial slots.  A marker of super-row s receives one partial from every block
// that contains s, in slot o = the block's other super-row (o = s for the
// diagonal block), so slot o of part[o][C][N] is written exactly once for every
// marker.  The S = #super-rows slots are summed in a fixed order downstream:
// results are deterministic and independent of the order blocks run in.
//
// Per unordered pair (FP64): 2 subtractions and a dot for |d|^2, the guard,
// v_rsq_f64 + one Newton step, |d|^-3, the cross product, then two
// multiply-adds per side and the DPP rotation of the source accumulator
// (shared by the lane's two targets): about 26 VALU slots per unordered pair
// against 23 per ordered pair in k_pair_f64.
#include "collop_launch.hpp"

#include <type_traits>

namespace collop {
namespace {

// Block t -> super-rows (a, b): the S diagonal blocks first (their ordered sums
// cost more per pair, and the grid's tail then holds only off-diagonal blocks),
// then the off-diagonal blocks (a < b), ordered by b then a.  The grid has
// n_off + S blocks.
__device__ __forceinline__ void block_rows(long long t, long long n_off, int& a, int& b) {
  const long long S = (long long)gridDim.x - n_off;
  if (t < S) {
    a = b = (int)t;
    return;
  }
  t -= S;
  long long bb = (long long)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
  while (bb * (bb - 1) / 2 > t) --bb;
  while ((bb + 1) * bb / 2 <= t) ++bb;
  b = (int)bb;
  a = (int)(t - bb * (bb - 1) / 2);
}

// |d|^2 below the coincidence threshold -> about 1e300 (high word of 1e300, low
// word of q), so that 1/|d|^3 underflows to exactly 0 and the pair adds nothing
// on either side.
__device__ __forceinline__ double guard_q64(double q) {
  const int hi = q < 1e-24 ? 0x7E37E43C : __double2hiint(q);
  return __hiloint2double(hi, __double2loint(q));
}
__device__ __forceinline__ float guard_q32(float q) { return q < 1e-12f ? 3.0e38f : q; }

// q^(-3/2) from y0 = v_rsq_f64(q) (relative error eps below 2^-24, measured
// 5.2e-8 on gfx950): with f = 1 - q y0^2, q^(-3/2) = y0^3 (1 - f)^(-3/2)
// = y0^3 (1 + 1.5 f + 1.875 f^2 + ...); the first-order form leaves
// 1.875 f^2 ~ 7.5 eps^2 < 2.1e-14 relative, as one Newton step on y followed
// by cubing (4.5 eps^2), in five operations instead of six.
__device__ __forceinline__ double rsqrt3_f64(double q) {
  const double y0 = __builtin_amdgcn_rsq(q);
  const double y02 = y0 * y0;
  const double e3 = __builtin_fma(-1.5 * q, y02, 1.5);  // 1.5 f
  const double z0 = y02 * y0;
  return __builtin_fma(z0, e3, z0);
}

// Lane i receives the value of lane (i + 1) mod 64 (every lane has a source,
// so no lane keeps an old value).
__device__ __forceinline__ float rot1(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x134, 0xf, 0xf, false));
}
__device__ __forceinline__ double rot1(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), 0x134, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), 0x134, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ void near_push(const NearOut& near, int g, int a) {
  const int idx = atomicAdd(near.count, 1);
  if (idx < near.cap) near.list[idx] = make_int2(g, a);
}

// LDS of one FP64 block: the sources of super-row b by tile -- (v1, v2),
// (G1, G2), w_masked of its 64 sources, each stored twice in a row so that
// source (i + k) mod 64 sits at offset i + k and the fully unrolled step loop
// reads it with two 16-byte and one 8-byte LDS load at immediate offsets from
// one address -- and their accumulators.
// With equal weights (POW2W) the tiles carry no weights: 40 KB at R = 4, four
// blocks per CU.
template <int R, bool METRIC, bool POW2W>
struct SymLds64Core {
  struct TileW {
    double2 v[128];
    double2 g[128];
    double w[128];
  };
  struct TileP {
    double2 v[128];
    double2 g[128];
  };
  using Tile = std::conditional_t<POW2W, TileP, TileW>;
  Tile L[2 * R];
  double Lacc[METRIC ? 5 : 2][128 * R];
};
template <int R, bool METRIC, bool NEAR, bool POW2W = false>
struct SymLds64 : SymLds64Core<R, METRIC, POW2W> {
  int Lin[128 * R];  // sources in the box (near-pair search)
};
template <int R, bool METRIC, bool POW2W>
struct SymLds64<R, METRIC, false, POW2W> : SymLds64Core<R, METRIC, POW2W> {};

// Equal power-of-two weights (POW2W): every marker carries the same weight w =
// 2^k, so the pair's two scale factors w_a c / |d|^3 and w_g c / |d|^3 are one,
// and w multiplies G exactly.  The off-diagonal blocks then take G' = w G and
// c' = w c (exact scaling: the same bits as w * (y^3 c)), and drop both weight
// products per pair.  A target outside the box (w_masked = 0) sits at the far
// point (kFar, kFar), a source outside the box or a padding slot past N at
// (-kFar, -kFar), both with G' = 0: every pair with a far marker gives
// |d|^-3 = 0 exactly (underflow) and c' = 0, so it adds nothing on either
// side, as the mask does, and no two far markers coincide.
constexpr double kFar = 1e150;

// Block (a, b) of the symmetric FP64 sum over src[0, N) into part[S][C][N]:
// super-rows of E = 128 R markers, R waves, Q = 2 targets per lane.  (Four
// targets per lane on R/2 waves halve the rotations per pair and measured the
// same time.)
template <int R, bool METRIC, bool NEAR, bool POW2W = false>
__device__ __forceinline__ void sym_block_f64(const Src64* __restrict__ src, int N, int a, int b,
                                              double* __restrict__ part, SymLds64<R, METRIC, NEAR, POW2W>& S,
                                              const NearOut& near) {
  static_assert(!(POW2W && (METRIC || NEAR)), "the equal-weight form is the plain pair sum only");
  constexpr int Q = 2, E = 128 * R, NTL = 2 * R, NW = R, C = METRIC ? 5 : 2;
  using Tile = typename SymLds64<R, METRIC, NEAR, POW2W>::Tile;
  auto& L = S.L;
  auto& Lacc = S.Lacc;
  const int lane = threadIdx.x & 63, r = threadIdx.x >> 6;
  const size_t cs = (size_t)C * N;
  double gv1[Q], gv2[Q], gg1[Q], gg2[Q], gw[Q];
  double f1[Q], f2[Q], m11[Q], m12[Q], m22[Q];
  bool gin[Q];
  int gm[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    gm[q] = a * E + r * 64 * Q + q * 64 + lane;
    gv1[q] = gv2[q] = gg1[q] = gg2[q] = gw[q] = 0.0;
    gin[q] = false;
    if (gm[q] < N) {
      const Src64 s = src[gm[q]];
      gv1[q] = s.v1;
      gv2[q] = s.v2;
      gg1[q] = s.g1;
      gg2[q] = s.g2;
      gw[q] = s.w;
      gin[q] = s.in_domain != 0.0;
    }
    f1[q] = f2[q] = m11[q] = m12[q] = m22[q] = 0.0;
  }
  if (a == b) {  // diagonal block: ordered sum over the super-row
    const int s0 = a * E, s1 = min(N, s0 + E);
#pragma unroll 2
    for (int s = s0; s < s1; ++s) {
      const double av1 = src[s].v1, av2 = src[s].v2, ag1 = src[s].g1, ag2 = src[s].g2, aw = src[s].w;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        double d1, d2, t;
        pair_term_f64<1>(gv1[q], gv2[q], gg1[q], gg2[q], av1, av2, ag1, ag2, aw, f1[q], f2[q], d1, d2, t);
        if constexpr (METRIC) {
          const double td2 = t * d2, td1 = t * d1;
          m11[q] = __builtin_fma(td2, d2, m11[q]);
          m12[q] = __builtin_fma(-td2, d1, m12[q]);
          m22[q] = __builtin_fma(td1, d1, m22[q]);
        }
        if constexpr (NEAR) {
          const double qd = __builtin_fma(d2, d2, d1 * d1);
          if (qd < near.qmax && gin[q] && qd >= 1e-24 && near.th2 * qd < aw * aw) near_push(near, gm[q], s);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      if (gm[q] >= N) continue;
      double* p = part + (size_t)a * cs + gm[q];
      p[0] = gin[q] ? f1[q] : 0.0;
      p[N] = gin[q] ? f2[q] : 0.0;
      if constexpr (METRIC) {
        p[2 * (size_t)N] = gin[q] ? m11[q] : 0.0;
        p[3 * (size_t)N] = gin[q] ? m12[q] : 0.0;
        p[4 * (size_t)N] = gin[q] ? m22[q] : 0.0;
      }
    }
    return;
  }
  if constexpr (POW2W) {
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      gv1[q] = gin[q] ? gv1[q] : kFar;
      gv2[q] = gin[q] ? gv2[q] : kFar;
      gg1[q] = gin[q] ? gw[q] * gg1[q] : 0.0;
      gg2[q] = gin[q] ? gw[q] * gg2[q] : 0.0;
    }
  }
  // sources of super-row b into the tiles, their accumulators cleared
  auto stage = [&]() {
    for (int e = threadIdx.x; e < E; e += 64 * NW) {
      const int m = b * E + e;
      double x1 = 0.0, x2 = 0.0, y1 = 0.0, y2 = 0.0, w = 0.0;
      int in = 0;
      if (m < N) {
        const Src64 s = src[m];
        x1 = s.v1;
        x2 = s.v2;
        y1 = s.g1;
        y2 = s.g2;
        w = s.w;
        in = s.in_domain != 0.0;
      }
      Tile& t = L[e >> 6];
      const int j = e & 63;
      if constexpr (POW2W) {  // padding past N too: the weight no longer masks it
        x1 = in ? x1 : -kFar;
        x2 = in ? x2 : -kFar;
        y1 = in ? w * y1 : 0.0;
        y2 = in ? w * y2 : 0.0;
      } else {
        t.w[j] = t.w[j + 64] = w;
      }
      t.v[j] = t.v[j + 64] = make_double2(x1, x2);
      t.g[j] = t.g[j + 64] = make_double2(y1, y2);
      if constexpr (NEAR) S.Lin[e] = in;
#pragma unroll
      for (int c = 0; c < C; ++c) Lacc[c][e] = 0.0;
    }
  };
  stage();
  __syncthreads();
  // GUARD = false: the pairs run without the coincidence guard while each lane
  // keeps the smallest high word of |d|^2; the caller reruns the block guarded
  // if any pair came near the threshold (the same bits: the guard leaves every
  // |d|^2 >= 1e-24 as it is).
  int qhi_min = 0x7FFFFFFF;
  auto tiles = [&](auto guard_tag) {
  constexpr bool GUARD = decltype(guard_tag)::value;
  for (int p = 0; p < NTL; ++p) {
    const int tile = (Q * r + p) % NTL;  // the waves hold distinct tiles at every phase
    const int base = tile * 64;
    const double2* pv = &L[tile].v[lane];
    const double2* pg = &L[tile].g[lane];
    const double* pw = nullptr;
    if constexpr (!POW2W) pw = &L[tile].w[lane];
    // step k: lane i meets source base + (i + k) mod 64 and holds its accumulator
    double b1 = Lacc[0][base + lane], b2 = Lacc[1][base + lane];
    double n11 = 0.0, n12 = 0.0, n22 = 0.0;
    if constexpr (METRIC) {
      n11 = Lacc[2][base + lane];
      n12 = Lacc[3][base + lane];
      n22 = Lacc[4][base + lane];
    }
    auto step = [&](int k) {
      const int ia = base + ((lane + k) & 63);  // source index (near pairs only)
      const double2 av = pv[k], ag = pg[k];
      const double av1 = av.x, av2 = av.y, ag1 = ag.x, ag2 = ag.y;
      double aw = 0.0;
      if constexpr (!POW2W) aw = pw[k];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const double d1 = gv1[q] - av1;
        const double d2 = gv2[q] - av2;
        const double u22 = d1 * d1;
        const double qd = __builtin_fma(d2, d2, u22);
        double y3;
        if constexpr (GUARD) {
          y3 = rsqrt3_f64(guard_q64(qd));
        } else {
          qhi_min = min(qhi_min, __double2hiint(qd));
          y3 = rsqrt3_f64(qd);
        }
        const double c = __builtin_fma(ag1 - gg1[q], d2, -((ag2 - gg2[q]) * d1));
        if constexpr (POW2W) {  // c is w c: one factor serves both sides
          const double pc = y3 * c;
          f1[q] = __builtin_fma(pc, d2, f1[q]);
          f2[q] = __builtin_fma(-pc, d1, f2[q]);
          b1 = __builtin_fma(-pc, d2, b1);
          b2 = __builtin_fma(pc, d1, b2);
        } else if constexpr (!METRIC) {
          const double pc = y3 * c;
          const double kg = aw * pc, ka = gw[q] * pc;
          f1[q] = __builtin_fma(kg, d2, f1[q]);
          f2[q] = __builtin_fma(-kg, d1, f2[q]);
          b1 = __builtin_fma(-ka, d2, b1);
          b2 = __builtin_fma(ka, d1, b2);
        } else {
          const double sg = aw * y3, sa = gw[q] * y3;
          const double kg = sg * c, ka = sa * c;
          f1[q] = __builtin_fma(kg, d2, f1[q]);
          f2[q] = __builtin_fma(-kg, d1, f2[q]);
          b1 = __builtin_fma(-ka, d2, b1);
          b2 = __builtin_fma(ka, d1, b2);
          // A += w U(d) = w (I |d|^2 - d d^T) / |d|^3  (functions.jl:278-281)
          const double u11 = d2 * d2, u12 = -(d1 * d2);
          m11[q] = __builtin_fma(sg, u11, m11[q]);
          m12[q] = __builtin_fma(sg, u12, m12[q]);
          m22[q] = __builtin_fma(sg, u22, m22[q]);
          n11 = __builtin_fma(sa, u11, n11);
          n12 = __builtin_fma(sa, u12, n12);
          n22 = __builtin_fma(sa, u22, n22);
        }
        if constexpr (NEAR) {  // near pairs w_source / |d| > theta, both directions
          if (qd < near.qmax && qd >= 1e-24) {
            const int ma = b * E + ia;
            if (gin[q] && near.th2 * qd < aw * aw) near_push(near, gm[q], ma);
            if (S.Lin[ia] && near.th2 * qd < gw[q] * gw[q]) near_push(near, ma, gm[q]);
          }
        }
      }
      b1 = rot1(b1);
      b2 = rot1(b2);
      if constexpr (METRIC) {
        n11 = rot1(n11);
        n12 = rot1(n12);
        n22 = rot1(n22);
      }
    };
    if constexpr (NEAR) {  // runs once per step of the solver: keep the code small
#pragma unroll 2
      for (int k = 0; k < 64; ++k) step(k);
    } else {
#pragma unroll
      for (int k = 0; k < 64; ++k) step(k);
    }
    // after 64 rotations lane i holds source base + i again
    Lacc[0][base + lane] = b1;
    Lacc[1][base + lane] = b2;
    if constexpr (METRIC) {
      Lacc[2][base + lane] = n11;
      Lacc[3][base + lane] = n12;
      Lacc[4][base + lane] = n22;
    }
    __syncthreads();
  }
  };
  if constexpr (METRIC || NEAR) {
    tiles(std::true_type{});
  } else {
    tiles(std::false_type{});
    // The vote: one word per wave in the tiles, which no wave reads after the
    // last phase's barrier.  High word of 1e-24 (0x3AF357C2 299A88EA): every
    // |d|^2 < 1e-24 has a high word <= it.
    int* vote = reinterpret_cast<int*>(&L[0]);
    const bool near_pair = __builtin_amdgcn_ballot_w64(qhi_min <= 0x3AF357C2) != 0;
    if (lane == 0) vote[r] = near_pair;
    __syncthreads();
    bool redo = false;
#pragma unroll
    for (int i = 0; i < NW; ++i) redo = redo || vote[i] != 0;
    if (redo) {  // block-uniform: every thread read the same words
      __syncthreads();
#pragma unroll
      for (int q = 0; q < Q; ++q) f1[q] = f2[q] = 0.0;
      stage();
      __syncthreads();
      tiles(std::true_type{});
    }
  }
  // targets of super-row a -> slot b; sources of super-row b -> slot a
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if (gm[q] >= N) continue;
    double* p = part + (size_t)b * cs + gm[q];
    p[0] = gin[q] ? f1[q] : 0.0;
    p[N] = gin[q] ? f2[q] : 0.0;
    if constexpr (METRIC) {
      p[2 * (size_t)N] = gin[q] ? m11[q] : 0.0;
      p[3 * (size_t)N] = gin[q] ? m12[q] : 0.0;
      p[4 * (size_t)N] = gin[q] ? m22[q] : 0.0;
    }
  }
  for (int e = threadIdx.x; e < E; e += 64 * NW) {
    const int m = b * E + e;
    if (m >= N) continue;
    const bool in = src[m].in_domain != 0.0;
    double* p = part + (size_t)a * cs + m;
#pragma unroll
    for (int c = 0; c < C; ++c) p[(size_t)c * N] = in ? Lacc[c][e] : 0.0;
  }
}

template <int R, bool METRIC, bool NEAR, bool POW2W = false>
__global__ __launch_bounds__(64 * R) void k_pair_sym_f64(const Src64* __restrict__ src, int N, long long n_off,
                                                         double* __restrict__ part, NearOut near) {
  __shared__ SymLds64<R, METRIC, NEAR, POW2W> S;
  if (near.live && *near.live == 0) return;
  int a, b;
  block_rows(blockIdx.x, n_off, a, b);
  sym_block_f64<R, METRIC, NEAR, POW2W>(src, N, a, b, part, S, near);
}

// FP32 terms, compensated sums: every 32 steps each accumulator's block sum is
// folded into a (hi, lo) pair with TwoSum (the targets' in VGPRs, the sources'
// in LDS), as in k_pair_f32c.
template <int R, bool METRIC, bool NEAR>
struct SymLds32 {
  float L[5][128 * R];  // sources of super-row b: v1, v2, G1, G2, w_masked
  float Lh[METRIC ? 5 : 2][128 * R], Ll[METRIC ? 5 : 2][128 * R];  // their (hi, lo) sums
  int Lin[NEAR ? 128 * R : 1];
};

// POW2W: the equal power-of-two weight form of sym_block_f64, in FP32 (2^k is
// exact there too; far points at +-kFar32); the metric sums take the weight
// once, at the end.
constexpr float kFar32 = 1e18f;

template <int R, bool METRIC, bool NEAR, bool POW2W = false>
__device__ __forceinline__ void sym_block_f32c(const Src32* __restrict__ src, int N, int a, int b,
                                               double* __restrict__ part, SymLds32<R, METRIC, NEAR>& S,
                                               const NearOut& near) {
  static_assert(!(POW2W && NEAR), "the equal-weight form has no near-pair search");
  constexpr int E = 128 * R, NTL = 2 * R, C = METRIC ? 5 : 2;
  auto& L = S.L;
  auto& Lh = S.Lh;
  auto& Ll = S.Ll;
  auto& Lin = S.Lin;
  const int lane = threadIdx.x & 63, r = threadIdx.x >> 6;
  const size_t cs = (size_t)C * N;
  float gv1[2], gv2[2], gg1[2], gg2[2], gw[2];
  float bs[2][C], hs[2][C], ls[2][C];  // target block sums and (hi, lo)
  bool gin[2];
  int gm[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    gm[q] = a * E + r * 128 + q * 64 + lane;
    gv1[q] = gv2[q] = gg1[q] = gg2[q] = gw[q] = 0.0f;
    gin[q] = false;
    if (gm[q] < N) {
      const Src32 s = src[gm[q]];
      gv1[q] = s.v1;
      gv2[q] = s.v2;
      gg1[q] = s.g1;
      gg2[q] = s.g2;
      gw[q] = s.w;
      gin[q] = s.in_domain != 0.0f;
    }
#pragma unroll
    for (int c = 0; c < C; ++c) bs[q][c] = hs[q][c] = ls[q][c] = 0.0f;
  }
  if (a == b) {  // diagonal block: ordered sum over the super-row, blocks of 32
    const int s0 = a * E, s1 = min(N, s0 + E);
    for (int sb = s0; sb < s1; sb += 32) {
      const int se = min(s1, sb + 32);
#pragma unroll 4
      for (int s = sb; s < se; ++s) {
        const float av1 = src[s].v1, av2 = src[s].v2, ag1 = src[s].g1, ag2 = src[s].g2, aw = src[s].w;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          if constexpr (METRIC) {
            float d1, d2, t;
            pair_term_f32_t(gv1[q], gv2[q], gg1[q], gg2[q], av1, av2, ag1, ag2, aw, bs[q][0], bs[q][1], d1, d2, t);
            const float td2 = t * d2, td1 = t * d1;
            bs[q][2] = __builtin_fmaf(td2, d2, bs[q][2]);
            bs[q][3] = __builtin_fmaf(-td2, d1, bs[q][3]);
            bs[q][4] = __builtin_fmaf(td1, d1, bs[q][4]);
            if constexpr (NEAR) {
              const float qd = __builtin_fmaf(d2, d2, d1 * d1);
              if (qd < (float)near.qmax && gin[q] && qd >= 1e-12f && (float)near.th2 * qd < aw * aw)
                near_push(near, gm[q], s);
            }
          } else {
            pair_term_f32(gv1[q], gv2[q], gg1[q], gg2[q], av1, av2, ag1, ag2, aw, bs[q][0], bs[q][1]);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int c = 0; c < C; ++c) {
          two_sum_fold(hs[q][c], ls[q][c], bs[q][c]);
          bs[q][c] = 0.0f;
        }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (gm[q] >= N) continue;
      double* p = part + (size_t)a * cs + gm[q];
#pragma unroll
      for (int c = 0; c < C; ++c) p[(size_t)c * N] = gin[q] ? (double)hs[q][c] + (double)ls[q][c] : 0.0;
    }
    return;
  }
  if constexpr (POW2W) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      gv1[q] = gin[q] ? gv1[q] : kFar32;
      gv2[q] = gin[q] ? gv2[q] : kFar32;
      gg1[q] = gin[q] ? gw[q] * gg1[q] : 0.0f;
      gg2[q] = gin[q] ? gw[q] * gg2[q] : 0.0f;
    }
  }
  for (int e = threadIdx.x; e < E; e += 64 * R) {
    const int m = b * E + e;
    float x1 = 0.0f, x2 = 0.0f, y1 = 0.0f, y2 = 0.0f, w = 0.0f;
    int in = 0;
    if (m < N) {
      const Src32 s = src[m];
      x1 = s.v1;
      x2 = s.v2;
      y1 = s.g1;
      y2 = s.g2;
      w = s.w;
      in = s.in_domain != 0.0f;
    }
    if constexpr (POW2W) {  // padding past N too: the weight no longer masks it
      x1 = in ? x1 : -kFar32;
      x2 = in ? x2 : -kFar32;
      y1 = in ? w * y1 : 0.0f;
      y2 = in ? w * y2 : 0.0f;
    }
    L[0][e] = x1;
    L[1][e] = x2;
    L[2][e] = y1;
    L[3][e] = y2;
    L[4][e] = w;
    if constexpr (NEAR) Lin[e] = in;
#pragma unroll
    for (int c = 0; c < C; ++c) Lh[c][e] = Ll[c][e] = 0.0f;
  }
  __syncthreads();
  // GUARD = false: unguarded pairs with the smallest |d|^2 kept per lane (its
  // bits: |d|^2 >= 0), rerun guarded if any came below 1e-12, as in the FP64 block
  int qbits_min = 0x7FFFFFFF;
  auto tiles = [&](auto guard_tag) {
  constexpr bool GUARD = decltype(guard_tag)::value;
  for (int p = 0; p < NTL; ++p) {
    const int base = ((2 * r + p) % NTL) * 64;
    for (int h = 0; h < 2; ++h) {
      float ab[C];  // source block sums (travel with the sources)
#pragma unroll
      for (int c = 0; c < C; ++c) ab[c] = 0.0f;
#pragma unroll 4
      for (int kk = 0; kk < 32; ++kk) {
        const int ia = base + ((lane + 32 * h + kk) & 63);
        const float av1 = L[0][ia], av2 = L[1][ia], ag1 = L[2][ia], ag2 = L[3][ia], aw = L[4][ia];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const float d1 = gv1[q] - av1;
          const float d2 = gv2[q] - av2;
          const float u22 = d1 * d1;
          const float qd = __builtin_fmaf(d2, d2, u22);
          float y;
          if constexpr (GUARD) {
            y = __builtin_amdgcn_rsqf(guard_q32(qd));
          } else {
            qbits_min = min(qbits_min, __float_as_int(qd));
            y = __builtin_amdgcn_rsqf(qd);
          }
          const float y3 = (y * y) * y;
          const float c = __builtin_fmaf(ag1 - gg1[q], d2, -((ag2 - gg2[q]) * d1));
          if constexpr (POW2W && !METRIC) {  // c is w c: one factor serves both sides
            const float pc = y3 * c;
            bs[q][0] = __builtin_fmaf(pc, d2, bs[q][0]);
            bs[q][1] = __builtin_fmaf(-pc, d1, bs[q][1]);
            ab[0] = __builtin_fmaf(-pc, d2, ab[0]);
            ab[1] = __builtin_fmaf(pc, d1, ab[1]);
          } else if constexpr (!METRIC) {
            const float pc = y3 * c;
            const float kg = aw * pc, ka = gw[q] * pc;
            bs[q][0] = __builtin_fmaf(kg, d2, bs[q][0]);
            bs[q][1] = __builtin_fmaf(-kg, d1, bs[q][1]);
            ab[0] = __builtin_fmaf(-ka, d2, ab[0]);
            ab[1] = __builtin_fmaf(ka, d1, ab[1]);
          } else {
            // POW2W: kg = ka = y^3 (w c), the same bits as (w y^3) c; the metric
            // sums run without the weight (sg = sa = y^3) and take it at the end
            const float sg = POW2W ? y3 : aw * y3, sa = POW2W ? y3 : gw[q] * y3;
            const float kg = POW2W ? y3 * c : sg * c, ka = POW2W ? kg : sa * c;
            bs[q][0] = __builtin_fmaf(kg, d2, bs[q][0]);
            bs[q][1] = __builtin_fmaf(-kg, d1, bs[q][1]);
            ab[0] = __builtin_fmaf(-ka, d2, ab[0]);
            ab[1] = __builtin_fmaf(ka, d1, ab[1]);
            const float u11 = d2 * d2, u12 = -(d1 * d2);
            bs[q][2] = __builtin_fmaf(sg, u11, bs[q][2]);
            bs[q][3] = __builtin_fmaf(sg, u12, bs[q][3]);
            bs[q][4] = __builtin_fmaf(sg, u22, bs[q][4]);
            ab[2] = __builtin_fmaf(sa, u11, ab[2]);
            ab[3] = __builtin_fmaf(sa, u12, ab[3]);
            ab[4] = __builtin_fmaf(sa, u22, ab[4]);
          }
          if constexpr (NEAR) {
            if (qd < (float)near.qmax && qd >= 1e-12f) {
              const int ma = b * E + ia;
              const float th2 = (float)near.th2;
              if (gin[q] && th2 * qd < aw * aw) near_push(near, gm[q], ma);
              if (Lin[ia] && th2 * qd < gw[q] * gw[q]) near_push(near, ma, gm[q]);
            }
          }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) ab[c] = rot1(ab[c]);
      }
      // lane i now holds source base + (i + 32 (h + 1)) mod 64
      const int ie = base + ((lane + 32 * (h + 1)) & 63);
#pragma unroll
      for (int c = 0; c < C; ++c) {
        float hi = Lh[c][ie], lo = Ll[c][ie];
        two_sum_fold(hi, lo, ab[c]);
        Lh[c][ie] = hi;
        Ll[c][ie] = lo;
      }
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int c = 0; c < C; ++c) {
          two_sum_fold(hs[q][c], ls[q][c], bs[q][c]);
          bs[q][c] = 0.0f;
        }
    }
    __syncthreads();
  }
  };
  if constexpr (NEAR) {
    tiles(std::true_type{});
  } else {
    tiles(std::false_type{});
    if (__syncthreads_or(qbits_min < 0x2B8CBCCC)) {  // bits of 1e-12f
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int c = 0; c < C; ++c) hs[q][c] = ls[q][c] = 0.0f;
      for (int e = threadIdx.x; e < E; e += 64 * R)
#pragma unroll
        for (int c = 0; c < C; ++c) Lh[c][e] = Ll[c][e] = 0.0f;
      __syncthreads();
      tiles(std::true_type{});
    }
  }
  // with equal weights the metric sums take their weight w = 2^k here (exact:
  // every product and TwoSum of the loop scaled by w gives the general form's bits)
  constexpr int CW = POW2W && METRIC ? 2 : C;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    if (gm[q] >= N) continue;
    double* p = part + (size_t)b * cs + gm[q];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      double x = (double)hs[q][c] + (double)ls[q][c];
      if (c >= CW) x *= (double)gw[q];
      p[(size_t)c * N] = gin[q] ? x : 0.0;
    }
  }
  for (int e = threadIdx.x; e < E; e += 64 * R) {
    const int m = b * E + e;
    if (m >= N) continue;
    const Src32 s = src[m];
    double* p = part + (size_t)a * cs + m;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      double x = (double)Lh[c][e] + (double)Ll[c][e];
      if (c >= CW) x *= (double)s.w;
      p[(size_t)c * N] = s.in_domain != 0.0f ? x : 0.0;
    }
  }
}

template <int R, bool METRIC, bool NEAR, bool POW2W = false>
__global__ __launch_bounds__(64 * R) void k_pair_sym_f32c(const Src32* __restrict__ src, int N, long long n_off,
                                                          double* __restrict__ part, NearOut near) {
  __shared__ SymLds32<R, METRIC, NEAR> S;
  if (near.live && *near.live == 0) return;
  int a, b;
  block_rows(blockIdx.x, n_off, a, b);
  sym_block_f32c<R, METRIC, NEAR, POW2W>(src, N, a, b, part, S, near);
}

// Batched ensemble: a grid-stride walk over the blocks (run, a, b) of every
// run's symmetric sum, FP64 or FP32c per run; each block runs the single-run
// block body on that run's records and partial slots, so a run's partials
// equal those of k_pair_sym_<R> over its markers alone.
template <int R>
__global__ __launch_bounds__(64 * R) void k_pair_sym_b(EnsView E, const int4* __restrict__ blocks, int nblocks,
                                                       const Src64* __restrict__ s64, const Src32* __restrict__ s32,
                                                       int force_f64, double* __restrict__ part) {
  constexpr size_t b64 = sizeof(SymLds64<R, false, false>), b32 = sizeof(SymLds32<R, false, false>);
  __shared__ __align__(16) unsigned char smem[b64 > b32 ? b64 : b32];
  const NearOut nz{};
  for (int t = blockIdx.x; t < nblocks; t += gridDim.x) {
    const int4 bl = blocks[t];
    const int r = bl.x;
    if (!force_f64 && !E.ctl[r].active) continue;
    const bool f64 = force_f64 || E.ctl[r].kind == 0;
    const EnsRun& Rn = E.runs[r];
    if (f64)
      sym_block_f64<R, false, false>(s64 + Rn.off, Rn.n, bl.y, bl.z, part + Rn.poff,
                                     *reinterpret_cast<SymLds64<R, false, false>*>(smem), nz);
    else
      sym_block_f32c<R, false, false>(s32 + Rn.off, Rn.n, bl.y, bl.z, part + Rn.poff,
                                      *reinterpret_cast<SymLds32<R, false, false>*>(smem), nz);
    __syncthreads();  // the next block reuses the LDS
  }
}

template <int R>
void sym_f64(const Src64* src, int N, bool metric, const NearOut* near, double* part, hipStream_t st,
             const int* live, bool pow2w) {
  const long long S = sym_rows(N, R), n_off = S * (S - 1) / 2;
  const dim3 grid((unsigned)(n_off + S)), block(64 * R);
  NearOut nz{};
  nz.live = live;
  if (!metric && pow2w)
    hipLaunchKernelGGL((k_pair_sym_f64<R, false, false, true>), grid, block, 0, st, src, N, n_off, part, nz);
  else if (!metric)
    hipLaunchKernelGGL((k_pair_sym_f64<R, false, false>), grid, block, 0, st, src, N, n_off, part, nz);
  else if (!near)
    hipLaunchKernelGGL((k_pair_sym_f64<R, true, false>), grid, block, 0, st, src, N, n_off, part, nz);
  else
  {
    NearOut nn = *near;
    nn.live = live;
    hipLaunchKernelGGL((k_pair_sym_f64<R, true, true>), grid, block, 0, st, src, N, n_off, part, nn);
  }
}

template <int R>
void sym_f32c(const Src32* src, int N, bool metric, const NearOut* near, double* part, hipStream_t st,
             const int* live, bool pow2w) {
  const long long S = sym_rows(N, R), n_off = S * (S - 1) / 2;
  const dim3 grid((unsigned)(n_off + S)), block(64 * R);
  NearOut nz{};
  nz.live = live;
  if (!metric && pow2w)
    hipLaunchKernelGGL((k_pair_sym_f32c<R, false, false, true>), grid, block, 0, st, src, N, n_off, part, nz);
  else if (!metric)
    hipLaunchKernelGGL((k_pair_sym_f32c<R, false, false>), grid, block, 0, st, src, N, n_off, part, nz);
  else if (!near && pow2w)
    hipLaunchKernelGGL((k_pair_sym_f32c<R, true, false, true>), grid, block, 0, st, src, N, n_off, part, nz);
  else if (!near)
    hipLaunchKernelGGL((k_pair_sym_f32c<R, true, false>), grid, block, 0, st, src, N, n_off, part, nz);
  else
  {
    NearOut nn = *near;
    nn.live = live;
    hipLaunchKernelGGL((k_pair_sym_f32c<R, true, true>), grid, block, 0, st, src, N, n_off, part, nn);
  }
}

}  // namespace

void launch_pair_sym_f64(const Src64* src, int N, int R, bool metric, const NearOut* near, double* part,
                         hipStream_t st, const int* live, bool pow2w) {
  if (N <= 0) return;
  switch (R) {
    case 2: sym_f64<2>(src, N, metric, near, part, st, live, pow2w); break;
    case 4: sym_f64<4>(src, N, metric, near, part, st, live, pow2w); break;
    default: sym_f64<8>(src, N, metric, near, part, st, live, pow2w); break;
  }
}

void launch_pair_sym_b(const EnsView& E, const EnsTiles& T, const Src64* s64, const Src32* s32, int force_f64,
                       double* part, hipStream_t st) {
  if (T.ntiles <= 0) return;
  const dim3 grid(T.grid);
  switch (T.sym) {
    case 2:
      hipLaunchKernelGGL(k_pair_sym_b<2>, grid, dim3(128), 0, st, E, T.tiles, T.ntiles, s64, s32, force_f64, part);
      break;
    case 8:
      hipLaunchKernelGGL(k_pair_sym_b<8>, grid, dim3(512), 0, st, E, T.tiles, T.ntiles, s64, s32, force_f64, part);
      break;
    default:
      hipLaunchKernelGGL(k_pair_sym_b<4>, grid, dim3(256), 0, st, E, T.tiles, T.ntiles, s64, s32, force_f64, part);
      break;
  }
}

void launch_pair_sym_f32c(const Src32* src, int N, int R, bool metric, const NearOut* near, double* part,
                          hipStream_t st, const int* live, bool pow2w) {
  if (N <= 0) return;
  switch (R) {
    case 2: sym_f32c<2>(src, N, metric, near, part, st, live, pow2w); break;
    case 4: sym_f32c<4>(src, N, metric, near, part, st, live, pow2w); break;
    default: sym_f32c<8>(src, N, metric, near, part, st, live, pow2w); break;
  }
}

}  // namespace collop
