// capi.hip -- the C-ABI of include/collop.h: context, mesh set-up, hook-level
// entry points and the device-resident implicit step.
//
// Host-side responsibilities only: argument checking, device buffers cached in
// the context (grow-only), the mesh tables derived once per mesh (1D mass
// matrices and their banded Cholesky factors, basis values at the Gauss points),
// and the Anderson state machine of solver.jl:300-423, which reads a handful of
// reduced scalars per iteration while every O(N) and O(N^2) pass runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <utility>
#include <vector>

#include "../../include/collop.h"
#include "anderson.hpp"
#include "collop_ctx.hpp"
#include "nk.hpp"

using namespace collop;

// ---- scratch of the allocation-free pair sum (collop_collision_f64_noalloc) ---
// The packed source records (8 doubles per marker) followed by kNoAllocSlices
// slices of 2N partial sums, for N <= kNoAllocMaxN: 21 MB, one copy per GPU,
// held until the process ends.  Its device address on each GPU is looked up
// once (collop_create) and only that address is used from the host.
constexpr int64_t kNoAllocMaxN = 65536;
constexpr int kNoAllocSlices = 16;
static __device__ __align__(256) double scratch[(size_t)kNoAllocMaxN * (8 + 2 * kNoAllocSlices)];
constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices] = {nullptr};

namespace {

int nk_setup(collop_ctx* ctx, int n, int mem, NkDevice& E, double** extra, int nextra);

// Gauss-Legendre rule on [0,1]: Newton's method on the three-term recurrence
// of P_n (any textbook; e.g. Press et al., Numerical Recipes, 4.6).
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double t = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = t;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0;
      dp = n * (t * p1 - p0) / (t * t - 1.0);
      const double d = p1 / dp;
      t -= d;
      if (std::fabs(d) < 1e-16) break;
    }
    double p0 = 1.0, p1 = t;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    if (n == 1) p0 = 1.0;
    dp = n * (t * p1 - p0) / (t * t - 1.0);
    x[n - 1 - i] = 0.5 * (1.0 + t);
    w[n - 1 - i] = 1.0 / ((1.0 - t * t) * dp * dp);
  }
}

// phi_j(xi) on one cell from the reference's extraction layout C[i + 3j + 9 cell]
void cell_basis(const double* C, int cell, double xi, double phi[3]) {
  const double om = 1.0 - xi;
  const double B[3] = {om * om, 2.0 * xi * om, xi * xi};
  for (int j = 0; j < 3; ++j)
    phi[j] = C[9 * cell + 3 * j] * B[0] + C[9 * cell + 3 * j + 1] * B[1] + C[9 * cell + 3 * j + 2] * B[2];
}

// 1D mass matrix of one direction, exact with a 3-point Gauss rule per cell,
// and its banded Cholesky factor stored as {1/L_ii, L_i,i-1, L_i,i-2}.
// Dense inverse of a 1D mass matrix from its banded Cholesky factor
// (chol = {1/L_ii, L_i,i-1, L_i,i-2}): column j solves M x = e_j; the result is
// symmetrised.  The 1D mass matrices are well conditioned, so applying the
// inverse costs no accuracy against the banded solve.
void mass_inverse_1d(const std::vector<double>& chol, int nd, std::vector<double>& inv) {
  inv.assign((size_t)nd * nd, 0.0);
  std::vector<double> x(nd);
  for (int j = 0; j < nd; ++j) {
    std::fill(x.begin(), x.end(), 0.0);
    x[j] = 1.0;
    for (int i = 0; i < nd; ++i) {  // L y = e_j
      double t = x[i];
      if (i >= 1) t -= chol[3 * i + 1] * x[i - 1];
      if (i >= 2) t -= chol[3 * i + 2] * x[i - 2];
      x[i] = t * chol[3 * i];
    }
    for (int i = nd - 1; i >= 0; --i) {  // L^T x = y
      double t = x[i];
      if (i + 1 < nd) t -= chol[3 * (i + 1) + 1] * x[i + 1];
      if (i + 2 < nd) t -= chol[3 * (i + 2) + 2] * x[i + 2];
      x[i] = t * chol[3 * i];
    }
    for (int i = 0; i < nd; ++i) inv[(size_t)i * nd + j] = x[i];
  }
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < i; ++j) {
      const double a = 0.5 * (inv[(size_t)i * nd + j] + inv[(size_t)j * nd + i]);
      inv[(size_t)i * nd + j] = inv[(size_t)j * nd + i] = a;
    }
}

bool mass_factor_1d(const std::vector<double>& bp, const double* C, const int* bs0, int nd, std::vector<double>& out) {
  const int ne = (int)bp.size() - 1;
  std::vector<double> gx, gw;
  gauss_legendre(3, gx, gw);
  std::vector<double> M((size_t)nd * 3, 0.0);  // M[i][k] = M(i, i-k), k = 0..2
  for (int e = 0; e < ne; ++e) {
    const double h = bp[e + 1] - bp[e];
    for (int q = 0; q < 3; ++q) {
      double phi[3];
      cell_basis(C, e, gx[q], phi);
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b <= a; ++b) {
          const int i = bs0[e] + a, j = bs0[e] + b;  // i >= j
          M[(size_t)i * 3 + (i - j)] += phi[a] * phi[b] * gw[q] * h;
        }
    }
  }
  out.assign((size_t)nd * 3, 0.0);
  std::vector<double> L((size_t)nd * 3, 0.0);  // L[i][k] = L(i, i-k)
  for (int j = 0; j < nd; ++j) {
    double s = M[(size_t)j * 3];
    for (int k = 1; k <= 2 && j - k >= 0; ++k) s -= L[(size_t)j * 3 + k] * L[(size_t)j * 3 + k];
    if (!(s > 0.0)) return false;
    const double ljj = std::sqrt(s);
    L[(size_t)j * 3] = ljj;
    for (int i = j + 1; i <= std::min(nd - 1, j + 2); ++i) {
      double t = M[(size_t)i * 3 + (i - j)];
      // sum_k L(i,k) L(j,k) over k < j within the band
      for (int k = std::max(0, i - 2); k < j; ++k) t -= L[(size_t)i * 3 + (i - k)] * L[(size_t)j * 3 + (j - k)];
      L[(size_t)i * 3 + (i - j)] = t / ljj;
    }
  }
  for (int i = 0; i < nd; ++i) {
    out[(size_t)i * 3] = 1.0 / L[(size_t)i * 3];
    out[(size_t)i * 3 + 1] = i >= 1 ? L[(size_t)i * 3 + 1] : 0.0;
    out[(size_t)i * 3 + 2] = i >= 2 ? L[(size_t)i * 3 + 2] : 0.0;
  }
  return true;
}

int upload(collop_ctx* ctx, const void* host, size_t bytes, const void** dev_out) {
  auto* b = new DBuf();
  ctx->mesh_bufs.push_back(b);
  int r = ensure(ctx, *b, bytes);
  if (r != COLLOP_OK) return r;
  HIPC(hipMemcpy(b->p, host, bytes, hipMemcpyHostToDevice));
  *dev_out = b->p;
  return COLLOP_OK;
}

void free_mesh(collop_ctx* ctx) {
  for (DBuf* b : ctx->mesh_bufs) {
    release(*b);
    delete b;
  }
  ctx->mesh_bufs.clear();
  ctx->has_mesh = false;
}

int need_mesh(collop_ctx* ctx) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!ctx->has_mesh) return fail(ctx, COLLOP_ERR_NO_MESH, "collop_set_mesh has not been called on this context");
  return COLLOP_OK;
}

int check_N(collop_ctx* ctx, int64_t N) {
  if (N < 0 || N > (int64_t)1 << 27) return fail(ctx, COLLOP_ERR_INVALID, "N must be in [0, 2^27]");
  return COLLOP_OK;
}

// Symmetric pair sum: R waves per block.  Automatic: 2 below N = 4096, else 4
// (measured fastest at N = 65536: 8256 blocks of 256 threads), or 8 where 4
// would need more than kSymMaxSlots partial slots.  0 (the ordered sum) when
// it is switched off or N needs more slots than that.
int sym_tiles_for(const collop_ctx* ctx, int64_t N) {
  if (!ctx->pair_sym || N <= 0) return 0;
  int R = ctx->sym_tiles;
  if (!R) R = N < 4096 ? 2 : (sym_rows((int)N, 4) <= kSymPreferSlots ? 4 : 8);
  return sym_rows((int)N, R) <= kSymMaxSlots ? R : 0;
}

// The tile a pair sum over N markers runs with, from the context's settings.
PairTile tile_for(const collop_ctx* ctx, const PairTile& cur, int N) {
  PairTile t = cur;
  t.sym = sym_tiles_for(ctx, N);
  if (t.sym) {
    t.slices = sym_rows(N, t.sym);
  } else if (!ctx->tile_auto) {
    t.slices = ctx->manual_slices;
  } else {
    // Enough workgroups (>= 8192, i.e. ~32 waves per SIMD on 256 CUs, measured
    // best at N = 65536) by cutting the source range into slices; multiples of 8
    // keep one slice set per XCD.
    const int per = 256 * t.gamma_per_lane;
    const int gb = std::max(1, (N + per - 1) / per);
    int s = (8192 + gb - 1) / gb;
    if (s >= 8) s = (s + 7) / 8 * 8;
    s = std::max(1, std::min(s, 64));
    while (s > 1 && (N + s - 1) / s < 16) s /= 2;  // keep slices non-trivial
    t.slices = s;
  }
  return t;
}

void choose_tile(collop_ctx* ctx, int N) {
  ctx->tile64 = tile_for(ctx, ctx->tile64, N);
  ctx->tile32 = tile_for(ctx, ctx->tile32, N);
}

// ---- profiling of the pair kernel (HIP events on the launch stream) -------
void prof_begin(collop_ctx* ctx, int kind) {
  if (!ctx->profile) return;
  if (ctx->ev_used == ctx->ev_pool.size()) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    ctx->ev_pool.push_back({a, b});
    ctx->ev_kind.push_back(0);
    ctx->ev_map.push_back(0);
  }
  ctx->ev_kind[ctx->ev_used] = kind;
  ctx->ev_map[ctx->ev_used] = ctx->prof_map;
  (void)hipEventRecord(ctx->ev_pool[ctx->ev_used].first, ctx->stream);
}

void prof_end(collop_ctx* ctx) {
  if (!ctx->profile) return;
  (void)hipEventRecord(ctx->ev_pool[ctx->ev_used].second, ctx->stream);
  ctx->ev_used++;
}

// Called after a stream synchronisation.  Launches of maps after map_cut of a
// device-resident solve (launched, then skipped once the solve had ended) are
// not counted.
void prof_collect(collop_ctx* ctx, int map_cut = 1 << 30) {
  if (!ctx->profile) return;
  for (size_t i = 0; i < ctx->ev_used; ++i) {
    if (ctx->ev_map[i] > map_cut) continue;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->ev_pool[i].first, ctx->ev_pool[i].second) == hipSuccess) {
      ctx->prof_ms[ctx->ev_kind[i]] += ms;
      ctx->prof_n[ctx->ev_kind[i]]++;
    }
  }
  ctx->ev_used = 0;
}

// Waits until the pinned sequence number reaches `target` (published by the
// device after its writes to pinned memory, which is coherent, so they reach
// the host without a synchronisation): the host polls instead of synchronising
// the stream, a round trip about 5 us shorter.  After 20 ms (long maps) it
// synchronises the stream; stat_syncs counts the polls that ran out.
int wait_seq(collop_ctx* ctx, unsigned target) {
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned it = 1;; ++it) {
    if ((int)(__atomic_load_n(ctx->h_seq, __ATOMIC_ACQUIRE) - target) >= 0) return COLLOP_OK;
    if ((it & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) break;
  }
  ctx->stat_syncs++;
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

// The map's Anderson statistics: one kernel reduces the epilogue partials into
// the pinned h_stat and then publishes a new sequence number (wait_seq).
int publish_stat(collop_ctx* ctx, int nparts) {
  launch_stat_publish(P<double>(ctx->b_statp), nparts, ctx->h_stat_dev, ctx->h_seq_dev, ++ctx->seq, ctx->stream);
  CHECK_LAUNCH();
  int r;
  if ((r = wait_seq(ctx, ctx->seq))) return r;
  if (__atomic_load_n(ctx->h_seq, __ATOMIC_ACQUIRE) != ctx->seq)
    return fail(ctx, COLLOP_ERR_HIP, "the statistics' sequence number was not published");
  return COLLOP_OK;
}

// ---- building blocks on device buffers -------------------------------------
// part <- pair sum of the packed sources (FP64: kind 0; FP32c: kind 1)
// True when all N weights are one power of two in [2^-60, 2^60]: the pair sums'
// equal-weight form then gives the same bits (scaling by 2^k is exact in FP32
// and FP64 over that range).
bool equal_pow2_weights(const double* w, int64_t N) {
  int e = 0;
  if (N <= 0 || std::frexp(w[0], &e) != 0.5 || e < -59 || e > 61) return false;
  for (int64_t i = 1; i < N; ++i)
    if (w[i] != w[0]) return false;
  return true;
}

// pow2w: the markers packed in b_s64/b_s32 satisfy equal_pow2_weights
int run_pair(collop_ctx* ctx, int kind, int N, double* part, bool pow2w, const int* live = nullptr) {
  pow2w = pow2w && ctx->pow2w_path;
  if (kind == 0) {
    prof_begin(ctx, 0);
    launch_pair_f64(P<Src64>(ctx->b_s64), N, ctx->tile64, part, ctx->stream, live, pow2w);
  } else {
    prof_begin(ctx, 1);
    launch_pair_f32c(P<Src32>(ctx->b_s32), N, ctx->tile32, part, ctx->stream, live, pow2w);
  }
  CHECK_LAUNCH();
  prof_end(ctx);
  return COLLOP_OK;
}

int ensure_scratch(collop_ctx* ctx, int64_t N) {
  const size_t n2 = (size_t)2 * std::max<int64_t>(N, 1);
  const MeshView& m = ctx->mv;
  ENSURE(ctx->b_v, n2 * 8);
  ENSURE(ctx->b_w, n2 / 2 * 8);
  ENSURE(ctx->b_G, n2 * 8);
  ENSURE(ctx->b_out, (n2 / 2) * 5 * 8);
  ENSURE(ctx->b_vmid, n2 * 8);
  ENSURE(ctx->b_dv, n2 * 8);
  ENSURE(ctx->b_dS, n2 * 8);
  ENSURE(ctx->b_s64, (n2 / 2) * sizeof(Src64));
  ENSURE(ctx->b_s32, (n2 / 2) * sizeof(Src32));
  const int Rs = sym_tiles_for(ctx, N);
  const int Smax = std::max(64, Rs ? sym_rows((int)N, Rs) : 0);
  ENSURE(ctx->b_part, (size_t)Smax * (n2 / 2) * 5 * 8);
  const int nsb = scatter_blocks((int)N);
  ENSURE(ctx->b_spart_sc, (size_t)2 * nsb * m.n_dofs * 8);
  ENSURE(ctx->b_mom, (size_t)nsb * 3 * 8);
  ENSURE(ctx->b_dpart, (size_t)gather_blocks((int)N) * 2 * 8);
  ENSURE(ctx->b_iota, (size_t)m.ne1 * m.ne2 * m.nq * m.nq * 8);
  ENSURE(ctx->b_spart_q, (size_t)2 * quad_blocks(m) * 8);
  ENSURE(ctx->b_vec, (size_t)4 * m.n_dofs * 8);  // [f_mid, f_in, r, L]
  ENSURE(ctx->b_scal, 16 * 8);                    // S0, S1, lambda, E, P1, P2, ...
  ENSURE(ctx->b_statp, (size_t)epi_blocks((int)N) * kStatWidth * 8);
  ENSURE(ctx->b_stmp, (size_t)2 * m.n_dofs * 8);
  ENSURE(ctx->b_lbpart, (size_t)lb_blocks((int)N) * kLbSums * 8);
  return COLLOP_OK;
}

}  // namespace

// ===========================================================================
extern "C" {

int collop_version(void) { return 100; }

int collop_create(int device, collop_ctx** out) {
  if (!out) return COLLOP_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return COLLOP_ERR_HIP;
  if (device < 0 || device >= n) return COLLOP_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return COLLOP_ERR_HIP;
  auto* c = new collop_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return COLLOP_ERR_HIP;
  }
  if (hipHostMalloc((void**)&c->h_stat, kStatWidth * sizeof(double), hipHostMallocCoherent) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return COLLOP_ERR_HIP;
  }
  if (hipHostGetDevicePointer((void**)&c->h_stat_dev, c->h_stat, 0) != hipSuccess ||
      hipHostMalloc((void**)&c->h_seq, 64, hipHostMallocCoherent) != hipSuccess) {
    (void)hipHostFree(c->h_stat);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return COLLOP_ERR_HIP;
  }
  *c->h_seq = 0;
  if (hipHostGetDevicePointer((void**)&c->h_seq_dev, c->h_seq, 0) != hipSuccess) {
    (void)hipHostFree(c->h_seq);
    (void)hipHostFree(c->h_stat);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return COLLOP_ERR_HIP;
  }
  // the allocation-free pair sum's scratch on this device (the current one)
  if (device < kMaxDevices && !g_scratch_dev[device]) {
    double* ptr = nullptr;
    if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) == hipSuccess && ptr) g_scratch_dev[device] = ptr;
  }
  *out = c;
  return COLLOP_OK;
}

int collop_destroy(collop_ctx* ctx) {
  if (!ctx) return COLLOP_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ensemble_free(ctx->ens);
  ctx->ens = nullptr;
  free_mesh(ctx);
  for (DBuf* b : {&ctx->b_v, &ctx->b_w, &ctx->b_G, &ctx->b_out, &ctx->b_vmid, &ctx->b_dv, &ctx->b_dS, &ctx->b_s64,
                  &ctx->b_s32, &ctx->b_part, &ctx->b_spart_sc, &ctx->b_mom, &ctx->b_dpart, &ctx->b_iota,
                  &ctx->b_spart_q, &ctx->b_vec, &ctx->b_scal, &ctx->b_statp, &ctx->b_stmp, &ctx->b_lbpart, &ctx->nk_V, &ctx->nk_vec,
                  &ctx->nk_part, &ctx->nk_h, &ctx->nk_A, &ctx->s_v, &ctx->s_w,
                  &ctx->s_v1, &ctx->s_Gv, &ctx->s_r, &ctx->s_rp, &ctx->s_Gp, &ctx->s_best, &ctx->s_dF, &ctx->s_dG,
                  &ctx->s_fc, &ctx->s_dotv, &ctx->s_Gvo, &ctx->s_Gk, &ctx->s_FA, &ctx->s_D, &ctx->s_Ge, &ctx->s_vmk, &ctx->s_xk, &ctx->s_oU,
                  &ctx->s_oUp, &ctx->s_of, &ctx->s_ofp, &ctx->s_odF, &ctx->s_odG, &ctx->b_near_list, &ctx->b_near_count,
                  &ctx->b_ntg, &ctx->b_noff, &ctx->b_nal, &ctx->b_nidx, &ctx->s_v1n, &ctx->m_params, &ctx->m_adam_m, &ctx->m_adam_v, &ctx->m_data, &ctx->m_rec,
                  &ctx->m_y, &ctx->m_gpart, &ctx->m_lpart, &ctx->m_tiles, &ctx->m_eval, &ctx->m_grad,
                  &ctx->b_cell_of, &ctx->b_cell_cnt, &ctx->b_cell_start, &ctx->b_cell_sorted, &ctx->b_links,
                  &ctx->b_link_count, &ctx->b_cl_off, &ctx->b_cl_idx, &ctx->b_cl_row, &ctx->b_cl_x,
                  &ctx->b_cl_root, &ctx->b_cl_have, &ctx->b_cl_fail, &ctx->b_cl_its, &ctx->b_actl})
    release(*b);
  for (auto& e : ctx->ev_pool) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  if (ctx->h_stat) (void)hipHostFree(ctx->h_stat);
  if (ctx->h_seq) (void)hipHostFree(ctx->h_seq);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return COLLOP_OK;
}

const char* collop_last_error(const collop_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int collop_set_mesh(collop_ctx* ctx, const double* bp1, int nb1, const double* bp2, int nb2, const double* C1,
                    const double* C2, const int32_t* bs1, const int32_t* bs2, int nd1, int nd2, int p_deg,
                    int n_quad) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (p_deg != 2)
    return fail(ctx, COLLOP_ERR_UNSUPPORTED, "the device kernels are specialised to P_DEG=2 (got " +
                                                 std::to_string(p_deg) + ")");
  if (!bp1 || !bp2 || !C1 || !C2 || !bs1 || !bs2) return fail(ctx, COLLOP_ERR_INVALID, "null mesh table");
  if (nb1 < 2 || nb2 < 2 || nb1 > 512 || nb2 > 512) return fail(ctx, COLLOP_ERR_INVALID, "breakpoint count out of range [2, 512]");
  if (n_quad < 1 || n_quad > 16) return fail(ctx, COLLOP_ERR_INVALID, "n_quad out of range [1, 16]");
  for (int i = 0; i + 1 < nb1; ++i)
    if (!(bp1[i] < bp1[i + 1])) return fail(ctx, COLLOP_ERR_INVALID, "bp1 must be strictly increasing");
  for (int i = 0; i + 1 < nb2; ++i)
    if (!(bp2[i] < bp2[i + 1])) return fail(ctx, COLLOP_ERR_INVALID, "bp2 must be strictly increasing");
  const int ne1 = nb1 - 1, ne2 = nb2 - 1;
  std::vector<int> b1(ne1), b2(ne2);
  for (int e = 0; e < ne1; ++e) b1[e] = bs1[e] - 1;
  for (int e = 0; e < ne2; ++e) b2[e] = bs2[e] - 1;
  for (int e = 0; e < ne1; ++e)
    if (b1[e] < 0 || b1[e] + 2 >= nd1) return fail(ctx, COLLOP_ERR_INVALID, "bs1 inconsistent with nd1");
  for (int e = 0; e < ne2; ++e)
    if (b2[e] < 0 || b2[e] + 2 >= nd2) return fail(ctx, COLLOP_ERR_INVALID, "bs2 inconsistent with nd2");
  const long n_dofs = (long)nd1 * nd2;
  if (n_dofs * 8 + 2 * 512 * 8 > 64 * 1024)
    return fail(ctx, COLLOP_ERR_UNSUPPORTED, "n_dofs too large for the LDS-privatised scatter (max 6144)");
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  free_mesh(ctx);
  ctx->has_state = false;
  ensemble_free(ctx->ens);  // its tables and captured graph refer to the old mesh
  ctx->ens = nullptr;

  std::vector<double> v1(bp1, bp1 + nb1), v2(bp2, bp2 + nb2);
  std::vector<double> ch1, ch2;
  if (!mass_factor_1d(v1, C1, b1.data(), nd1, ch1) || !mass_factor_1d(v2, C2, b2.data(), nd2, ch2))
    return fail(ctx, COLLOP_ERR_INVALID, "1D mass matrix is not positive definite (check the extraction tables)");
  std::vector<double> mi1, mi2;
  mass_inverse_1d(ch1, nd1, mi1);
  mass_inverse_1d(ch2, nd2, mi2);
  std::vector<double> qx, qw;
  gauss_legendre(n_quad, qx, qw);
  std::vector<double> ph1((size_t)ne1 * 3 * n_quad), ph2((size_t)ne2 * 3 * n_quad);
  for (int e = 0; e < ne1; ++e)
    for (int a = 0; a < n_quad; ++a) {
      double phi[3];
      cell_basis(C1, e, qx[a], phi);
      for (int j = 0; j < 3; ++j) ph1[(size_t)(3 * e + j) * n_quad + a] = phi[j];
    }
  for (int e = 0; e < ne2; ++e)
    for (int a = 0; a < n_quad; ++a) {
      double phi[3];
      cell_basis(C2, e, qx[a], phi);
      for (int j = 0; j < 3; ++j) ph2[(size_t)(3 * e + j) * n_quad + a] = phi[j];
    }
  auto erange = [](const std::vector<int>& bs, int nd, std::vector<int>& out) {
    out.assign((size_t)2 * nd, 0);
    for (int i = 0; i < nd; ++i) {
      out[2 * i] = std::numeric_limits<int>::max();
      out[2 * i + 1] = -1;
    }
    for (int e = 0; e < (int)bs.size(); ++e)
      for (int j = 0; j < 3; ++j) {
        const int i = bs[e] + j;
        out[2 * i] = std::min(out[2 * i], e);
        out[2 * i + 1] = std::max(out[2 * i + 1], e);
      }
    for (int i = 0; i < nd; ++i)
      if (out[2 * i + 1] < 0) out[2 * i] = 0;  // empty support (cannot happen for valid tables)
  };
  std::vector<int> er1, er2;
  erange(b1, nd1, er1);
  erange(b2, nd2, er2);

  MeshView m{};
  int r;
  const void* d;
#define UP(field, vec, T)                                                      \
  r = upload(ctx, (vec).data(), (vec).size() * sizeof(T), &d);                 \
  if (r != COLLOP_OK) return r;                                                \
  m.field = static_cast<const T*>(d);
  UP(bp1, v1, double)
  UP(bp2, v2, double)
  std::vector<double> c1(C1, C1 + (size_t)9 * ne1), c2(C2, C2 + (size_t)9 * ne2);
  UP(C1, c1, double)
  UP(C2, c2, double)
  UP(bs1, b1, int)
  UP(bs2, b2, int)
  UP(chol1, ch1, double)
  UP(chol2, ch2, double)
  UP(minv1, mi1, double)
  UP(minv2, mi2, double)
  UP(phiq1, ph1, double)
  UP(phiq2, ph2, double)
  UP(qw, qw, double)
  UP(erange1, er1, int)
  UP(erange2, er2, int)
#undef UP
  m.nb1 = nb1;
  m.nb2 = nb2;
  m.ne1 = ne1;
  m.ne2 = ne2;
  m.nd1 = nd1;
  m.nd2 = nd2;
  m.n_dofs = (int)n_dofs;
  m.nq = n_quad;
  m.lo1 = bp1[0];
  m.hi1 = bp1[nb1 - 1];
  m.lo2 = bp2[0];
  m.hi2 = bp2[nb2 - 1];
  m.inv_h1 = ne1 / (m.hi1 - m.lo1);
  m.inv_h2 = ne2 / (m.hi2 - m.lo2);
  auto is_uniform = [](const std::vector<double>& bp) {
    const int ne = (int)bp.size() - 1;
    const double h = (bp.back() - bp.front()) / ne;
    for (int i = 0; i < ne; ++i)
      if (std::fabs((bp[i + 1] - bp[i]) - h) > 1e-6 * h) return 0;
    return 1;
  };
  m.uniform1 = is_uniform(v1);
  m.uniform2 = is_uniform(v2);
  ctx->mv = m;
  ctx->h_bp1 = v1;
  ctx->h_bp2 = v2;
  ctx->has_mesh = true;
  return COLLOP_OK;
}

// ---- hook-level ------------------------------------------------------------

static int collision_common(collop_ctx* ctx, int kind, const double* v, const double* w, const double* G,
                            double* dot_v, double* A, int64_t N) {
  int r = need_mesh(ctx);
  if (r) return r;
  if ((r = check_N(ctx, N))) return r;
  if (N == 0) return COLLOP_OK;
  const bool metric = kind >= 2;  // 2: metric FP64, 3: metric FP32c
  if (!v || !w || !G || !dot_v || (metric && !A)) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  if ((r = ensure_scratch(ctx, N))) return r;
  const int n = (int)N;
  choose_tile(ctx, n);
  HIPC(hipMemcpyAsync(ctx->b_v.p, v, 2 * N * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->b_w.p, w, N * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->b_G.p, G, 2 * N * 8, hipMemcpyHostToDevice, ctx->stream));
  const MeshView& m = ctx->mv;
  launch_pack_src(P<double>(ctx->b_v), P<double>(ctx->b_G), P<double>(ctx->b_w), n, m.lo1, m.hi1, m.lo2, m.hi2,
                  (kind == 1 || kind == 3) ? nullptr : P<Src64>(ctx->b_s64),
                  (kind == 1 || kind == 3) ? P<Src32>(ctx->b_s32) : nullptr, ctx->stream);
  CHECK_LAUNCH();
  const bool pow2w = equal_pow2_weights(w, N);
  int S;
  if (kind == 2) {
    prof_begin(ctx, 2);
    launch_pair_metric_f64(P<Src64>(ctx->b_s64), n, ctx->tile64, P<double>(ctx->b_part), ctx->stream);
    CHECK_LAUNCH();
    prof_end(ctx);
    S = ctx->tile64.slices;
  } else if (kind == 3) {
    prof_begin(ctx, 3);
    launch_pair_metric_f32c(P<Src32>(ctx->b_s32), n, ctx->tile32, P<double>(ctx->b_part), ctx->stream, nullptr,
                            pow2w && ctx->pow2w_path);
    CHECK_LAUNCH();
    prof_end(ctx);
    S = ctx->tile32.slices;
  } else {
    if ((r = run_pair(ctx, kind, n, P<double>(ctx->b_part), pow2w))) return r;
    S = kind == 0 ? ctx->tile64.slices : ctx->tile32.slices;
  }
  const int64_t len = (metric ? 5 : 2) * N;
  launch_slice_sum(P<double>(ctx->b_part), S, len, P<double>(ctx->b_out), ctx->stream);
  CHECK_LAUNCH();
  HIPC(hipMemcpyAsync(dot_v, ctx->b_out.p, 2 * N * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (metric) HIPC(hipMemcpyAsync(A, P<double>(ctx->b_out) + 2 * N, 3 * N * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  prof_collect(ctx);
  return COLLOP_OK;
}

int collop_collision_f64(collop_ctx* ctx, const double* v, const double* w, const double* G, double* dot_v,
                         int64_t N) {
  return collision_common(ctx, 0, v, w, G, dot_v, nullptr, N);
}

int collop_collision_f32c(collop_ctx* ctx, const double* v, const double* w, const double* G, double* dot_v,
                          int64_t N) {
  return collision_common(ctx, 1, v, w, G, dot_v, nullptr, N);
}

int collop_collision_metric_f64(collop_ctx* ctx, const double* v, const double* w, const double* G, double* dot_v,
                                double* A, int64_t N) {
  return collision_common(ctx, 2, v, w, G, dot_v, A, N);
}

int collop_collision_metric_f32c(collop_ctx* ctx, const double* v, const double* w, const double* G, double* dot_v,
                                 double* A, int64_t N) {
  return collision_common(ctx, 3, v, w, G, dot_v, A, N);
}

// The pair sum of collop_collision_f64 on caller-owned device arrays, with no
// allocation: the packed sources and the slice partials live in `scratch`
// (ordered form, kNoAllocSlices slices), launched on `stream` (null: the
// context's) without synchronisation.  Calls that overlap on one GPU race.
int collop_collision_f64_noalloc(collop_ctx* ctx, const double* v, const double* w, const double* G, double* dot_v,
                                 int64_t N, void* stream) {
  int r = need_mesh(ctx);
  if (r) return r;
  if (N < 0 || N > kNoAllocMaxN)
    return fail(ctx, COLLOP_ERR_INVALID,
                "N = " + std::to_string(N) + " exceeds the allocation-free scratch (at most 65536 markers)");
  if (N == 0) return COLLOP_OK;
  if (!v || !w || !G || !dot_v) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  double* base = ctx->device < kMaxDevices ? g_scratch_dev[ctx->device] : nullptr;
  if (!base) return fail(ctx, COLLOP_ERR_HIP, "the scratch array's device address is unknown on this GPU");
  HIPC(hipSetDevice(ctx->device));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
  const int n = (int)N;
  const MeshView& m = ctx->mv;
  Src64* src = reinterpret_cast<Src64*>(base);
  double* part = base + (size_t)8 * N;
  launch_pack_src(v, G, w, n, m.lo1, m.hi1, m.lo2, m.hi2, src, nullptr, st);
  CHECK_LAUNCH();
  launch_pair_f64(src, n, PairTile{1, kNoAllocSlices, 1, 0}, part, st);
  CHECK_LAUNCH();
  launch_slice_sum(part, kNoAllocSlices, 2 * N, dot_v, st);
  CHECK_LAUNCH();
  return COLLOP_OK;
}

static int rhs_common(collop_ctx* ctx, const double* v, const double* w, int64_t N, double* out, bool solve) {
  int r = need_mesh(ctx);
  if (r) return r;
  if ((r = check_N(ctx, N))) return r;
  if (!out || (N > 0 && (!v || !w))) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  if ((r = ensure_scratch(ctx, N))) return r;
  const int n = (int)N;
  const MeshView& m = ctx->mv;
  if (N > 0) {
    HIPC(hipMemcpyAsync(ctx->b_v.p, v, 2 * N * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(ctx->b_w.p, w, N * 8, hipMemcpyHostToDevice, ctx->stream));
  }
  const int nsb = scatter_blocks(n);
  double wsum = 0.0;
  for (int64_t i = 0; i < N; ++i) wsum += std::fabs(w[i]);
  launch_scatter(m, nullptr, P<double>(ctx->b_v), P<double>(ctx->b_w), n, 1, 1, wsum, nullptr, nullptr,
                 P<long long>(ctx->b_spart_sc), nullptr, ctx->stream);
  CHECK_LAUNCH();
  double* vec = P<double>(ctx->b_vec);
  if (solve) {
    launch_reduce_solve(m, P<long long>(ctx->b_spart_sc), nsb, wsum, nullptr, 1, vec, ctx->stream);
  } else {
    launch_rhs(m, P<long long>(ctx->b_spart_sc), nsb, wsum, vec, ctx->stream);
  }
  CHECK_LAUNCH();
  HIPC(hipMemcpyAsync(out, vec, (size_t)m.n_dofs * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

int collop_l2_rhs(collop_ctx* ctx, const double* v, const double* w, int64_t N, double* rhs) {
  return rhs_common(ctx, v, w, N, rhs, false);
}

int collop_l2_project(collop_ctx* ctx, const double* v, const double* w, int64_t N, double* f_coeffs) {
  return rhs_common(ctx, v, w, N, f_coeffs, true);
}

static int gather_common(collop_ctx* ctx, const double* v, int64_t N, const double* coeff, int mode, double fs_floor,
                         double g_max, double* out) {
  int r = need_mesh(ctx);
  if (r) return r;
  if ((r = check_N(ctx, N))) return r;
  if (N == 0) return COLLOP_OK;
  if (!v || !coeff || !out) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  if ((r = ensure_scratch(ctx, N))) return r;
  const MeshView& m = ctx->mv;
  double* vec = P<double>(ctx->b_vec);
  HIPC(hipMemcpyAsync(ctx->b_v.p, v, 2 * N * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(vec, coeff, (size_t)m.n_dofs * 8, hipMemcpyHostToDevice, ctx->stream));
  launch_gather(m, P<double>(ctx->b_v), vec, nullptr, nullptr, (int)N, mode, fs_floor, g_max, P<double>(ctx->b_out),
                nullptr, ctx->stream);
  CHECK_LAUNCH();
  HIPC(hipMemcpyAsync(out, ctx->b_out.p, 2 * N * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

int collop_grad_gather(collop_ctx* ctx, const double* v, int64_t N, const double* L, double* G) {
  return gather_common(ctx, v, N, L, 0, 0.0, 0.0, G);
}

int collop_loggrad_gather(collop_ctx* ctx, const double* v, int64_t N, const double* f_coeffs, double fs_floor,
                          double g_max, double* g) {
  if (ctx && !(fs_floor > 0.0 && g_max > 0.0))
    return fail(ctx, COLLOP_ERR_INVALID, "fs_floor and g_max must be positive");
  return gather_common(ctx, v, N, f_coeffs, 2, fs_floor, g_max, g);
}

int collop_quadrature(collop_ctx* ctx, const double* f_coeffs, int use_logsq, double* r_out, double* entropy) {
  int r = need_mesh(ctx);
  if (r) return r;
  if (!f_coeffs) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  if ((r = ensure_scratch(ctx, 1))) return r;
  const MeshView& m = ctx->mv;
  double* vec = P<double>(ctx->b_vec);
  double* sc = P<double>(ctx->b_scal);
  HIPC(hipMemcpyAsync(vec, f_coeffs, (size_t)m.n_dofs * 8, hipMemcpyHostToDevice, ctx->stream));
  launch_quad1(m, vec, vec, 1, use_logsq, P<double>(ctx->b_iota), P<double>(ctx->b_spart_q), ctx->stream);
  CHECK_LAUNCH();
  launch_quad2(m, P<double>(ctx->b_iota), vec + 2 * m.n_dofs, ctx->stream);
  CHECK_LAUNCH();
  launch_sum_rows(P<double>(ctx->b_spart_q), quad_blocks(m), 1, 0, 1, sc, ctx->stream);
  CHECK_LAUNCH();
  if (r_out) HIPC(hipMemcpyAsync(r_out, vec + 2 * m.n_dofs, (size_t)m.n_dofs * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (entropy) HIPC(hipMemcpyAsync(entropy, sc, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

}  // extern "C"

namespace {

// Run-record diagnostics of f_s = sum fc_k phi_k (functions.jl:300-343) from
// device arrays: out[1] = int max(-f_s, 0); with markers (v, w), also out[0] =
// ||f_s - f_p||_2, f_p their cell histogram density.
int diagnostics_dev(collop_ctx* ctx, const double* v, const double* w, int n, double w_abs_sum, const double* fc,
                    double* out) {
  double* sc = P<double>(ctx->b_scal);
  launch_fs_diagnostics(ctx->mv, v, w, n, w_abs_sum, fc, P<long long>(ctx->b_spart_sc), P<double>(ctx->b_stmp),
                        P<double>(ctx->b_spart_q), sc + 8, ctx->stream);
  CHECK_LAUNCH();
  double h[2];
  HIPC(hipMemcpyAsync(h, sc + 8, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  if (v) out[0] = std::sqrt(h[1]);
  out[1] = h[0];
  return COLLOP_OK;
}

}  // namespace

extern "C" {

int collop_negative_part_l1(collop_ctx* ctx, const double* f_coeffs, double* neg_part) {
  int r = need_mesh(ctx);
  if (r) return r;
  if (!f_coeffs || !neg_part) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  if ((r = ensure_scratch(ctx, 1))) return r;
  double* vec = P<double>(ctx->b_vec);
  HIPC(hipMemcpyAsync(vec, f_coeffs, (size_t)ctx->mv.n_dofs * 8, hipMemcpyHostToDevice, ctx->stream));
  double out[2] = {0.0, 0.0};
  if ((r = diagnostics_dev(ctx, nullptr, nullptr, 0, 0.0, vec, out))) return r;
  *neg_part = out[1];
  return COLLOP_OK;
}

int collop_fs_minus_fp_l2(collop_ctx* ctx, const double* v, const double* w, int64_t N, const double* f_coeffs,
                          double* fs_minus_fp) {
  int r = need_mesh(ctx);
  if (r) return r;
  if ((r = check_N(ctx, N))) return r;
  if (!f_coeffs || !fs_minus_fp || (N > 0 && (!v || !w))) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  if ((r = ensure_scratch(ctx, N))) return r;
  double wsum = 0.0;
  for (int64_t i = 0; i < N; ++i) wsum += std::fabs(w[i]);
  if (N > 0) {
    HIPC(hipMemcpyAsync(ctx->b_v.p, v, 2 * N * 8, hipMemcpyHostToDevice, ctx->stream));
    HIPC(hipMemcpyAsync(ctx->b_w.p, w, N * 8, hipMemcpyHostToDevice, ctx->stream));
  }
  double* vec = P<double>(ctx->b_vec);
  HIPC(hipMemcpyAsync(vec, f_coeffs, (size_t)ctx->mv.n_dofs * 8, hipMemcpyHostToDevice, ctx->stream));
  double out[2] = {0.0, 0.0};
  if ((r = diagnostics_dev(ctx, P<double>(ctx->b_v), P<double>(ctx->b_w), (int)N, wsum, vec, out))) return r;
  *fs_minus_fp = out[0];
  return COLLOP_OK;
}

int collop_lb_rhs(collop_ctx* ctx, const double* v, const double* w, const double* g, int64_t N, double nu,
                  double* dot_v, double* mult) {
  int r = need_mesh(ctx);
  if (r) return r;
  if ((r = check_N(ctx, N))) return r;
  if (N == 0) return COLLOP_OK;
  if (!v || !w || !g || !dot_v) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  if (!std::isfinite(nu)) return fail(ctx, COLLOP_ERR_INVALID, "nu must be finite");
  HIPC(hipSetDevice(ctx->device));
  if ((r = ensure_scratch(ctx, N))) return r;
  HIPC(hipMemcpyAsync(ctx->b_v.p, v, 2 * N * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->b_w.p, w, N * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->b_G.p, g, 2 * N * 8, hipMemcpyHostToDevice, ctx->stream));
  double* coef = P<double>(ctx->b_scal) + 10;
  launch_lb_rhs(ctx->mv, P<double>(ctx->b_v), nullptr, P<double>(ctx->b_w), nullptr, (int)N, nullptr, 0,
                P<double>(ctx->b_scal), nu, P<double>(ctx->b_G), P<double>(ctx->b_lbpart), coef,
                P<double>(ctx->b_out), ctx->stream);
  CHECK_LAUNCH();
  double h[5];
  HIPC(hipMemcpyAsync(dot_v, ctx->b_out.p, 2 * N * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipMemcpyAsync(h, coef, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  if (mult) {
    mult[0] = h[1];
    mult[1] = h[2];
    mult[2] = h[3];
  }
  return COLLOP_OK;
}

int collop_gmres_dense(collop_ctx* ctx, int n, const double* A, const double* b, double rtol, double atol, int itmax,
                       int mem, double* x, int* products, double* rnorm) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (n < 1 || !A || !b || !x) return fail(ctx, COLLOP_ERR_INVALID, "n >= 1 and non-null arrays required");
  if (mem < 1 || mem > kNkMaxKrylov || itmax < 0 || !(rtol >= 0.0) || !(atol >= 0.0))
    return fail(ctx, COLLOP_ERR_INVALID, "need 1 <= mem <= 64, itmax >= 0, rtol, atol >= 0");
  HIPC(hipSetDevice(ctx->device));
  NkDevice E;
  double* ex[2];
  int r = nk_setup(ctx, n, mem, E, ex, 2);
  if (r) return r;
  ENSURE(ctx->nk_A, (size_t)n * n * 8);
  HIPC(hipMemcpyAsync(ctx->nk_A.p, A, (size_t)n * n * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ex[1], b, (size_t)n * 8, hipMemcpyHostToDevice, ctx->stream));
  const double* Ad = P<double>(ctx->nk_A);
  NkDevice::Op op = [&](double* y, const double* u) -> int {
    launch_dense_mv(y, Ad, u, n, ctx->stream);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  };
  int k = 0;
  double rn = 0.0;
  if (E.gmres(ex[0], op, ex[1], mem, rtol, atol, itmax, &k, &rn))
    return fail(ctx, COLLOP_ERR_HIP, std::string("GMRES: ") + hipGetErrorString(hipGetLastError()));
  HIPC(hipMemcpyAsync(x, ex[0], (size_t)n * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  if (products) *products = k;
  if (rnorm) *rnorm = rn;
  return COLLOP_OK;
}

int collop_mass_solve(collop_ctx* ctx, const double* b, double* x) {
  int r = need_mesh(ctx);
  if (r) return r;
  if (!b || !x) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  if ((r = ensure_scratch(ctx, 1))) return r;
  const MeshView& m = ctx->mv;
  double* vec = P<double>(ctx->b_vec);
  HIPC(hipMemcpyAsync(vec, b, (size_t)m.n_dofs * 8, hipMemcpyHostToDevice, ctx->stream));
  launch_reduce_solve(m, nullptr, 0, 1.0, vec, 1, vec + m.n_dofs, ctx->stream);
  CHECK_LAUNCH();
  HIPC(hipMemcpyAsync(x, vec + m.n_dofs, (size_t)m.n_dofs * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

}  // extern "C"

// ===========================================================================
// Device-resident step
namespace {

// Projection of the state markers, its entropy S0 and the momentum / energy
// (main.jl:298-303).  Leaves fc in s_fc and S0, E, P1, P2 in scal[0], [3..5].
int project_state(collop_ctx* ctx, int use_logsq) {
  const MeshView& m = ctx->mv;
  const int n = (int)ctx->N;
  const int nsb = scatter_blocks(n);
  double* sc = P<double>(ctx->b_scal);
  launch_scatter(m, nullptr, P<double>(ctx->s_v), P<double>(ctx->s_w), n, 1, 1, ctx->w_abs_sum, nullptr, nullptr,
                 P<long long>(ctx->b_spart_sc), P<double>(ctx->b_mom), ctx->stream);
  CHECK_LAUNCH();
  launch_reduce_solve(m, P<long long>(ctx->b_spart_sc), nsb, ctx->w_abs_sum, nullptr, 1, P<double>(ctx->s_fc),
                      ctx->stream);
  CHECK_LAUNCH();
  launch_quad1(m, P<double>(ctx->s_fc), P<double>(ctx->s_fc), 1, use_logsq, nullptr, P<double>(ctx->b_spart_q),
               ctx->stream);
  CHECK_LAUNCH();
  launch_sum_rows(P<double>(ctx->b_spart_q), quad_blocks(m), 1, 0, 1, sc + 0, ctx->stream);
  CHECK_LAUNCH();
  // momentum / energy: columns of mom[nsb][3]
  launch_sum_rows(P<double>(ctx->b_mom), nsb, 3, 1, 3, sc + 3, ctx->stream);
  CHECK_LAUNCH();
  return COLLOP_OK;
}

// G = grad(M^{-1} r(f_s)) at x, with f_s = fc (main.jl:240-242); result in b_G.
int entropy_seed_gradient(collop_ctx* ctx, const double* x, const double* fc, int use_logsq) {
  const MeshView& m = ctx->mv;
  double* vec = P<double>(ctx->b_vec);
  launch_quad1(m, fc, fc, 1, use_logsq, P<double>(ctx->b_iota), nullptr, ctx->stream);
  CHECK_LAUNCH();
  launch_quad2(m, P<double>(ctx->b_iota), vec + 2 * m.n_dofs, ctx->stream);
  CHECK_LAUNCH();
  launch_reduce_solve(m, nullptr, 0, 1.0, vec + 2 * m.n_dofs, 1, vec + 3 * m.n_dofs, ctx->stream);
  CHECK_LAUNCH();
  launch_gather(m, x, vec + 3 * m.n_dofs, nullptr, nullptr, (int)ctx->N, 0, 0.0, 0.0, P<double>(ctx->b_G), nullptr,
                ctx->stream);
  CHECK_LAUNCH();
  return COLLOP_OK;
}

// landau_geff! (solver.jl:86-113) at vin: v_mid, dv, the entropy gradient at
// v_mid and the Gonzalez correction, then G_eff -- packed into pair-source
// records (s64 / s32, either may be null), copied to geff_out, and/or turned
// into the frozen-metric product D = A (G_eff - Gk).
int geff_dev(collop_ctx* ctx, const collop_step_params& p, const double* vin, Src64* s64, Src32* s32,
             double* geff_out, const double* Gk, const double* A, double* D) {
  const MeshView& m = ctx->mv;
  const int n = (int)ctx->N;
  const int nsb = scatter_blocks(n);
  double* vec = P<double>(ctx->b_vec);
  double* sc = P<double>(ctx->b_scal);
  // v_mid, dv and the rhs of v_mid (and of v_in for the Gonzalez entropy S(v_in))
  launch_scatter(m, P<double>(ctx->s_v), vin, P<double>(ctx->s_w), n, 2, p.use_gonzalez ? 2 : 1, ctx->w_abs_sum,
                 P<double>(ctx->b_vmid), P<double>(ctx->b_dv), P<long long>(ctx->b_spart_sc), nullptr, ctx->stream);
  CHECK_LAUNCH();
  // the mass solves run split over the columns and rows of the coefficient
  // array (bitwise the single-workgroup solve) when the 1D spaces allow it
  const bool split = m.nd1 <= 256 && m.nd2 <= 256;
  if (split)
    launch_solve_split(m, P<long long>(ctx->b_spart_sc), nsb, ctx->w_abs_sum, nullptr, p.use_gonzalez ? 2 : 1, vec,
                       P<double>(ctx->b_stmp), ctx->stream);
  else
    launch_reduce_solve(m, P<long long>(ctx->b_spart_sc), nsb, ctx->w_abs_sum, nullptr, p.use_gonzalez ? 2 : 1,
                        vec, ctx->stream);
  CHECK_LAUNCH();
  // r(f_mid) and S(f_in)
  launch_quad1(m, vec, vec + m.n_dofs, p.use_gonzalez ? 2 : 1, p.use_logsq, P<double>(ctx->b_iota),
               P<double>(ctx->b_spart_q), ctx->stream);
  CHECK_LAUNCH();
  launch_quad2(m, P<double>(ctx->b_iota), vec + 2 * m.n_dofs, ctx->stream);
  CHECK_LAUNCH();
  const int nqb = quad_blocks(m);
  if (split) {  // with the Gonzalez entropy S(f_in) summed in the same launch
    launch_solve_split(m, nullptr, 0, 1.0, vec + 2 * m.n_dofs, 1, vec + 3 * m.n_dofs, P<double>(ctx->b_stmp),
                       ctx->stream, p.use_gonzalez ? P<double>(ctx->b_spart_q) + nqb : nullptr, nqb,
                       p.use_gonzalez ? sc + 1 : nullptr);
    CHECK_LAUNCH();
  } else {
    launch_reduce_solve(m, nullptr, 0, 1.0, vec + 2 * m.n_dofs, 1, vec + 3 * m.n_dofs, ctx->stream);
    CHECK_LAUNCH();
    if (p.use_gonzalez) {
      launch_sum_rows(P<double>(ctx->b_spart_q) + nqb, nqb, 1, 0, 1, sc + 1, ctx->stream);
      CHECK_LAUNCH();
    }
  }
  // dS = -w grad L(v_mid) and the partial sums of dv.dS, |dv|^2
  launch_gather(m, P<double>(ctx->b_vmid), vec + 3 * m.n_dofs, P<double>(ctx->s_w), P<double>(ctx->b_dv), n, 1, 0.0,
                0.0, P<double>(ctx->b_dS), P<double>(ctx->b_dpart), ctx->stream);
  CHECK_LAUNCH();
  launch_geff_pack(P<double>(ctx->b_vmid), P<double>(ctx->b_dS), P<double>(ctx->b_dv), P<double>(ctx->s_w), n,
                   P<double>(ctx->b_dpart), gather_blocks(n), sc + 1, sc + 0, p.use_gonzalez, m.lo1, m.hi1, m.lo2,
                   m.hi2, s64, s32, sc + 2, ctx->stream, geff_out, Gk, A, D);
  CHECK_LAUNCH();
  return COLLOP_OK;
}

// Anderson epilogue of a map whose value is v0e + dte * (slice sum of part)
// (S slices of 2N), then the statistics, published to the pinned host buffer.
int epilogue_run(collop_ctx* ctx, const EpiArgs& a) {
  const int n = (int)ctx->N;
  launch_epilogue(a, n, ctx->stream);
  CHECK_LAUNCH();
  return publish_stat(ctx, epi_blocks(n));
}

int epilogue_dev(collop_ctx* ctx, const double* part, int S, const double* v0e, double dte, const double* vin,
                 double* Gv, int col, int hist, bool clusters = false) {
  EpiArgs a{};
  if (clusters) {
    a.cl_row = P<int>(ctx->b_cl_row);
    a.cl_x = P<double>(ctx->b_cl_x);
  }
  a.part = part;
  a.S = S;
  a.v0 = v0e;
  a.v1 = vin;
  a.dt = dte;
  a.Gv = Gv;
  a.r = P<double>(ctx->s_r);
  a.rp = P<double>(ctx->s_rp);
  a.Gp = P<double>(ctx->s_Gp);
  a.dF = P<double>(ctx->s_dF);
  a.dG = P<double>(ctx->s_dG);
  a.col = col;
  a.hist = hist;
  a.stat_part = P<double>(ctx->b_statp);
  return epilogue_run(ctx, a);
}

// Near pairs found by the step's first outer map -> CSR by target (sorted, so
// the correction sums in a fixed order).  Too many pairs: no near field.
int build_near(collop_ctx* ctx) {
  int cnt = 0;
  HIPC(hipMemcpy(&cnt, ctx->b_near_count.p, sizeof(int), hipMemcpyDeviceToHost));
  ctx->near_nt = 0;
  if (cnt <= 0 || cnt > kNearCap) return COLLOP_OK;
  std::vector<int2> L(cnt);
  HIPC(hipMemcpy(L.data(), ctx->b_near_list.p, (size_t)cnt * sizeof(int2), hipMemcpyDeviceToHost));
  std::sort(L.begin(), L.end(), [](const int2& a, const int2& b) { return a.x != b.x ? a.x < b.x : a.y < b.y; });
  std::vector<int> tg, off, al(cnt);
  for (int i = 0; i < cnt; ++i) {
    if (i == 0 || L[i].x != L[i - 1].x) {
      tg.push_back(L[i].x);
      off.push_back(i);
    }
    al[i] = L[i].y;
  }
  off.push_back(cnt);
  HIPC(hipMemcpy(ctx->b_ntg.p, tg.data(), tg.size() * 4, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(ctx->b_noff.p, off.data(), off.size() * 4, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(ctx->b_nal.p, al.data(), al.size() * 4, hipMemcpyHostToDevice));
  ctx->near_nt = (int)tg.size();
  launch_near_index(ctx->near_nt, P<int>(ctx->b_ntg), (int)ctx->N, P<int>(ctx->b_nidx), ctx->stream);
  CHECK_LAUNCH();
  return COLLOP_OK;
}

// Stiff clusters of the step (extension): every pair of the state's markers
// with max(w_i, w_j) / |d| > theta, found cell by cell on the device, joined
// closest pair first unless the cluster would exceed kmax markers; clusters of
// two or more markers in CSR, members ascending, ordered by first member.  Too
// many links or members for the buffers: no clusters this step.
int build_clusters_dev(collop_ctx* ctx, double theta, int kmax) {
  const MeshView& m = ctx->mv;
  const int n = (int)ctx->N;
  if (ctx->cl_total > 0) {
    launch_cluster_rows(P<int>(ctx->b_cl_idx), ctx->cl_total, false, P<int>(ctx->b_cl_row), ctx->stream);
    CHECK_LAUNCH();
  }
  ctx->ncl = ctx->cl_total = 0;
  if (!(theta > 0.0)) return COLLOP_OK;
  launch_cell_sort(m, P<double>(ctx->s_v), n, P<int>(ctx->b_cell_of), P<int>(ctx->b_cell_cnt),
                   P<int>(ctx->b_cell_start), P<int>(ctx->b_cell_sorted), ctx->stream);
  CHECK_LAUNCH();
  launch_near_cells(m, P<double>(ctx->s_v), P<double>(ctx->s_w), n, theta, ctx->w_max, P<int>(ctx->b_cell_start),
                    P<int>(ctx->b_cell_sorted), P<NearLink>(ctx->b_links), P<int>(ctx->b_link_count), kLinkCap,
                    ctx->stream);
  CHECK_LAUNCH();
  int cnt = 0;
  HIPC(hipMemcpyAsync(&cnt, ctx->b_link_count.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  if (cnt <= 0 || cnt > kLinkCap) return COLLOP_OK;
  std::vector<NearLink> L(cnt);
  HIPC(hipMemcpy(L.data(), ctx->b_links.p, (size_t)cnt * sizeof(NearLink), hipMemcpyDeviceToHost));
  std::sort(L.begin(), L.end(), [](const NearLink& a, const NearLink& b) {
    return a.q != b.q ? a.q < b.q : (a.i != b.i ? a.i < b.i : a.j < b.j);
  });
  std::vector<int> nodes;
  nodes.reserve(2 * cnt);
  for (const NearLink& l : L) {
    nodes.push_back(l.i);
    nodes.push_back(l.j);
  }
  std::sort(nodes.begin(), nodes.end());
  nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
  auto id = [&](int a) { return (int)(std::lower_bound(nodes.begin(), nodes.end(), a) - nodes.begin()); };
  std::vector<int> par(nodes.size()), sz(nodes.size(), 1);
  for (size_t t = 0; t < par.size(); ++t) par[t] = (int)t;
  auto find = [&](int a) {
    while (par[a] != a) a = par[a] = par[par[a]];
    return a;
  };
  const int cap = std::min(std::max(2, kmax), kClusterMax);
  for (const NearLink& l : L) {
    int a = find(id(l.i)), b = find(id(l.j));
    if (a == b || sz[a] + sz[b] > cap) continue;
    if (a > b) std::swap(a, b);
    par[b] = a;
    sz[a] += sz[b];
  }
  std::vector<std::vector<int>> groups(nodes.size());
  for (size_t t = 0; t < nodes.size(); ++t) groups[find((int)t)].push_back(nodes[t]);
  std::vector<int> off(1, 0), idx;
  for (auto& g : groups) {  // groups[root] lists members ascending; roots are first members
    if (g.size() < 2) continue;
    idx.insert(idx.end(), g.begin(), g.end());
    off.push_back((int)idx.size());
  }
  if (idx.empty() || (int)idx.size() > kClMembersCap) return COLLOP_OK;
  ctx->ncl = (int)off.size() - 1;
  ctx->cl_total = (int)idx.size();
  HIPC(hipMemcpyAsync(ctx->b_cl_off.p, off.data(), off.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->b_cl_idx.p, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  launch_cluster_rows(P<int>(ctx->b_cl_idx), ctx->cl_total, true, P<int>(ctx->b_cl_row), ctx->stream);
  CHECK_LAUNCH();
  launch_fill_int(P<int>(ctx->b_cl_have), ctx->ncl, 0, ctx->stream);
  CHECK_LAUNCH();
  launch_fill_int(P<int>(ctx->b_cl_its), ctx->ncl, 0, ctx->stream);
  CHECK_LAUNCH();
  launch_fill_int(P<int>(ctx->b_cl_fail), 1, 0, ctx->stream);
  CHECK_LAUNCH();

  // the host lists must outlive the asynchronous copies
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

// The cluster solves of an FP64 map (after its pair sum): every cluster's rows
// solved exactly with the map's v_mid, G_eff, lambda, L and partials.
int cluster_solve_dev(collop_ctx* ctx, const collop_step_params& p, const double* part, int S) {
  ClusterArgs a{};
  a.m = ctx->mv;
  a.off = P<int>(ctx->b_cl_off);
  a.idx = P<int>(ctx->b_cl_idx);
  a.N = (int)ctx->N;
  a.dt = p.dt;
  a.v0 = P<double>(ctx->s_v);
  a.w = P<double>(ctx->s_w);
  a.vmid = P<double>(ctx->b_vmid);
  a.dS = P<double>(ctx->b_dS);
  a.dv = P<double>(ctx->b_dv);
  a.lam = P<double>(ctx->b_scal) + 2;
  a.L = P<double>(ctx->b_vec) + 3 * ctx->mv.n_dofs;
  a.part = part;
  a.S = S;
  a.mroot = P<double>(ctx->b_cl_root);
  a.have_root = P<int>(ctx->b_cl_have);
  a.x_out = P<double>(ctx->b_cl_x);
  a.fail_count = P<int>(ctx->b_cl_fail);
  a.newton_its = P<int>(ctx->b_cl_its);
  launch_cluster_solve(a, ctx->ncl, ctx->stream);
  CHECK_LAUNCH();
  return COLLOP_OK;
}

// One Picard map (solver.jl:176-227, Landau branch): Gv = v0 + dt vdot(v_mid, G_eff),
// followed by the epilogue of the Anderson iteration.
// With stiff clusters (FP64 maps only), each cluster's rows take the cluster
// solve's value (Anderson mixes that map) and the plain map's residual goes to
// the statistics beside it.
int picard_map_dev(collop_ctx* ctx, const collop_step_params& p, int kind, const double* vin, int col, int hist) {
  int r = geff_dev(ctx, p, vin, kind == 0 ? P<Src64>(ctx->b_s64) : nullptr, kind == 1 ? P<Src32>(ctx->b_s32) : nullptr,
                   nullptr, nullptr, nullptr, nullptr);
  if (r) return r;
  if ((r = run_pair(ctx, kind, (int)ctx->N, P<double>(ctx->b_part), ctx->w_pow2))) return r;
  const int S = kind == 0 ? ctx->tile64.slices : ctx->tile32.slices;
  const bool cl = kind == 0 && ctx->ncl > 0;
  if (cl && (r = cluster_solve_dev(ctx, p, P<double>(ctx->b_part), S))) return r;
  return epilogue_dev(ctx, P<double>(ctx->b_part), S, P<double>(ctx->s_v), p.dt, vin, P<double>(ctx->s_Gv), col, hist,
                      cl);
}

// Outer map of defect correction (solver.jl:590-594): G(v1) and the metric A in
// one pair pass (compute_collision_metric!, functions.jl:249-290), FP64 (kind 0)
// or FP32-compensated (kind 1); keeps G_eff(v1) in s_Gk, [F, A] in s_FA and
// G(v1) in s_Gvo.
int outer_map_dev(collop_ctx* ctx, const collop_step_params& p, int kind, bool find_near, bool metric) {
  const int n = (int)ctx->N;
  int r = geff_dev(ctx, p, P<double>(ctx->s_v1), kind == 0 ? P<Src64>(ctx->b_s64) : nullptr,
                   kind == 1 ? P<Src32>(ctx->b_s32) : nullptr, P<double>(ctx->s_Gk), nullptr, nullptr, nullptr);
  if (r) return r;
  const PairTile& t = kind == 0 ? ctx->tile64 : ctx->tile32;
  if (!metric) {  // the plain pair sum; A of an earlier outer map stays in s_FA + 2N
    if ((r = run_pair(ctx, kind, n, P<double>(ctx->b_part), ctx->w_pow2))) return r;
    launch_slice_sum(P<double>(ctx->b_part), t.slices, (int64_t)2 * n, P<double>(ctx->s_FA), ctx->stream);
    CHECK_LAUNCH();
    return epilogue_dev(ctx, P<double>(ctx->s_FA), 1, P<double>(ctx->s_v), p.dt, P<double>(ctx->s_v1),
                        P<double>(ctx->s_Gvo), -1, 0);
  }
  const double th2 = p.dc_near_theta * p.dc_near_theta;
  NearOut near{th2, P<int>(ctx->b_near_count), P<int2>(ctx->b_near_list), kNearCap,
               th2 > 0.0 ? ctx->w_max * ctx->w_max / th2 : 0.0};
  if (find_near) HIPC(hipMemsetAsync(ctx->b_near_count.p, 0, sizeof(int), ctx->stream));
  prof_begin(ctx, 2 + kind);
  if (kind == 0)
    launch_pair_metric_f64(P<Src64>(ctx->b_s64), n, t, P<double>(ctx->b_part), ctx->stream, find_near ? &near : nullptr);
  else
    launch_pair_metric_f32c(P<Src32>(ctx->b_s32), n, t, P<double>(ctx->b_part), ctx->stream,
                            find_near ? &near : nullptr, ctx->w_pow2 && ctx->pow2w_path);
  CHECK_LAUNCH();
  prof_end(ctx);
  launch_slice_sum(P<double>(ctx->b_part), t.slices, (int64_t)5 * n, P<double>(ctx->s_FA), ctx->stream);
  CHECK_LAUNCH();
  return epilogue_dev(ctx, P<double>(ctx->s_FA), 1, P<double>(ctx->s_v), p.dt, P<double>(ctx->s_v1),
                      P<double>(ctx->s_Gvo), -1, 0);
}

// Inner map of defect correction (solver.jl:598-607): Phi(u) = G(v_k) - dt A_k
// (G_eff(u) - G_k), no pair sum; Phi(u) lands in s_Gv with the Anderson epilogue.
// The same map with its small kernels fused (split mass solves only): the
// pending Anderson update applied by the scatter (the updated iterate lands in
// s_v1n, then takes s_v1's place), quadrature passes 1 and 2 in one launch, and
// G_eff, D = A (G_eff - G_k) and the near-pair corrections inside the
// epilogue.  The same arithmetic as the unfused map.
int frozen_map_fused(collop_ctx* ctx, const collop_step_params& p, int col, int hist, const UpdateArgs* upd) {
  const MeshView& m = ctx->mv;
  const int n = (int)ctx->N;
  const int nsb = scatter_blocks(n);
  double* vec = P<double>(ctx->b_vec);
  double* sc = P<double>(ctx->b_scal);
  const int fields = p.use_gonzalez ? 2 : 1;
  launch_scatter(m, P<double>(ctx->s_v), P<double>(ctx->s_v1), P<double>(ctx->s_w), n, 2, fields, ctx->w_abs_sum,
                 P<double>(ctx->b_vmid), P<double>(ctx->b_dv), P<long long>(ctx->b_spart_sc), nullptr, ctx->stream,
                 upd, upd ? P<double>(ctx->s_v1n) : nullptr);
  CHECK_LAUNCH();
  if (upd) std::swap(ctx->s_v1, ctx->s_v1n);
  launch_solve_split(m, P<long long>(ctx->b_spart_sc), nsb, ctx->w_abs_sum, nullptr, fields, vec,
                     P<double>(ctx->b_stmp), ctx->stream);
  CHECK_LAUNCH();
  const int nqb = quad_blocks(m);
  launch_quad_r(m, vec, p.use_gonzalez ? vec + m.n_dofs : nullptr, p.use_logsq, vec + 2 * m.n_dofs,
                P<double>(ctx->b_spart_q), ctx->stream);
  CHECK_LAUNCH();
  launch_solve_split(m, nullptr, 0, 1.0, vec + 2 * m.n_dofs, 1, vec + 3 * m.n_dofs, P<double>(ctx->b_stmp),
                     ctx->stream, p.use_gonzalez ? P<double>(ctx->b_spart_q) : nullptr, nqb,
                     p.use_gonzalez ? sc + 1 : nullptr);
  CHECK_LAUNCH();
  launch_gather(m, P<double>(ctx->b_vmid), vec + 3 * m.n_dofs, P<double>(ctx->s_w), P<double>(ctx->b_dv), n, 1, 0.0,
                0.0, P<double>(ctx->b_dS), P<double>(ctx->b_dpart), ctx->stream);
  CHECK_LAUNCH();
  EpiArgs a{};
  a.part = nullptr;
  a.S = 0;
  a.v0 = P<double>(ctx->s_Gvo);
  a.v1 = P<double>(ctx->s_v1);
  a.dt = -p.dt;
  a.Gv = P<double>(ctx->s_Gv);
  a.r = P<double>(ctx->s_r);
  a.rp = P<double>(ctx->s_rp);
  a.Gp = P<double>(ctx->s_Gp);
  a.dF = P<double>(ctx->s_dF);
  a.dG = P<double>(ctx->s_dG);
  a.col = col;
  a.hist = hist;
  a.stat_part = P<double>(ctx->b_statp);
  FrozenArgs f{};
  f.dS = P<double>(ctx->b_dS);
  f.dv = P<double>(ctx->b_dv);
  f.w = P<double>(ctx->s_w);
  f.Gk = P<double>(ctx->s_Gk);
  f.A = P<double>(ctx->s_FA) + 2 * (size_t)n;
  f.dpart = P<double>(ctx->b_dpart);
  f.nparts = gather_blocks(n);
  f.S1 = sc + 1;
  f.S0 = sc + 0;
  f.use_gonzalez = p.use_gonzalez;
  if (ctx->near_nt > 0) {
    f.nidx = P<int>(ctx->b_nidx);
    f.off = P<int>(ctx->b_noff);
    f.al = P<int>(ctx->b_nal);
    f.vm = P<double>(ctx->b_vmid);
    f.vmk = P<double>(ctx->s_vmk);
  }
  launch_epilogue_frozen(a, f, n, ctx->stream);
  CHECK_LAUNCH();
  return publish_stat(ctx, epi_blocks(n));
}

int frozen_map_dev(collop_ctx* ctx, const collop_step_params& p, int col, int hist) {
  const int n = (int)ctx->N;
  const bool near = ctx->near_nt > 0;
  int r = geff_dev(ctx, p, P<double>(ctx->s_v1), nullptr, nullptr, near ? P<double>(ctx->s_Ge) : nullptr,
                   P<double>(ctx->s_Gk), P<double>(ctx->s_FA) + 2 * (size_t)n, P<double>(ctx->s_D));
  if (r) return r;
  if (near) {
    launch_near(ctx->near_nt, P<int>(ctx->b_ntg), P<int>(ctx->b_noff), P<int>(ctx->b_nal), n, P<double>(ctx->s_w),
                P<double>(ctx->b_vmid), P<double>(ctx->s_Ge), P<double>(ctx->s_vmk), P<double>(ctx->s_Gk),
                P<double>(ctx->s_D), ctx->stream);
    CHECK_LAUNCH();
  }
  return epilogue_dev(ctx, P<double>(ctx->s_D), 1, P<double>(ctx->s_Gvo), -p.dt, P<double>(ctx->s_v1),
                      P<double>(ctx->s_Gv), col, hist);
}

// One Lenard-Bernstein Picard map (solver.jl:176-227, LB branch): the
// projections of v_mid (and of v_in for Gonzalez), g = clamped grad f / f at
// v_mid (LOGGRAD_FN), lambda and the drift multipliers, v_dot = -nu (G_eff + A
// + B v_mid), then Gv = v0 + dt v_dot with the Anderson epilogue.
int lb_map_dev(collop_ctx* ctx, const collop_step_params& p, const double* vin, int col, int hist) {
  const MeshView& m = ctx->mv;
  const int n = (int)ctx->N;
  const int nsb = scatter_blocks(n);
  double* vec = P<double>(ctx->b_vec);
  double* sc = P<double>(ctx->b_scal);
  const int fields = p.use_gonzalez ? 2 : 1;
  launch_scatter(m, P<double>(ctx->s_v), vin, P<double>(ctx->s_w), n, 2, fields, ctx->w_abs_sum,
                 P<double>(ctx->b_vmid), P<double>(ctx->b_dv), P<long long>(ctx->b_spart_sc), nullptr, ctx->stream);
  CHECK_LAUNCH();
  if (m.nd1 <= 256 && m.nd2 <= 256)
    launch_solve_split(m, P<long long>(ctx->b_spart_sc), nsb, ctx->w_abs_sum, nullptr, fields, vec,
                       P<double>(ctx->b_stmp), ctx->stream);
  else
    launch_reduce_solve(m, P<long long>(ctx->b_spart_sc), nsb, ctx->w_abs_sum, nullptr, fields, vec, ctx->stream);
  CHECK_LAUNCH();
  if (p.use_gonzalez) {  // S(f_in) partials (solver.jl:196-197)
    launch_quad1(m, vec + m.n_dofs, nullptr, 1, p.use_logsq, nullptr, P<double>(ctx->b_spart_q), ctx->stream);
    CHECK_LAUNCH();
  }
  launch_lb_rhs(m, P<double>(ctx->b_vmid), vec, P<double>(ctx->s_w), P<double>(ctx->b_dv), n,
                p.use_gonzalez ? P<double>(ctx->b_spart_q) : nullptr, p.use_gonzalez ? quad_blocks(m) : 0, sc + 0,
                p.nu, P<double>(ctx->b_G), P<double>(ctx->b_lbpart), sc + 10, P<double>(ctx->b_dS), ctx->stream);
  CHECK_LAUNCH();
  return epilogue_dev(ctx, P<double>(ctx->b_dS), 1, P<double>(ctx->s_v), p.dt, vin, P<double>(ctx->s_Gv), col, hist);
}

// The same LB map with its small kernels fused (split mass solves only): the
// entropy quadrature's blocks inside the gather launch, and the reduction,
// lambda, multipliers and v_dot inside the epilogue.  The same arithmetic as
// lb_map_dev on s_v1.  (The Anderson update stays a launch of its own: with the
// LB window of 8 columns, applying it inside the scatter, whose two projections
// would each re-read the columns, costs more than the launch it saves.)
int lb_map_fused(collop_ctx* ctx, const collop_step_params& p, int col, int hist) {
  const MeshView& m = ctx->mv;
  const int n = (int)ctx->N;
  double* vec = P<double>(ctx->b_vec);
  double* sc = P<double>(ctx->b_scal);
  const int fields = p.use_gonzalez ? 2 : 1;
  launch_scatter(m, P<double>(ctx->s_v), P<double>(ctx->s_v1), P<double>(ctx->s_w), n, 2, fields, ctx->w_abs_sum,
                 P<double>(ctx->b_vmid), P<double>(ctx->b_dv), P<long long>(ctx->b_spart_sc), nullptr, ctx->stream);
  CHECK_LAUNCH();
  launch_solve_split(m, P<long long>(ctx->b_spart_sc), scatter_blocks(n), ctx->w_abs_sum, nullptr, fields, vec,
                     P<double>(ctx->b_stmp), ctx->stream);
  CHECK_LAUNCH();
  launch_lb_gather_q(m, P<double>(ctx->b_vmid), vec, p.use_gonzalez ? vec + m.n_dofs : nullptr, p.use_logsq,
                     P<double>(ctx->s_w), P<double>(ctx->b_dv), n, P<double>(ctx->b_G), P<double>(ctx->b_lbpart),
                     P<double>(ctx->b_spart_q), ctx->stream);
  CHECK_LAUNCH();
  EpiArgs a{};
  a.part = nullptr;
  a.S = 0;
  a.v0 = P<double>(ctx->s_v);
  a.v1 = P<double>(ctx->s_v1);
  a.dt = p.dt;
  a.Gv = P<double>(ctx->s_Gv);
  a.r = P<double>(ctx->s_r);
  a.rp = P<double>(ctx->s_rp);
  a.Gp = P<double>(ctx->s_Gp);
  a.dF = P<double>(ctx->s_dF);
  a.dG = P<double>(ctx->s_dG);
  a.col = col;
  a.hist = hist;
  a.stat_part = P<double>(ctx->b_statp);
  LbEpiArgs l{};
  l.x = P<double>(ctx->b_vmid);
  l.g = P<double>(ctx->b_G);
  l.dv = P<double>(ctx->b_dv);
  l.w = P<double>(ctx->s_w);
  l.part = P<double>(ctx->b_lbpart);
  l.nparts = lb_blocks(n);
  l.qpart = p.use_gonzalez ? P<double>(ctx->b_spart_q) : nullptr;
  l.nq = p.use_gonzalez ? quad_blocks(m) : 0;
  l.S0 = sc + 0;
  l.nu = p.nu;
  l.coef = sc + 10;
  launch_epilogue_lb(a, l, n, ctx->stream);
  CHECK_LAUNCH();
  return publish_stat(ctx, epi_blocks(n));
}

// The LB predictor (main.jl:233-238): g at v from the state's projection, the
// multipliers at v, v1 = v + dt v_dot.
int lb_predictor_dev(collop_ctx* ctx, const collop_step_params& p) {
  const int n = (int)ctx->N;
  double* sc = P<double>(ctx->b_scal);
  launch_lb_rhs(ctx->mv, P<double>(ctx->s_v), P<double>(ctx->s_fc), P<double>(ctx->s_w), nullptr, n, nullptr, 0,
                sc + 0, p.nu, P<double>(ctx->b_G), P<double>(ctx->b_lbpart), sc + 10, P<double>(ctx->b_dS),
                ctx->stream);
  CHECK_LAUNCH();
  launch_euler(P<double>(ctx->b_dS), 1, P<double>(ctx->s_v), p.dt, n, P<double>(ctx->s_v1), P<double>(ctx->s_dotv),
               ctx->stream);
  CHECK_LAUNCH();
  return COLLOP_OK;
}

// The device Newton-Krylov engine over vectors of length n with a basis of mem
// vectors; nvec extra work vectors follow the engine's own in nk_vec.
int nk_setup(collop_ctx* ctx, int n, int mem, NkDevice& E, double** extra, int nextra) {
  const size_t nb = (size_t)std::max(n, 1) * 8;
  ENSURE(ctx->nk_V, nb * (size_t)(mem + 1));
  ENSURE(ctx->nk_vec, nb * (size_t)(8 + nextra));
  ENSURE(ctx->nk_part, nk_part_size() * 8);
  ENSURE(ctx->nk_h, (size_t)(kNkMaxKrylov + 1) * 8);
  double* v = P<double>(ctx->nk_vec);
  const size_t ld = (size_t)std::max(n, 1);
  E.n = n;
  E.st = ctx->stream;
  E.V = P<double>(ctx->nk_V);
  E.ldv = ld;
  E.part = P<double>(ctx->nk_part);
  E.hdev = P<double>(ctx->nk_h);
  E.w = v;
  E.Ft = v + ld;
  E.xt = v + 2 * ld;
  E.xp = v + 3 * ld;
  E.Fp = v + 4 * ld;
  E.d = v + 5 * ld;
  E.xb = v + 6 * ld;
  E.Fb = v + 7 * ld;
  for (int i = 0; i < nextra; ++i) extra[i] = v + (8 + i) * ld;
  return COLLOP_OK;
}

// step_newton! (solver.jl:437-486): inexact Newton on F(v) = v - G(v) from the
// initial guess in s_v1, every residual one device Picard map (Landau or LB);
// with exit_picard_step, v1 <- v - F(v) = G(v).
int newton_solve(collop_ctx* ctx, const collop_step_params& p, collop_step_stats* out) {
  const int n2 = 2 * (int)ctx->N;
  const bool lb = p.collision_model == COLLOP_MODEL_LB;
  NkDevice E;
  double* ex[2];
  int r = nk_setup(ctx, n2, p.nk_krylov_max, E, ex, 2);
  if (r) return r;
  double* x = ex[0];
  double* F = ex[1];
  HIPC(hipMemcpyAsync(x, ctx->s_v1.p, (size_t)n2 * 8, hipMemcpyDeviceToDevice, ctx->stream));
  double nv;
  if (E.norm(x, &nv)) return fail(ctx, COLLOP_ERR_HIP, "Newton-Krylov: device norm failed");
  const double eff_tol = std::max(p.tol * (nv + 1e-30), p.abs_floor);
  const double fd_h = p.nk_fd_rel > 0 ? p.nk_fd_rel * std::max(nv, 1.0) : (p.nk_fd_h > 0 ? p.nk_fd_h : 1e-6);
  int maps = 0;
  NkDevice::Resid Fn = [&](const double* xv, double* Fv, double* nrm) -> int {
    const int rr = lb ? lb_map_dev(ctx, p, xv, -1, 0) : picard_map_dev(ctx, p, 0, xv, -1, 0);
    if (rr) return rr;
    ++maps;
    launch_neg(Fv, P<double>(ctx->s_r), n2, ctx->stream);  // F = v - G(v) = -(G(v) - v)
    *nrm = std::sqrt(ctx->h_stat[0]);
    return std::isfinite(*nrm) ? 0 : -2;
  };
  NkResult res{};
  const int st = E.newton(x, Fn, F, p.nk_krylov_max, eff_tol, p.max_iter, fd_h, p.nk_eta_max, &res);
  if (st == -2) return fail(ctx, COLLOP_ERR_NONFINITE, "non-finite residual in a Newton-Krylov map " +
                                                             std::to_string(maps));
  if (st) return fail(ctx, COLLOP_ERR_HIP, std::string("Newton-Krylov: ") + hipGetErrorString(hipGetLastError()));
  if (p.exit_picard_step)
    launch_axpby(P<double>(ctx->s_v1), x, -1.0, F, n2, ctx->stream);
  else
    HIPC(hipMemcpyAsync(ctx->s_v1.p, x, (size_t)n2 * 8, hipMemcpyDeviceToDevice, ctx->stream));
  CHECK_LAUNCH();
  out->iters = res.n_evals;
  out->iters_f32 = 0;
  out->n_restart = res.n_newton;
  out->exit_reason = res.status;
  out->residual = res.rnorm;
  out->r0 = res.rnorm0;
  out->inner_maps = 0;
  return COLLOP_OK;
}

// One line per map of a traced solve (COLLOP_TRACE set).
void trace_map(const AndersonRun& an, const char* tag) {
  if (an.last_f32)
    std::fprintf(stderr, "  [%s f32] k=%d |r|=%.3e floor=%.3e best=%.3e\n", tag, an.maps(), an.last_nrm_r,
                 an.last_floor32, an.best32());
  else
    std::fprintf(stderr, "  [%s f64] k=%d |r|=%.3e (mixed %.3e) tol=%.3e\n", tag, an.maps(), an.last_nrm_map,
                 an.last_nrm_r, an.last_eff_tol);
}

int copy_d2d(collop_ctx* ctx, DBuf& dst, const DBuf& src, size_t bytes) {
  HIPC(hipMemcpyAsync(dst.p, src.p, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return COLLOP_OK;
}

int validate_params(collop_ctx* ctx, const collop_step_params* p) {
  if (!p) return fail(ctx, COLLOP_ERR_INVALID, "null params");
  if (!(p->dt > 0.0) || !std::isfinite(p->dt)) return fail(ctx, COLLOP_ERR_INVALID, "dt must be positive");
  if (p->m_anderson < 1 || p->m_anderson > kMaxWindow)
    return fail(ctx, COLLOP_ERR_INVALID, "m_anderson must be in [1, 8]");
  if (p->max_iter < 1) return fail(ctx, COLLOP_ERR_INVALID, "max_iter must be >= 1");
  if (p->stag_window < 1) return fail(ctx, COLLOP_ERR_INVALID, "stag_window must be >= 1");
  if (!(p->damping > 0.0 && p->damping <= 1.0)) return fail(ctx, COLLOP_ERR_INVALID, "damping must be in (0, 1]");
  if (p->precision != COLLOP_PREC_F64 && p->precision != COLLOP_PREC_F32C)
    return fail(ctx, COLLOP_ERR_INVALID, "unknown precision");
  if (p->warmstart != COLLOP_WARM_EULER && p->warmstart != COLLOP_WARM_MLP)
    return fail(ctx, COLLOP_ERR_INVALID, "unknown warmstart");
  if (p->warmstart == COLLOP_WARM_MLP && !(ctx->m_has_params && ctx->m_has_norm))
    return fail(ctx, COLLOP_ERR_NO_MODEL, "no warm-start model loaded on this context (collop_mlp_set_params / "
                                          "collop_mlp_set_norm)");
  if (p->solver != COLLOP_SOLVER_ANDERSON && p->solver != COLLOP_SOLVER_DEFECT && p->solver != COLLOP_SOLVER_NEWTON)
    return fail(ctx, COLLOP_ERR_INVALID, "unknown solver");
  if (p->solver == COLLOP_SOLVER_NEWTON &&
      (p->precision != COLLOP_PREC_F64 || p->nk_krylov_max < 1 || p->nk_krylov_max > kNkMaxKrylov ||
       !(p->nk_eta_max > 0.0 && p->nk_eta_max < 1.0) || !(p->nk_fd_h >= 0.0) || !(p->nk_fd_rel >= 0.0) ||
       !std::isfinite(p->nk_fd_h) || !std::isfinite(p->nk_fd_rel)))
    return fail(ctx, COLLOP_ERR_INVALID,
                "Newton-Krylov runs FP64 maps and needs 1 <= nk_krylov_max <= 64, 0 < nk_eta_max < 1 and finite "
                "nk_fd_h, nk_fd_rel >= 0");
  if (p->solver == COLLOP_SOLVER_DEFECT &&
      (!(p->dc_eta > 0.0 && p->dc_eta < 1.0) || p->dc_max_inner < 1 || p->dc_stag_window < 1 ||
       !(p->dc_eta_f64 == 0.0 || (p->dc_eta_f64 > 0.0 && p->dc_eta_f64 < 1.0))))
    return fail(ctx, COLLOP_ERR_INVALID,
                "defect correction needs 0 < dc_eta < 1, dc_eta_f64 = 0 or in (0, 1), dc_max_inner >= 1, "
                "dc_stag_window >= 1");
  if (p->solver == COLLOP_SOLVER_DEFECT &&
      (p->dc_outer_m < 0 || p->dc_outer_m > kMaxWindow || !(p->dc_near_theta >= 0.0) ||
       !std::isfinite(p->dc_near_theta) || p->dc_metric_reuse < 0 || p->dc_metric_reuse > 2))
    return fail(ctx, COLLOP_ERR_INVALID,
                "defect correction needs 0 <= dc_outer_m <= 8, dc_near_theta >= 0 and dc_metric_reuse in {0, 1, 2}");
  if (p->dc_fallback != 0 && p->dc_fallback != 1) return fail(ctx, COLLOP_ERR_INVALID, "dc_fallback must be 0 or 1");
  if (p->dc_sticky != 0 && p->dc_sticky != 1) return fail(ctx, COLLOP_ERR_INVALID, "dc_sticky must be 0 or 1");
  if (p->dc_sticky && !(p->solver == COLLOP_SOLVER_DEFECT && p->dc_fallback))
    return fail(ctx, COLLOP_ERR_INVALID, "dc_sticky needs the defect-correction solver with dc_fallback = 1");
  if (p->collision_model != COLLOP_MODEL_LANDAU && p->collision_model != COLLOP_MODEL_LB)
    return fail(ctx, COLLOP_ERR_INVALID, "unknown collision_model");
  if (p->collision_model == COLLOP_MODEL_LB &&
      (p->solver == COLLOP_SOLVER_DEFECT || p->precision != COLLOP_PREC_F64 || p->warmstart != COLLOP_WARM_EULER ||
       p->record_training || !std::isfinite(p->nu)))
    return fail(ctx, COLLOP_ERR_INVALID,
                "the Lenard-Bernstein model runs the Anderson or Newton-Krylov solver with FP64 maps and the Euler "
                "predictor (no pair sum, no metric, no learned warm start) and needs a finite nu");
  if (p->record_training && ctx->m_size + ctx->N > ctx->m_cap)
    return fail(ctx, COLLOP_ERR_INVALID, "training dataset full (collop_mlp_dataset_alloc)");
  return COLLOP_OK;
}

// step_defect! (solver.jl:569-634) from the initial guess in s_v1; on return
// s_v1 holds the solution.  Outer iteration k: the full map G(v_k) with the
// metric A_k; then Anderson (solver.jl:507-545: window m_anderson, damping,
// no stagnation logic) on the frozen-metric map Phi_k to
// max(dc_eta ||G(v_k) - v_k||, abs_floor / 2), at most dc_max_inner maps.
// exit_reason of defect_solve when the step continues with the Anderson solve
constexpr int kDefectHandover = 3;
// the hand-over's outer contraction test: an outer iteration must reduce the
// residual at least by this factor, FP64 or FP32c (the latter while the
// residual is above kHandoverFloorMargin times the FP32 floor).  An Anderson
// map reduces it by about half in the steps with stiff near pairs, and an
// outer iteration (a pair sum with the metric plus its inner maps) costs about
// 1.3 Anderson maps: 0.5^1.3 = 0.4.
constexpr double kHandoverRatio = 0.4;
constexpr double kHandoverRatio32 = 0.4;
constexpr double kHandoverFloorMargin = 4.0;

int defect_solve(collop_ctx* ctx, const collop_step_params& p, collop_step_stats* out) {
  const int n = (int)ctx->N;
  const size_t n2b = (size_t)2 * n * 8;
  const int mw = p.m_anderson;
  double nrm0 = 0.0, nrm_best = std::numeric_limits<double>::infinity();
  double nrm_best_window = std::numeric_limits<double>::infinity(), res_out = 0.0;
  int exit_reason = 2, k_done = p.max_iter, inner = 0, r;
  const bool trace = std::getenv("COLLOP_TRACE") != nullptr;
  // F32C: FP32-compensated outer maps (F and A) until the outer residual reaches
  // the FP32 rounding floor of the step increment, is predicted to reach it at
  // the next map, or stops improving; then FP64 outer maps, which alone decide
  // convergence, stagnation and the result.
  bool use32 = p.precision == COLLOP_PREC_F32C;
  const double floor_factor = p.f32_floor_factor > 0.0 ? p.f32_floor_factor : 16.0;
  double best32 = std::numeric_limits<double>::infinity(), prev32 = 0.0;
  bool switch_after = false;
  int since_best32 = 0, iters32 = 0, k64 = 0;
  double prev64 = 0.0;
  // outer Anderson window (extension): Anderson with beta = 1 on v -> T(v), the
  // inner solution, with the epilogue's column bookkeeping
  const int mo = p.dc_outer_m;
  // the fused frozen map needs the split solves
  const bool fused = ctx->fused_maps && ctx->mv.nd1 <= 256 && ctx->mv.nd2 <= 256;
  int ohist = 0, oslot = 0;
  bool oprev = false;
  double oATA[kMaxWindow * kMaxWindow] = {0};
  ctx->near_nt = 0;
  // the best iterate starts as the initial guess
  if ((r = copy_d2d(ctx, ctx->s_best, ctx->s_v1, n2b))) return r;
  for (int k = 1; k <= p.max_iter; ++k) {
    // the hand-over needs at least one iteration left for the Anderson solve
    const bool fallback = p.dc_fallback && k < p.max_iter;
    const bool find_near = k == 1 && p.dc_near_theta > 0.0;
    // the metric A: every outer map (mode 0); FP64 maps after an FP32c phase
    // reuse the last FP32c A (mode 1); the step's first outer map only (mode 2)
    const bool metric = k == 1 || p.dc_metric_reuse == 0 || (p.dc_metric_reuse == 1 && (use32 || iters32 == 0));
    if ((r = outer_map_dev(ctx, p, use32 ? 1 : 0, find_near, metric))) return r;
    if (find_near && (r = build_near(ctx))) return r;
    const double nr = std::sqrt(ctx->h_stat[0]), nv = std::sqrt(ctx->h_stat[1]);
    if (!std::isfinite(nr) || !std::isfinite(nv))
      return fail(ctx, COLLOP_ERR_NONFINITE, "non-finite residual in outer map " + std::to_string(k));
    if (trace) std::fprintf(stderr, "  [defect %s] k=%d |r|=%.3e inner so far %d\n", use32 ? "f32" : "f64", k, nr, inner);
    if (k == 1) nrm0 = nr;
    if (use32) {
      iters32++;
      const double floor32 = floor_factor * 5.96e-8 * std::sqrt(ctx->h_stat[2]);
      if (nr < best32) {
        best32 = nr;
        since_best32 = 0;
      } else {
        since_best32++;
      }
      // with the hand-over, an FP32c outer iteration that does not reduce the
      // residual by kHandoverRatio32, well above the FP32 floor, hands the step
      // to the Anderson solve from the outer iterate
      if (fallback && prev32 > 0.0 && nr > kHandoverRatio32 * prev32 && nr > kHandoverFloorMargin * floor32) {
        exit_reason = kDefectHandover;
        k_done = k;
        res_out = nr;
        break;
      }
      if (nr < std::max(floor32, std::max(p.tol * (nv + 1e-30), p.abs_floor)) || since_best32 >= 2 ||
          k >= p.max_iter / 2) {
        use32 = false;  // re-evaluate this iterate with an FP64 outer map
        oprev = false;  // and restart the outer window on FP64 maps
        ohist = oslot = 0;
        continue;
      }
      // When the outer contraction so far, nr / prev32, predicts that the next
      // FP32c map would only reach the FP32 floor, finish this iteration and
      // evaluate the next iterate with an FP64 map: an FP32c map at the floor
      // would serve only to detect the floor.
      if (prev32 > 0.0 && nr < prev32 && nr * (nr / prev32) < floor32) switch_after = true;
      prev32 = nr;
    } else {
      k64++;
      if (nr < nrm_best) {
        nrm_best = nr;
        if ((r = copy_d2d(ctx, ctx->s_best, ctx->s_Gvo, n2b))) return r;
      }
      if (nr < std::max(p.tol * (nv + 1e-30), p.abs_floor)) {
        std::swap(ctx->s_v1, ctx->s_Gvo);
        exit_reason = 0;
        k_done = k;
        res_out = nr;
        break;
      }
      // with the hand-over, an FP64 outer iteration that does not reduce the
      // residual by kHandoverRatio hands the step to the Anderson solve from
      // the best iterate (its Picard image)
      if (fallback && k64 >= 2 && nr > kHandoverRatio * prev64) {
        std::swap(ctx->s_v1, ctx->s_best);
        exit_reason = kDefectHandover;
        k_done = k;
        res_out = nrm_best;
        break;
      }
      prev64 = nr;
    }
    if (!use32 && k64 > p.dc_stag_window && k64 % p.dc_stag_window == 0) {
      if ((nrm_best_window - nrm_best) / (nrm_best_window + 1e-30) < p.stag_rel_tol) {
        std::swap(ctx->s_v1, ctx->s_best);
        exit_reason = fallback ? kDefectHandover : 1;
        k_done = k;
        res_out = nrm_best;
        break;
      }
      nrm_best_window = nrm_best;
    }
    if (ctx->near_nt > 0 && (r = copy_d2d(ctx, ctx->s_vmk, ctx->b_vmid, n2b))) return r;  // v_mid of v_k
    if ((mo > 0 || p.dc_fallback) && (r = copy_d2d(ctx, ctx->s_xk, ctx->s_v1, n2b))) return r;
    // inner Anderson on u = Phi_k(u), u = s_v1 (solver.jl:507-545)
    const double eta = (!use32 && p.dc_eta_f64 > 0.0) ? p.dc_eta_f64 : p.dc_eta;
    const double tol_in = std::max(eta * nr, 0.5 * p.abs_floor);
    int hist = 0, slot = 0;
    UpdateArgs pend{};
    bool has_pend = false;
    double ATA[kMaxWindow * kMaxWindow] = {0};
    bool inner_ok = false;
    for (int j = 1; j <= p.dc_max_inner; ++j) {
      const int col = j == 1 ? -1 : slot % mw;
      const int hist_new = j == 1 ? 0 : std::min(hist + 1, mw);
      if (j == 1) {
        // Phi_k(v_k) = G(v_k): at u = v_k the frozen-metric correction and the
        // near-pair differences vanish exactly, so the first inner iterate's map
        // value, residual and statistics are the outer map's (s_Gvo, s_r, h_stat)
        if ((r = copy_d2d(ctx, ctx->s_Gv, ctx->s_Gvo, n2b))) return r;
      } else if (fused) {
        if ((r = frozen_map_fused(ctx, p, col, hist_new, has_pend ? &pend : nullptr))) return r;
        has_pend = false;
      } else if ((r = frozen_map_dev(ctx, p, col, hist_new))) {
        return r;
      }
      inner++;
      const double* st = ctx->h_stat;
      const double ri = std::sqrt(st[0]);
      if (!std::isfinite(ri)) return fail(ctx, COLLOP_ERR_NONFINITE, "non-finite residual in a frozen-metric map");
      if (ri < tol_in) {
        std::swap(ctx->s_v1, ctx->s_Gv);  // u <- Phi(u)
        inner_ok = true;
        break;
      }
      // the next frozen map's scatter applies the update (the loop's last
      // iterate is updated here, as is every one when the maps are not fused)
      const bool defer = fused && j < p.dc_max_inner;
      if (j == 1) {
        if (defer) {
          pend = UpdateArgs{};
          pend.Gv = P<double>(ctx->s_Gv);
          pend.dG = P<double>(ctx->s_dG);
          pend.hist = 0;
          pend.beta = p.damping;
          has_pend = true;
        } else {
          launch_update(P<double>(ctx->s_v1), P<double>(ctx->s_Gv), P<double>(ctx->s_dG), nullptr, 0, p.damping, n,
                        ctx->stream);
          CHECK_LAUNCH();
        }
      } else {
        slot = slot % mw + 1;
        hist = hist_new;
        for (int i = 0; i < hist; ++i) {
          ATA[col * kMaxWindow + i] = st[3 + i];
          ATA[i * kMaxWindow + col] = st[3 + i];
        }
        double A[kMaxWindow * kMaxWindow], b[kMaxWindow], g[kMaxWindow];
        double dm = 0.0;
        for (int i = 0; i < hist; ++i) {
          for (int c = 0; c < hist; ++c) A[i * hist + c] = ATA[i * kMaxWindow + c];
          b[i] = st[3 + kMaxWindow + i];
          dm += A[i * hist + i];
        }
        const double l2 = p.reg_factor * dm / hist + 1e-30;
        for (int i = 0; i < hist; ++i) A[i * hist + i] += l2;
        small_solve(hist, A, b, g);
        if (defer) {
          pend = UpdateArgs{};
          pend.Gv = P<double>(ctx->s_Gv);
          pend.dG = P<double>(ctx->s_dG);
          for (int i = 0; i < hist; ++i) pend.g[i] = g[i];
          pend.hist = hist;
          pend.beta = p.damping;
          has_pend = true;
        } else {
          launch_update(P<double>(ctx->s_v1), P<double>(ctx->s_Gv), P<double>(ctx->s_dG), g, hist, p.damping, n,
                        ctx->stream);
          CHECK_LAUNCH();
        }
      }
      std::swap(ctx->s_rp, ctx->s_r);
      std::swap(ctx->s_Gp, ctx->s_Gv);
    }
    if (has_pend) return fail(ctx, COLLOP_ERR_HIP, "internal: an Anderson update was left pending");
    if (!inner_ok && fallback) {
      // the frozen-metric solve did not converge (stiff near pairs): the
      // Anderson solve takes over from the outer iterate x_k
      std::swap(ctx->s_v1, ctx->s_xk);
      exit_reason = kDefectHandover;
      k_done = k;
      res_out = nr;
      break;
    }
    if (mo > 0) {  // outer Anderson: f = U - x_k with U the inner solution
      EpiArgs a{};
      a.part = nullptr;
      a.S = 0;
      a.v0 = P<double>(ctx->s_v1);
      a.v1 = P<double>(ctx->s_xk);
      a.dt = 0.0;
      a.Gv = P<double>(ctx->s_oU);
      a.r = P<double>(ctx->s_of);
      a.rp = P<double>(ctx->s_ofp);
      a.Gp = P<double>(ctx->s_oUp);
      a.dF = P<double>(ctx->s_odF);
      a.dG = P<double>(ctx->s_odG);
      a.col = oprev ? oslot % mo : -1;
      a.hist = oprev ? std::min(ohist + 1, mo) : 0;
      a.stat_part = P<double>(ctx->b_statp);
      if ((r = epilogue_run(ctx, a))) return r;
      if (oprev) {
        const double* st = ctx->h_stat;
        oslot = oslot % mo + 1;
        ohist = a.hist;
        for (int i = 0; i < ohist; ++i) {
          oATA[a.col * kMaxWindow + i] = st[3 + i];
          oATA[i * kMaxWindow + a.col] = st[3 + i];
        }
        double A[kMaxWindow * kMaxWindow], b[kMaxWindow], g[kMaxWindow];
        double dm = 0.0;
        for (int i = 0; i < ohist; ++i) {
          for (int c = 0; c < ohist; ++c) A[i * ohist + c] = oATA[i * kMaxWindow + c];
          b[i] = st[3 + kMaxWindow + i];
          dm += A[i * ohist + i];
        }
        const double l2 = p.reg_factor * dm / ohist + 1e-30;
        for (int i = 0; i < ohist; ++i) A[i * ohist + i] += l2;
        small_solve(ohist, A, b, g);
        launch_update(P<double>(ctx->s_v1), P<double>(ctx->s_oU), P<double>(ctx->s_odG), g, ohist, 1.0, n,
                      ctx->stream);
        CHECK_LAUNCH();
      }
      std::swap(ctx->s_of, ctx->s_ofp);
      std::swap(ctx->s_oU, ctx->s_oUp);
      oprev = true;
    }
    if (switch_after) {  // the predicted switch: FP64 outer maps from the next iterate on
      switch_after = false;
      use32 = false;
      oprev = false;
      ohist = oslot = 0;
    }
  }
  if (exit_reason == 2) {
    std::swap(ctx->s_v1, ctx->s_best);
    res_out = nrm_best;
  }
  out->iters = k_done;
  out->iters_f32 = iters32;
  out->n_restart = 0;
  out->exit_reason = exit_reason;
  out->residual = res_out;
  out->r0 = nrm0;
  out->inner_maps = inner;
  return COLLOP_OK;
}

// The Anderson solve of a Landau step with its control on the device
// (solver.jl:300-423; no stiff clusters): per map the host only launches work.
// The kernels of map k read the plan {live, kind, col, hist} that the control
// kernel of map k - 1 wrote, and the control kernel of map k runs the state
// machine (the window's least squares included) and plans map k + 1.  While the
// maps are FP32c the host launches map k + 1 once map k has published (its
// precision is decided there); once they are FP64, which they then stay, it
// launches map k + 1 while map k runs.  The launches of a map after the exit
// return at once.  Same arithmetic, and so the same bits, as the host-controlled
// loop.
int anderson_device(collop_ctx* ctx, const collop_step_params& p, AndersonRun& an) {
  const int n = (int)ctx->N;
  ENSURE(ctx->b_actl, sizeof(AndCtl));
  AndCtl* dc = P<AndCtl>(ctx->b_actl);
  ctx->h_actl.resize(sizeof(AndCtl));
  const AndersonPlan pl0 = an.plan();
  AndCtl h{};
  h.gate[0] = 1;
  h.gate[1] = pl0.kind;
  h.gate[2] = pl0.col;
  h.gate[3] = pl0.hist;
  h.maps = 0;
  h.err = 0;
  h.run = an;
  std::memcpy(ctx->h_actl.data(), &h, sizeof(AndCtl));
  HIPC(hipMemcpyAsync(dc, ctx->h_actl.data(), sizeof(AndCtl), hipMemcpyHostToDevice, ctx->stream));
  const unsigned seq0 = ctx->seq;
  volatile unsigned* pub = ctx->h_seq;
  bool f32 = pl0.kind == 1;
  int k = 0, seen = 0, r;
  while (k < p.max_iter) {
    const int need = f32 ? k : k - 1;
    if (need > seen) {
      if ((r = wait_seq(ctx, seq0 + (unsigned)need))) return r;
      seen = (int)(__atomic_load_n(ctx->h_seq, __ATOMIC_ACQUIRE) - seq0);
      if (pub[1] == 0 || pub[3] != 0) break;  // the solve has ended
      if (seen >= k) f32 = pub[2] == 1;       // precision of map seen + 1
    }
    const int kind = f32 ? 1 : 0;
    ctx->prof_map = ++k;
    if ((r = geff_dev(ctx, p, P<double>(ctx->s_v1), kind == 0 ? P<Src64>(ctx->b_s64) : nullptr,
                      kind == 1 ? P<Src32>(ctx->b_s32) : nullptr, nullptr, nullptr, nullptr, nullptr)))
      return r;
    if ((r = run_pair(ctx, kind, n, P<double>(ctx->b_part), ctx->w_pow2, &dc->gate[0]))) return r;
    EpiArgs a{};
    a.part = P<double>(ctx->b_part);
    a.S = ctx->tile64.slices;
    a.S32 = ctx->tile32.slices;
    a.gate = &dc->gate[0];
    a.v0 = P<double>(ctx->s_v);
    a.v1 = P<double>(ctx->s_v1);
    a.dt = p.dt;
    a.Gv = P<double>(ctx->s_Gv);
    a.r = P<double>(ctx->s_r);
    a.rp = P<double>(ctx->s_rp);
    a.Gp = P<double>(ctx->s_Gp);
    a.dF = P<double>(ctx->s_dF);
    a.dG = P<double>(ctx->s_dG);
    a.stat_part = P<double>(ctx->b_statp);
    launch_epilogue(a, n, ctx->stream);
    CHECK_LAUNCH();
    launch_anderson_ctl(P<double>(ctx->b_statp), epi_blocks(n), dc, k, ctx->h_seq_dev, seq0, ctx->stream);
    CHECK_LAUNCH();
    launch_anderson_apply(dc, k, P<double>(ctx->s_v1), P<double>(ctx->s_Gv), P<double>(ctx->s_dG),
                          P<double>(ctx->s_r), P<double>(ctx->s_rp), P<double>(ctx->s_Gp), P<double>(ctx->s_best), n,
                          p.exit_picard_step, ctx->stream);
    CHECK_LAUNCH();
  }
  ctx->prof_map = 0;
  HIPC(hipMemcpyAsync(ctx->h_actl.data(), dc, sizeof(AndCtl), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  std::memcpy(&h, ctx->h_actl.data(), sizeof(AndCtl));
  ctx->seq = seq0 + (unsigned)h.maps;
  prof_collect(ctx, h.maps);
  if (h.err) return fail(ctx, COLLOP_ERR_NONFINITE, "non-finite residual in Picard map " + std::to_string(h.maps));
  an = h.run;
  return COLLOP_OK;
}

// One implicit step on the device-resident state (main.jl:226-307 with
// step_anderson!, solver.jl:300-423).
// A step of the hand-over memory (dc_sticky) runs step_anderson! from the
// start with FP64 maps and predictor, as the hand-over continues a step.
constexpr int kStickyOuter = 2;  // a hand-over within this many outer iterations arms the memory

int step_once(collop_ctx* ctx, const collop_step_params& p_in, collop_step_stats* out) {
  const bool skip_dc = p_in.solver == COLLOP_SOLVER_DEFECT && p_in.dc_sticky && ctx->dc_skip;
  collop_step_params p = p_in;
  if (skip_dc) {
    p.solver = COLLOP_SOLVER_ANDERSON;
    p.precision = COLLOP_PREC_F64;
    p.cl_theta = 0.0;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const long syncs0 = ctx->stat_syncs;
  const int n = (int)ctx->N;
  const size_t n2b = (size_t)2 * n * 8;
  int r;
  const MeshView& m = ctx->mv;
  const bool lb = p.collision_model == COLLOP_MODEL_LB;
  if (lb) {
    if ((r = lb_predictor_dev(ctx, p))) return r;
  } else {
    // ---- predictor: v1 = v + dt vdot(v, G), G = grad(M^{-1} r(f_s(v)))
    if ((r = entropy_seed_gradient(ctx, P<double>(ctx->s_v), P<double>(ctx->s_fc), p.use_logsq))) return r;
    // (the predictor is only the initial guess: F32C runs take its pair sum in FP32c)
    const int pk = p.precision == COLLOP_PREC_F32C ? 1 : 0;
    launch_pack_src(P<double>(ctx->s_v), P<double>(ctx->b_G), P<double>(ctx->s_w), n, m.lo1, m.hi1, m.lo2, m.hi2,
                    pk == 0 ? P<Src64>(ctx->b_s64) : nullptr, pk == 1 ? P<Src32>(ctx->b_s32) : nullptr, ctx->stream);
    CHECK_LAUNCH();
    if ((r = run_pair(ctx, pk, n, P<double>(ctx->b_part), ctx->w_pow2))) return r;
    launch_euler(P<double>(ctx->b_part), pk == 0 ? ctx->tile64.slices : ctx->tile32.slices, P<double>(ctx->s_v),
                 p.dt, n, P<double>(ctx->s_v1), P<double>(ctx->s_dotv), ctx->stream);
    CHECK_LAUNCH();
  }
  // ---- warm-start hook (main.jl:249-254): features of v, optional record,
  // optional MLP correction of the predictor
  float* rec = nullptr;
  if (p.record_training) {
    if (ctx->m_size + n > ctx->m_cap) return fail(ctx, COLLOP_ERR_INVALID, "training dataset full");
    rec = P<float>(ctx->m_data) + (size_t)ctx->m_size * 8;
  } else if (p.warmstart == COLLOP_WARM_MLP) {
    ENSURE(ctx->m_rec, (size_t)n * 8 * sizeof(float));
    rec = P<float>(ctx->m_rec);
  }
  if (rec) {
    launch_gather(m, P<double>(ctx->s_v), P<double>(ctx->s_fc), nullptr, nullptr, n, 3, 0.0, 0.0,
                  P<double>(ctx->b_out), nullptr, ctx->stream);
    CHECK_LAUNCH();
    launch_features(P<double>(ctx->s_v), P<double>(ctx->b_G), P<double>(ctx->b_out), n, rec, ctx->stream);
    CHECK_LAUNCH();
  }
  if (p.warmstart == COLLOP_WARM_MLP) {
    ENSURE(ctx->m_y, (size_t)n * 2 * sizeof(float));
    MlpArgs a{};
    a.params = P<float>(ctx->m_params);
    a.rec = rec;
    a.nsamples = n;
    a.tiles = nullptr;
    a.ntiles = (n + 15) / 16;
    for (int j = 0; j < 6; ++j) {
      a.mu[j] = ctx->m_mu[j];
      a.isd[j] = ctx->m_isd[j];
    }
    a.oscale[0] = ctx->m_oscale[0];
    a.oscale[1] = ctx->m_oscale[1];
    a.mode = 0;
    a.y = P<float>(ctx->m_y);
    launch_mlp(a, std::min(a.ntiles, 2048), ctx->stream);
    CHECK_LAUNCH();
    launch_warm_apply(P<double>(ctx->s_v), P<double>(ctx->s_dotv), P<float>(ctx->m_y), ctx->m_oscale[0],
                      ctx->m_oscale[1], p.dt, n, P<double>(ctx->s_v1), ctx->stream);
    CHECK_LAUNCH();
  }

  // stiff clusters of the step (Anderson, Landau): found on the state's markers
  const bool want_cl = !lb && p.solver == COLLOP_SOLVER_ANDERSON && p.cl_theta > 0.0;
  if (want_cl || ctx->cl_total > 0) {
    if ((r = build_clusters_dev(ctx, want_cl ? p.cl_theta : 0.0, p.cl_kmax))) return r;
  }
  collop_step_stats dc{};
  if (p.solver == COLLOP_SOLVER_DEFECT) {
    if ((r = defect_solve(ctx, p, &dc))) return r;
  } else if (p.solver == COLLOP_SOLVER_NEWTON) {
    if ((r = newton_solve(ctx, p, &dc))) return r;
  }
  // ---- Anderson (solver.jl:300-423).  A defect-correction step that handed
  // over (dc_fallback) continues here from its outer iterate with the
  // iterations it left, in FP64 maps: a hand-over means stiff near pairs, where
  // FP32c maps slow the Anderson solve down (their rounding is amplified along
  // the stiff modes).
  const bool handover = p.solver == COLLOP_SOLVER_DEFECT && dc.exit_reason == kDefectHandover;
  collop_step_params pa = p;
  if (handover) {
    pa.solver = COLLOP_SOLVER_ANDERSON;
    pa.max_iter = p.max_iter - dc.iters;  // >= 1: a hand-over leaves at least one
    pa.precision = COLLOP_PREC_F64;
  }
  const bool anderson = pa.solver == COLLOP_SOLVER_ANDERSON;
  AndersonRun an;
  an.start(pa, ctx->ncl > 0);
  // The best iterate starts as the initial guess (solver.jl:331).
  if (anderson && (r = copy_d2d(ctx, ctx->s_best, ctx->s_v1, n2b))) return r;
  const char* trace = std::getenv("COLLOP_TRACE") ? "run" : nullptr;
  const bool lb_fused = lb && ctx->fused_maps && m.nd1 <= 256 && m.nd2 <= 256;
  if (anderson && ctx->dev_solve && !lb && ctx->ncl == 0 && !trace) {
    if ((r = anderson_device(ctx, pa, an))) return r;
  }
  while (anderson && !an.done) {
    const AndersonPlan pl = an.plan();
    if (lb_fused) {
      if ((r = lb_map_fused(ctx, pa, pl.col, pl.hist))) return r;
    } else if (lb) {
      if ((r = lb_map_dev(ctx, pa, P<double>(ctx->s_v1), pl.col, pl.hist))) return r;
    } else if ((r = picard_map_dev(ctx, pa, pl.kind, P<double>(ctx->s_v1), pl.col, pl.hist))) {
      return r;
    }
    AndersonAct act;
    if (an.after_map(ctx->h_stat, act) != COLLOP_OK)
      return fail(ctx, COLLOP_ERR_NONFINITE, "non-finite residual in Picard map " + std::to_string(an.maps()));
    if (trace) trace_map(an, trace);
    // the best iterate is saved by the update's launch when one follows, else by a copy
    const bool best_in_update = act.copy_best && act.exit == kContinue && act.update;
    if (act.copy_best && !best_in_update &&
        (r = copy_d2d(ctx, ctx->s_best, p.exit_picard_step ? ctx->s_Gv : ctx->s_v1, n2b)))
      return r;
    if (act.exit == kConverged) {
      // the exit Picard step (solver.jl:351).  With stiff clusters it is the
      // map Anderson mixes: its cluster rows are the clusters' exact solutions
      // at the final iterate, where the plain map would amplify the final
      // residual by the clusters' stiffness.
      if (p.exit_picard_step) std::swap(ctx->s_v1, ctx->s_Gv);
      break;
    }
    if (act.exit != kContinue) {
      std::swap(ctx->s_v1, ctx->s_best);
      break;
    }
    if (act.update) {
      launch_update(P<double>(ctx->s_v1), P<double>(ctx->s_Gv), P<double>(ctx->s_dG), act.gamma, act.h, act.beta, n,
                    ctx->stream, best_in_update ? P<double>(p.exit_picard_step ? ctx->s_Gv : ctx->s_v1) : nullptr,
                    best_in_update ? P<double>(ctx->s_best) : nullptr);
      CHECK_LAUNCH();
    }
    if (act.rotate) {  // rp <- r, Gp <- Gv (pointer rotation instead of copies)
      std::swap(ctx->s_rp, ctx->s_r);
      std::swap(ctx->s_Gp, ctx->s_Gv);
    }
  }
  if (trace && ctx->ncl > 0) {  // per cluster size: clusters, Newton iterations of the step, failures
    std::vector<int> off(ctx->ncl + 1), its(ctx->ncl), have(ctx->ncl);
    HIPC(hipMemcpy(off.data(), ctx->b_cl_off.p, off.size() * 4, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(its.data(), ctx->b_cl_its.p, its.size() * 4, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(have.data(), ctx->b_cl_have.p, have.size() * 4, hipMemcpyDeviceToHost));
    int cnt[kClusterMax + 1] = {0}, mx[kClusterMax + 1] = {0}, bad[kClusterMax + 1] = {0};
    long sum[kClusterMax + 1] = {0};
    for (int c = 0; c < ctx->ncl; ++c) {
      const int k = off[c + 1] - off[c];
      cnt[k]++;
      sum[k] += its[c];
      mx[k] = std::max(mx[k], its[c]);
      bad[k] += have[c] < 0;
    }
    for (int k = 2; k <= kClusterMax; ++k)
      if (cnt[k])
        std::fprintf(stderr, "  [clusters] size %2d: %4d clusters, Newton its mean %.1f max %d, unsolved at exit %d\n", k,
                     cnt[k], (double)sum[k] / cnt[k], mx[k], bad[k]);
  }
  // v <- v1 and the new projection / entropy / moments (main.jl:296-303)
  std::swap(ctx->s_v, ctx->s_v1);
  if (p.record_training) {
    launch_targets(P<double>(ctx->s_v1), P<double>(ctx->s_v), P<double>(ctx->s_dotv), p.dt, n, rec, ctx->stream);
    CHECK_LAUNCH();
    ctx->m_size += n;
  }
  if ((r = project_state(ctx, p.use_logsq))) return r;
  double sc[6];
  int cl_fail = 0;
  HIPC(hipMemcpyAsync(sc, ctx->b_scal.p, 6 * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (ctx->ncl > 0) HIPC(hipMemcpyAsync(&cl_fail, ctx->b_cl_fail.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  prof_collect(ctx);
  if (p_in.dc_sticky)  // the memory: armed by an early hand-over, cleared by a converged Anderson step
    ctx->dc_skip = skip_dc ? an.exit_reason != 0 : handover && dc.iters <= kStickyOuter;
  if (out) {
    out->handover = handover ? dc.iters : 0;
    out->dc_skipped = skip_dc ? 1 : 0;
    if (handover) {
      out->iters = dc.iters + an.k_done;
      out->iters_f32 = dc.iters_f32 + an.iters32;
      out->n_restart = an.n_restart;
      out->exit_reason = an.exit_reason;
      out->residual = an.res_out;
      out->r0 = dc.r0;
      out->inner_maps = dc.inner_maps;
    } else if (p.solver != COLLOP_SOLVER_ANDERSON) {
      *out = dc;
    } else {
      out->iters = an.k_done;
      out->iters_f32 = an.iters32;
      out->n_restart = an.n_restart;
      out->exit_reason = an.exit_reason;
      out->residual = an.res_out;
      out->r0 = an.nrm_r0;
      out->inner_maps = 0;
    }
    out->n_clusters = ctx->ncl;
    out->cl_fail = cl_fail;
    out->stat_syncs = (int32_t)(ctx->stat_syncs - syncs0);
    out->entropy = sc[0];
    out->energy = 0.5 * sc[5];
    out->momentum[0] = sc[3];
    out->momentum[1] = sc[4];
    out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return COLLOP_OK;
}

}  // namespace

extern "C" {

int collop_state_set(collop_ctx* ctx, const double* v, const double* w, int64_t N) {
  int r = need_mesh(ctx);
  if (r) return r;
  if ((r = check_N(ctx, N))) return r;
  if (N < 2) return fail(ctx, COLLOP_ERR_INVALID, "the implicit step needs N >= 2 markers");
  if (!v || !w) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  double wsum = 0.0, wmax = 0.0;
  bool same_w = true;
  for (int64_t i = 0; i < N; ++i) {
    if (!(w[i] > 0.0) || !std::isfinite(w[i]))
      return fail(ctx, COLLOP_ERR_INVALID, "weights must be positive and finite (G_eff divides by w)");
    wsum += w[i];
    wmax = std::max(wmax, w[i]);
    same_w = same_w && w[i] == w[0];
  }
  ctx->w_pow2 = same_w && equal_pow2_weights(w, 1);
  ctx->dc_skip = false;
  HIPC(hipSetDevice(ctx->device));
  if ((r = ensure_scratch(ctx, N))) return r;
  const size_t n2b = (size_t)2 * N * 8;
  for (DBuf* b : {&ctx->s_v, &ctx->s_v1, &ctx->s_Gv, &ctx->s_r, &ctx->s_rp, &ctx->s_Gp, &ctx->s_best, &ctx->s_dotv,
                  &ctx->s_Gvo, &ctx->s_Gk, &ctx->s_D, &ctx->s_Ge, &ctx->s_vmk, &ctx->s_xk, &ctx->s_oU, &ctx->s_oUp,
                  &ctx->s_of, &ctx->s_ofp})
    ENSURE(*b, n2b);
  ENSURE(ctx->s_FA, (size_t)5 * N * 8);
  ENSURE(ctx->s_odF, n2b * kMaxWindow);
  ENSURE(ctx->s_odG, n2b * kMaxWindow);
  ENSURE(ctx->b_near_list, (size_t)kNearCap * sizeof(int2));
  ENSURE(ctx->b_near_count, 64);
  ENSURE(ctx->b_ntg, (size_t)kNearCap * 4);
  ENSURE(ctx->b_noff, (size_t)(kNearCap + 1) * 4);
  ENSURE(ctx->b_nal, (size_t)kNearCap * 4);
  ENSURE(ctx->b_nidx, (size_t)N * 4);
  ENSURE(ctx->s_v1n, n2b);
  ENSURE(ctx->s_w, (size_t)N * 8);
  ENSURE(ctx->s_dF, n2b * kMaxWindow);
  ENSURE(ctx->s_dG, n2b * kMaxWindow);
  ENSURE(ctx->s_fc, (size_t)ctx->mv.n_dofs * 8);
  // stiff clusters: no cluster until a step finds them
  const int ncell = ctx->mv.ne1 * ctx->mv.ne2;
  ENSURE(ctx->b_cell_of, (size_t)N * 4);
  ENSURE(ctx->b_cell_sorted, (size_t)N * 4);
  ENSURE(ctx->b_cell_cnt, (size_t)ncell * 4);
  ENSURE(ctx->b_cell_start, (size_t)(ncell + 1) * 4);
  ENSURE(ctx->b_links, (size_t)kLinkCap * sizeof(NearLink));
  ENSURE(ctx->b_link_count, 64);
  ENSURE(ctx->b_cl_idx, (size_t)kClMembersCap * 4);
  ENSURE(ctx->b_cl_off, (size_t)(kClMembersCap / 2 + 1) * 4);
  ENSURE(ctx->b_cl_x, (size_t)2 * kClMembersCap * 8);
  ENSURE(ctx->b_cl_root, (size_t)2 * kClMembersCap * 8);
  ENSURE(ctx->b_cl_have, (size_t)(kClMembersCap / 2) * 4);
  ENSURE(ctx->b_cl_its, (size_t)(kClMembersCap / 2) * 4);
  ENSURE(ctx->b_cl_fail, 64);
  ENSURE(ctx->b_cl_row, (size_t)N * 4);
  launch_fill_int(P<int>(ctx->b_cl_row), (int)N, -1, ctx->stream);
  CHECK_LAUNCH();
  ctx->cl_row_n = N;
  ctx->ncl = ctx->cl_total = 0;
  ctx->N = N;
  ctx->w_abs_sum = wsum;
  ctx->w_max = wmax;
  choose_tile(ctx, (int)N);
  HIPC(hipMemcpyAsync(ctx->s_v.p, v, n2b, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemcpyAsync(ctx->s_w.p, w, (size_t)N * 8, hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemsetAsync(ctx->s_rp.p, 0, n2b, ctx->stream));
  HIPC(hipMemsetAsync(ctx->s_Gp.p, 0, n2b, ctx->stream));
  if ((r = project_state(ctx, 0))) return r;
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->has_state = true;
  return COLLOP_OK;
}

int collop_state_get(collop_ctx* ctx, double* v) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!ctx->has_state) return fail(ctx, COLLOP_ERR_NO_STATE, "no device-resident state (collop_state_set)");
  if (!v) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipMemcpyAsync(v, ctx->s_v.p, (size_t)2 * ctx->N * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

int collop_state_coeffs(collop_ctx* ctx, double* f_coeffs) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!ctx->has_state) return fail(ctx, COLLOP_ERR_NO_STATE, "no device-resident state (collop_state_set)");
  if (!f_coeffs) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipMemcpyAsync(f_coeffs, ctx->s_fc.p, (size_t)ctx->mv.n_dofs * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

int collop_state_diagnostics(collop_ctx* ctx, double* out) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!ctx->has_state) return fail(ctx, COLLOP_ERR_NO_STATE, "no device-resident state (collop_state_set)");
  if (!out) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  return diagnostics_dev(ctx, P<double>(ctx->s_v), P<double>(ctx->s_w), (int)ctx->N, ctx->w_abs_sum,
                         P<double>(ctx->s_fc), out);
}

int collop_run_steps(collop_ctx* ctx, const collop_step_params* params, int n_steps, collop_step_stats* stats) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!ctx->has_state) return fail(ctx, COLLOP_ERR_NO_STATE, "no device-resident state (collop_state_set)");
  int r = validate_params(ctx, params);
  if (r) return r;
  if (n_steps < 0) return fail(ctx, COLLOP_ERR_INVALID, "n_steps must be >= 0");
  HIPC(hipSetDevice(ctx->device));
  choose_tile(ctx, (int)ctx->N);
  // the entropy of the current state uses the same integrand variant as the steps
  if ((r = project_state(ctx, params->use_logsq))) return r;
  for (int s = 0; s < n_steps; ++s) {
    r = step_once(ctx, *params, stats ? &stats[s] : nullptr);
    if (r) return r;
  }
  return COLLOP_OK;
}

int collop_step(collop_ctx* ctx, const collop_step_params* params, double* v_inout, const double* w, int64_t N,
                collop_step_stats* stats) {
  int r = collop_state_set(ctx, v_inout, w, N);
  if (r) return r;
  if ((r = collop_run_steps(ctx, params, 1, stats))) return r;
  return collop_state_get(ctx, v_inout);
}

// ---- warm-start MLP ----------------------------------------------------------
int collop_mlp_set_params(collop_ctx* ctx, const float* params, int n) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!params || n != kMlpParams) return fail(ctx, COLLOP_ERR_INVALID, "expected 4738 MLP parameters");
  for (int i = 0; i < n; ++i)
    if (!std::isfinite(params[i])) return fail(ctx, COLLOP_ERR_INVALID, "non-finite MLP parameter");
  HIPC(hipSetDevice(ctx->device));
  ENSURE(ctx->m_params, kMlpParams * sizeof(float));
  ENSURE(ctx->m_adam_m, kMlpParams * sizeof(float));
  ENSURE(ctx->m_adam_v, kMlpParams * sizeof(float));
  HIPC(hipMemcpyAsync(ctx->m_params.p, params, kMlpParams * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  HIPC(hipMemsetAsync(ctx->m_adam_m.p, 0, kMlpParams * sizeof(float), ctx->stream));
  HIPC(hipMemsetAsync(ctx->m_adam_v.p, 0, kMlpParams * sizeof(float), ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  ctx->m_adam_step = 0;
  ctx->m_has_params = true;
  return COLLOP_OK;
}

int collop_mlp_get_params(collop_ctx* ctx, float* params, int n) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!ctx->m_has_params) return fail(ctx, COLLOP_ERR_NO_MODEL, "no MLP parameters on this context");
  if (!params || n != kMlpParams) return fail(ctx, COLLOP_ERR_INVALID, "expected 4738 MLP parameters");
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipMemcpyAsync(params, ctx->m_params.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

int collop_mlp_set_norm(collop_ctx* ctx, const float* mu, const float* sd, const float* out_scale) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!mu || !sd || !out_scale) return fail(ctx, COLLOP_ERR_INVALID, "null normalisation array");
  for (int j = 0; j < 6; ++j) {
    if (!(sd[j] > 0.0f) || !std::isfinite(mu[j])) return fail(ctx, COLLOP_ERR_INVALID, "bad feature normalisation");
    ctx->m_mu[j] = mu[j];
    ctx->m_isd[j] = 1.0f / sd[j];
  }
  for (int j = 0; j < 2; ++j) {
    if (!(out_scale[j] > 0.0f)) return fail(ctx, COLLOP_ERR_INVALID, "bad output scale");
    ctx->m_oscale[j] = out_scale[j];
  }
  ctx->m_has_norm = true;
  return COLLOP_OK;
}

int collop_mlp_dataset_alloc(collop_ctx* ctx, int64_t capacity) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (capacity < 0 || capacity > ((int64_t)1 << 31)) return fail(ctx, COLLOP_ERR_INVALID, "capacity out of range");
  HIPC(hipSetDevice(ctx->device));
  ENSURE(ctx->m_data, (size_t)std::max<int64_t>(capacity, 1) * 8 * sizeof(float));
  ctx->m_cap = capacity;
  ctx->m_size = 0;
  return COLLOP_OK;
}

int collop_mlp_dataset_size(collop_ctx* ctx, int64_t* n) {
  if (!ctx || !n) return COLLOP_ERR_INVALID;
  *n = ctx->m_size;
  return COLLOP_OK;
}

int collop_mlp_dataset_read(collop_ctx* ctx, int64_t first, int64_t count, float* out) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!out || first < 0 || count < 0 || first + count > ctx->m_size)
    return fail(ctx, COLLOP_ERR_INVALID, "dataset range out of bounds");
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipMemcpyAsync(out, P<float>(ctx->m_data) + (size_t)first * 8, (size_t)count * 8 * sizeof(float),
                      hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

static void mlp_args(collop_ctx* ctx, MlpArgs& a) {
  a.params = P<float>(ctx->m_params);
  for (int j = 0; j < 6; ++j) {
    a.mu[j] = ctx->m_mu[j];
    a.isd[j] = ctx->m_isd[j];
  }
  a.oscale[0] = ctx->m_oscale[0];
  a.oscale[1] = ctx->m_oscale[1];
}

// Sum of the per-workgroup loss partials (fixed order, host side).
static int mlp_pass(collop_ctx* ctx, MlpArgs& a, int grid, double* loss_sum) {
  ENSURE(ctx->m_lpart, (size_t)grid * sizeof(double));
  a.loss_part = P<double>(ctx->m_lpart);
  launch_mlp(a, grid, ctx->stream);
  CHECK_LAUNCH();
  if (loss_sum) {
    std::vector<double> lp(grid);
    HIPC(hipMemcpyAsync(lp.data(), a.loss_part, grid * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPC(hipStreamSynchronize(ctx->stream));
    double t = 0.0;
    for (double x : lp) t += x;
    *loss_sum = t;
  }
  return COLLOP_OK;
}

int collop_mlp_train(collop_ctx* ctx, int epochs, int batch, double lr, uint64_t seed, double* loss_per_epoch) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!(ctx->m_has_params && ctx->m_has_norm)) return fail(ctx, COLLOP_ERR_NO_MODEL, "set MLP parameters and norm first");
  if (epochs < 0 || batch < 16 || batch % 16 || !(lr > 0.0)) return fail(ctx, COLLOP_ERR_INVALID, "bad training knobs");
  const int64_t ntiles = ctx->m_size / 16;
  if (ntiles < 1) return fail(ctx, COLLOP_ERR_INVALID, "the training dataset is empty");
  HIPC(hipSetDevice(ctx->device));
  const int tpb = batch / 16;
  const int grid = std::min(tpb, 512);
  ENSURE(ctx->m_tiles, (size_t)ntiles * sizeof(int));
  ENSURE(ctx->m_gpart, (size_t)grid * kMlpParams * sizeof(float));
  std::vector<int> perm((size_t)ntiles);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;  // splitmix64 stream for the shuffles
  auto next = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  for (int e = 0; e < epochs; ++e) {
    for (int64_t i = 0; i < ntiles; ++i) perm[i] = (int)i;
    for (int64_t i = ntiles - 1; i > 0; --i) std::swap(perm[i], perm[next() % (uint64_t)(i + 1)]);
    HIPC(hipMemcpyAsync(ctx->m_tiles.p, perm.data(), (size_t)ntiles * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    for (int64_t b0 = 0; b0 < ntiles; b0 += tpb) {
      MlpArgs a{};
      mlp_args(ctx, a);
      a.rec = P<float>(ctx->m_data);
      a.nsamples = ctx->m_size;
      a.tiles = P<int>(ctx->m_tiles) + b0;
      a.ntiles = (int)std::min<int64_t>(tpb, ntiles - b0);
      a.inv_batch = 2.0f / (16.0f * a.ntiles);
      a.mode = 2;
      a.grad_part = P<float>(ctx->m_gpart);
      const int g = std::min(a.ntiles, grid);
      int r = mlp_pass(ctx, a, g, nullptr);
      if (r) return r;
      launch_adam(P<float>(ctx->m_gpart), g, P<float>(ctx->m_params), P<float>(ctx->m_adam_m),
                  P<float>(ctx->m_adam_v), nullptr, (float)lr, (int)++ctx->m_adam_step, 1, ctx->stream);
      CHECK_LAUNCH();
    }
    if (loss_per_epoch) {
      MlpArgs a{};
      mlp_args(ctx, a);
      a.rec = P<float>(ctx->m_data);
      a.nsamples = ctx->m_size;
      a.ntiles = (int)ntiles;
      a.mode = 1;
      double tot = 0.0;
      int r = mlp_pass(ctx, a, 1024, &tot);
      if (r) return r;
      loss_per_epoch[e] = tot / (double)(ntiles * 16);
    }
  }
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

int collop_mlp_eval(collop_ctx* ctx, const float* rec, int64_t n, int mode, float* y, double* loss, float* grad) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!(ctx->m_has_params && ctx->m_has_norm)) return fail(ctx, COLLOP_ERR_NO_MODEL, "set MLP parameters and norm first");
  if (!rec || n < 1 || n > ((int64_t)1 << 28) || mode < 0 || mode > 2 || (mode == 0 && !y) ||
      (mode >= 1 && !loss) || (mode == 2 && !grad))
    return fail(ctx, COLLOP_ERR_INVALID, "bad arguments");
  HIPC(hipSetDevice(ctx->device));
  ENSURE(ctx->m_eval, (size_t)n * 8 * sizeof(float));
  ENSURE(ctx->m_y, (size_t)n * 2 * sizeof(float));
  HIPC(hipMemcpyAsync(ctx->m_eval.p, rec, (size_t)n * 8 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  MlpArgs a{};
  mlp_args(ctx, a);
  a.rec = P<float>(ctx->m_eval);
  a.nsamples = n;
  a.ntiles = (int)((n + 15) / 16);
  a.inv_batch = 2.0f / (float)n;
  a.mode = mode;
  a.y = P<float>(ctx->m_y);
  const int grid = std::min(a.ntiles, 256);
  ENSURE(ctx->m_gpart, (size_t)grid * kMlpParams * sizeof(float));
  a.grad_part = P<float>(ctx->m_gpart);
  double tot = 0.0;
  int r = mlp_pass(ctx, a, grid, mode >= 1 ? &tot : nullptr);
  if (r) return r;
  if (mode == 0) HIPC(hipMemcpyAsync(y, ctx->m_y.p, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  if (mode >= 1) *loss = tot / (double)n;
  if (mode == 2) {
    ENSURE(ctx->m_grad, kMlpParams * sizeof(float));
    launch_adam(a.grad_part, grid, nullptr, nullptr, nullptr, P<float>(ctx->m_grad), 0.0f, 1, 0, ctx->stream);
    CHECK_LAUNCH();
    HIPC(hipMemcpyAsync(grad, ctx->m_grad.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

// ---- extras (not part of the reference's hook set) ----------------------
// Ordered pair-sum tile (selects the ordered sum): gamma_per_lane in {1,2},
// slices in [1, 64], newton in {1,2}; slices = 0 restores the automatic choice.
int collop_set_pair_tile(collop_ctx* ctx, int gamma_per_lane, int slices, int newton) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (gamma_per_lane != 1 && gamma_per_lane != 2) return fail(ctx, COLLOP_ERR_INVALID, "gamma_per_lane must be 1 or 2");
  if (newton != 1 && newton != 2) return fail(ctx, COLLOP_ERR_INVALID, "newton must be 1 or 2");
  if (slices < 0 || slices > 64) return fail(ctx, COLLOP_ERR_INVALID, "slices must be in [0, 64]");
  ctx->tile64.gamma_per_lane = ctx->tile32.gamma_per_lane = gamma_per_lane;
  ctx->tile64.newton = newton;
  ctx->tile_auto = slices == 0;
  ctx->manual_slices = slices;
  ctx->pair_sym = false;
  if (ctx->has_state) choose_tile(ctx, (int)ctx->N);
  return COLLOP_OK;
}

int collop_get_pair_mode(collop_ctx* ctx, int64_t N, int* mode, int* tiles, int* slots) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (N < 0 || N > (int64_t)1 << 27) return fail(ctx, COLLOP_ERR_INVALID, "N must be in [0, 2^27]");
  const int R = sym_tiles_for(ctx, N);
  if (mode) *mode = R ? COLLOP_PAIR_SYMMETRIC : COLLOP_PAIR_ORDERED;
  if (tiles) *tiles = R;
  if (slots) *slots = tile_for(ctx, ctx->tile64, (int)N).slices;
  return COLLOP_OK;
}

int collop_set_pair_mode(collop_ctx* ctx, int mode, int tiles) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (mode != COLLOP_PAIR_ORDERED && mode != COLLOP_PAIR_SYMMETRIC)
    return fail(ctx, COLLOP_ERR_INVALID, "mode must be COLLOP_PAIR_ORDERED or COLLOP_PAIR_SYMMETRIC");
  if (tiles != 0 && tiles != 2 && tiles != 4 && tiles != 8)
    return fail(ctx, COLLOP_ERR_INVALID, "tiles must be 0 (automatic), 2, 4 or 8");
  ctx->pair_sym = mode == COLLOP_PAIR_SYMMETRIC;
  ctx->sym_tiles = tiles;
  if (mode == COLLOP_PAIR_ORDERED) {  // the ordered sum with its automatic tile
    ctx->tile_auto = true;
    ctx->tile64.gamma_per_lane = ctx->tile32.gamma_per_lane = 1;
    ctx->tile64.newton = 1;
  }
  if (ctx->has_state) {
    int r;
    if ((r = ensure_scratch(ctx, ctx->N))) return r;
    choose_tile(ctx, (int)ctx->N);
  }
  return COLLOP_OK;
}

int collop_set_fused_maps(collop_ctx* ctx, int on) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (on != 0 && on != 1) return fail(ctx, COLLOP_ERR_INVALID, "on must be 0 or 1");
  ctx->fused_maps = on == 1;
  return COLLOP_OK;
}

int collop_set_weight_path(collop_ctx* ctx, int on) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (on != 0 && on != 1) return fail(ctx, COLLOP_ERR_INVALID, "on must be 0 or 1");
  ctx->pow2w_path = on == 1;
  return COLLOP_OK;
}

int collop_set_device_solve(collop_ctx* ctx, int on) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (on != 0 && on != 1) return fail(ctx, COLLOP_ERR_INVALID, "on must be 0 or 1");
  ctx->dev_solve = on == 1;
  return COLLOP_OK;
}

// HIP-event timing of every pair-kernel launch (kind 0 = FP64, 1 = FP32c,
// 2 = FP64 with the metric, 3 = FP32c with the metric).
int collop_profile(collop_ctx* ctx, int enable, double* ms_total, long* launches) {
  if (!ctx) return COLLOP_ERR_INVALID;
  for (int k = 0; k < 4; ++k) {
    if (ms_total) ms_total[k] = ctx->prof_ms[k];
    if (launches) launches[k] = ctx->prof_n[k];
    ctx->prof_ms[k] = 0.0;
    ctx->prof_n[k] = 0;
  }
  ctx->profile = enable != 0;
  ctx->ev_used = 0;
  return COLLOP_OK;
}

}  // extern "C"
