// ensemble.hip -- a batched ensemble of independent Landau runs with uneven N,
// advanced together: every Anderson iteration is one replay of a captured
// hipGraph whose pair sum is a single persistent launch over all runs.
//
// Layout: the markers of all runs are concatenated (run r occupies
// [off_r, off_r + n_r), off_r a multiple of 256, so no 256-thread workgroup
// straddles two runs) in one SoA array v[2][NT].  Per-run quantities live in
// small per-run arrays; a per-run control block (RunCtl, uploaded before each
// replay) tells every kernel whether the run maps in this iteration, in which
// precision, which Anderson column it writes, which of the two ping-pong
// residual buffers is current, and which update (Anderson step, damped step,
// exit copy) to apply to the run's iterate first.  Kernel arguments therefore
// never change between replays.
//
// The pair kernel walks a static tile list (run, target block, source slice)
// sorted longest-first (LPT) with a grid-stride loop; the γ-tile size (threads
// per workgroup, one target per lane) and the number of resident waves per CU
// (the grid size) come from the autotuner (collop_ensemble_tune).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "collop_ctx.hpp"

namespace collop {

struct RunCtl {
  int active;       // the run maps in this iteration
  int kind;         // 0 FP64 maps, 1 FP32c maps
  int col;          // Anderson column written by the epilogue (-1: none)
  int hist;         // columns the epilogue dots with
  int cur;          // ping-pong buffer the map writes (r, Gv); the other is (rp, Gp)
  int upd;          // update before the map: 0 none, 1 Anderson/damped step, 2 v1 <- Gv, 3 v1 <- best
  int upd_hist;     // columns of the Anderson step (0: damped Picard step)
  int copy_best;    // before the update: best <- Gv (exit_picard_step) or v1
  int exit_picard;  // exit_picard_step of the run's parameters
  int pad;
  double beta;
  double gamma[kMaxWindow];
};

struct EnsView {
  int R, NT;
  const int* off;      // [R]
  const int* n;        // [R]
  const int* blk_run;  // [NT / 256]
  RunCtl* ctl;         // [R]
};

constexpr int kEnsBlock = 256;     // markers per element-wise workgroup
constexpr int kEnsScatter = 1024;  // markers per scatter workgroup (4 per thread)

namespace {

// ---- element-wise helpers ----------------------------------------------------
struct Where {
  int run, local;  // local index within the run, or -1 for padding
};

__device__ __forceinline__ Where where(const EnsView& e, int i) {
  const int r = e.blk_run[i / kEnsBlock];
  const int loc = i - e.off[r];
  return Where{r, loc < e.n[r] ? loc : -1};
}

// ---- scatter (v_mid / v_in or the state), fixed-point LDS rows ---------------
// grid.x = scatter blocks (sblk_run / sblk_first tables), grid.y = fields.
__global__ __launch_bounds__(kScatterThreads) void k_ens_scatter(MeshView m, EnsView e, const int* __restrict__ sblk_run,
                                                                 const int* __restrict__ sblk_first,
                                                                 const double* __restrict__ v0,
                                                                 const double* __restrict__ vin,
                                                                 const double* __restrict__ w,
                                                                 const double* __restrict__ scale, int map_mode,
                                                                 double* __restrict__ vmid, double* __restrict__ dv,
                                                                 long long* __restrict__ part,
                                                                 double* __restrict__ mom) {
  const int r = sblk_run[blockIdx.x];
  if (map_mode && !e.ctl[r].active) return;
  extern __shared__ unsigned long long lds_u[];
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kScatterThreads / 64];
  const int field = blockIdx.y;
  for (int k = threadIdx.x; k < m.n_dofs; k += blockDim.x) lds_u[k] = 0ull;
  stage_bp(m, sb1, sb2);
  __syncthreads();
  const int NT = e.NT, n = e.n[r], o = e.off[r];
  const double sc = scale[r];
  double p1 = 0.0, p2 = 0.0, e2 = 0.0;
  for (int loc = sblk_first[blockIdx.x] + threadIdx.x; loc < min(n, sblk_first[blockIdx.x] + kEnsScatter);
       loc += kScatterThreads) {
    const int i = o + loc;
    double x, y;
    if (map_mode) {
      const double a1 = v0[i], a2 = v0[NT + i], b1 = vin[i], b2 = vin[NT + i];
      const double m1 = 0.5 * (a1 + b1), m2 = 0.5 * (a2 + b2);
      if (field == 0) {
        vmid[i] = m1;
        vmid[NT + i] = m2;
        dv[i] = b1 - a1;
        dv[NT + i] = b2 - a2;
        x = m1;
        y = m2;
      } else {
        x = b1;
        y = b2;
      }
    } else {
      x = vin[i];
      y = vin[NT + i];
    }
    const double wa = w[i];
    p1 += wa * x;
    p2 += wa * y;
    e2 += wa * (x * x + y * y);
    Located L;
    if (!locate(m, sb1, sb2, x, y, L)) continue;
    double f1[3], f2[3];
    basis3(m.C1, L.i, L.xi1, f1);
    basis3(m.C2, L.j, L.xi2, f2);
    const int s1 = m.bs1[L.i], s2 = m.bs2[L.j];
    const double ws = wa * sc;
#pragma unroll
    for (int j2 = 0; j2 < 3; ++j2)
#pragma unroll
      for (int j1 = 0; j1 < 3; ++j1) {
        const long long q = __double2ll_rn(ws * f1[j1] * f2[j2]);
        atomicAdd(&lds_u[(s1 + j1) + (s2 + j2) * m.nd1], (unsigned long long)q);
      }
  }
  __syncthreads();
  long long* out = part + ((size_t)field * gridDim.x + blockIdx.x) * m.n_dofs;
  for (int k = threadIdx.x; k < m.n_dofs; k += blockDim.x) out[k] = (long long)lds_u[k];
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

// ---- per-run mass solve of the exact row sums (or of a direct vector) --------
// grid = (R, fields); out / direct at vec[(r * 4 + slot) * n_dofs].
__global__ __launch_bounds__(1024) void k_ens_solve(MeshView m, EnsView e, const int* __restrict__ srow0,
                                                    const int* __restrict__ srows, const long long* __restrict__ part,
                                                    int total_rows, const double* __restrict__ scale, int in_slot,
                                                    int out_slot, int map_mode, double* __restrict__ vec) {
  const int r = blockIdx.x, field = blockIdx.y;
  if (map_mode && !e.ctl[r].active) return;
  extern __shared__ double X[];
  const int nd = m.n_dofs;
  if (part) {
    const long long* P0 = part + ((size_t)field * total_rows + srow0[r]) * nd;
    const double inv = 1.0 / scale[r];
    for (int k = threadIdx.x; k < nd; k += blockDim.x) {
      unsigned long long s = 0ull;
      for (int b = 0; b < srows[r]; ++b) s += (unsigned long long)P0[(size_t)b * nd + k];
      X[k] = (double)(long long)s * inv;
    }
  } else {
    const double* src = vec + ((size_t)r * 4 + in_slot) * nd;
    for (int k = threadIdx.x; k < nd; k += blockDim.x) X[k] = src[k];
  }
  __syncthreads();
  for (int l = threadIdx.x; l < m.nd2; l += blockDim.x) band_solve_line(X + (size_t)l * m.nd1, m.nd1, 1, m.chol1);
  __syncthreads();
  for (int l = threadIdx.x; l < m.nd1; l += blockDim.x) band_solve_line(X + l, m.nd2, m.nd1, m.chol2);
  __syncthreads();
  double* dst = vec + ((size_t)r * 4 + out_slot + field) * nd;
  for (int k = threadIdx.x; k < nd; k += blockDim.x) dst[k] = X[k];
}

// ---- quadrature: grid (qblocks, fields, R); field f reads vec slot f ----------
constexpr int kQT = 256;
__global__ __launch_bounds__(kQT) void k_ens_quad1(MeshView m, EnsView e, const double* __restrict__ vec, int c_slot,
                                                   int use_logsq, int map_mode, double* __restrict__ iota,
                                                   double* __restrict__ spart) {
  __shared__ double red[kQT / 64];
  const int r = blockIdx.z, field = blockIdx.y;
  if (map_mode && !e.ctl[r].active) return;
  const double* c = vec + ((size_t)r * 4 + c_slot + field) * m.n_dofs;
  const int npt = m.nq * m.nq, nel = m.ne1 * m.ne2;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  double s_loc = 0.0;
  if (tid < nel * npt) {
    const int el = tid / npt, pt = tid - el * npt;
    const int e1 = el % m.ne1, e2 = el / m.ne1, a = pt % m.nq, b = pt / m.nq;
    const int s1 = m.bs1[e1], s2 = m.bs2[e2];
    double f = 0.0;
#pragma unroll
    for (int j2 = 0; j2 < 3; ++j2) {
      double row = 0.0;
#pragma unroll
      for (int j1 = 0; j1 < 3; ++j1) row += c[(s1 + j1) + (s2 + j2) * m.nd1] * m.phiq1[(3 * e1 + j1) * m.nq + a];
      f += row * m.phiq2[(3 * e2 + j2) * m.nq + b];
    }
    const double wq = m.qw[a] * m.qw[b] * (m.bp1[e1 + 1] - m.bp1[e1]) * (m.bp2[e2 + 1] - m.bp2[e2]);
    double lg = 0.0;
    bool ok;
    if (use_logsq) {
      ok = f * f > 1e-60;
      if (ok) lg = 0.5 * log(f * f);
    } else {
      ok = f > 1e-30;
      if (ok) lg = log(f);
    }
    if (iota && field == 0) iota[(size_t)r * nel * npt + (size_t)el * npt + pt] = ok ? (1.0 + lg) * wq : 0.0;
    s_loc = ok ? -f * lg * wq : 0.0;
  }
  if (spart) {
    const double s = block_sum<kQT>(s_loc, red);
    if (threadIdx.x == 0) spart[((size_t)r * 2 + field) * gridDim.x + blockIdx.x] = s;
  }
}

__global__ void k_ens_quad2(MeshView m, EnsView e, const double* __restrict__ iota, int map_mode,
                            double* __restrict__ vec) {
  const int r = blockIdx.y;
  if (map_mode && !e.ctl[r].active) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m.n_dofs) return;
  const int i1 = k % m.nd1, i2 = k / m.nd1, npt = m.nq * m.nq;
  const double* io0 = iota + (size_t)r * m.ne1 * m.ne2 * npt;
  double acc = 0.0;
  for (int e2 = m.erange2[2 * i2]; e2 <= m.erange2[2 * i2 + 1]; ++e2) {
    const double* p2 = m.phiq2 + (3 * e2 + i2 - m.bs2[e2]) * m.nq;
    for (int e1 = m.erange1[2 * i1]; e1 <= m.erange1[2 * i1 + 1]; ++e1) {
      const double* p1 = m.phiq1 + (3 * e1 + i1 - m.bs1[e1]) * m.nq;
      const double* io = io0 + (size_t)(e1 + e2 * m.ne1) * npt;
      for (int b = 0; b < m.nq; ++b) {
        double row = 0.0;
        for (int a = 0; a < m.nq; ++a) row += p1[a] * io[a + m.nq * b];
        acc += row * p2[b];
      }
    }
  }
  vec[((size_t)r * 4 + 2) * m.n_dofs + k] = acc;
}

// out[r * ostride + c] = sum_{j < cnt[r]} part[(row0[r] + j) * stride + c], c < ncomp
// (fixed order; one workgroup per (run, component)).
__global__ __launch_bounds__(256) void k_ens_rowsum(EnsView e, const double* __restrict__ part, const int* __restrict__ row0,
                                                    const int* __restrict__ cnt, int stride, int fixed_cnt,
                                                    int fixed_row0_mul, int map_mode, double* __restrict__ out,
                                                    int ostride) {
  __shared__ double red[4];
  const int r = blockIdx.x, c = blockIdx.y;
  if (map_mode && !e.ctl[r].active) return;
  const int n = cnt ? cnt[r] : fixed_cnt;
  const int r0 = row0 ? row0[r] : r * fixed_row0_mul;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) s += part[(size_t)(r0 + j) * stride + c];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) out[(size_t)r * ostride + c] = s;
}

// ---- gathers: mode 0 G, mode 1 dS = -w G + per-block (dv.dS, |dv|^2) ---------
__global__ __launch_bounds__(kEnsBlock) void k_ens_gather(MeshView m, EnsView e, const double* __restrict__ x,
                                                          const double* __restrict__ vec, int L_slot,
                                                          const double* __restrict__ w,
                                                          const double* __restrict__ dv, int mode, int map_mode,
                                                          double* __restrict__ out, double* __restrict__ dpart) {
  const Where wh = where(e, blockIdx.x * kEnsBlock + threadIdx.x);
  if (map_mode && !e.ctl[wh.run].active) return;
  __shared__ double sb1[kMaxBp], sb2[kMaxBp];
  __shared__ double red[kEnsBlock / 64];
  stage_bp(m, sb1, sb2);
  __syncthreads();
  const int NT = e.NT, i = blockIdx.x * kEnsBlock + threadIdx.x;
  const double* coeff = vec + ((size_t)wh.run * 4 + L_slot) * m.n_dofs;
  double dot = 0.0, nrm = 0.0;
  if (wh.local >= 0) {
    double o1 = 0.0, o2 = 0.0;
    Located L;
    if (locate(m, sb1, sb2, x[i], x[NT + i], L)) {
      double f1[3], d1[3], f2[3], d2[3];
      basis3_grad(m.C1, L.i, L.xi1, f1, d1);
      basis3_grad(m.C2, L.j, L.xi2, f2, d2);
      const double ih1 = 1.0 / L.h1, ih2 = 1.0 / L.h2;
      const int s1 = m.bs1[L.i], s2 = m.bs2[L.j];
#pragma unroll
      for (int j2 = 0; j2 < 3; ++j2)
#pragma unroll
        for (int j1 = 0; j1 < 3; ++j1) {
          const double c = coeff[(s1 + j1) + (s2 + j2) * m.nd1];
          o1 += c * (d1[j1] * f2[j2]) * ih1;
          o2 += c * (f1[j1] * d2[j2]) * ih2;
        }
    }
    if (mode == 1) {
      const double wa = w[i];
      o1 = -wa * o1;
      o2 = -wa * o2;
      const double e1 = dv[i], e2 = dv[NT + i];
      dot = e1 * o1 + e2 * o2;
      nrm = e1 * e1 + e2 * e2;
    }
    out[i] = o1;
    out[NT + i] = o2;
  }
  if (mode == 1) {
    const double s1 = block_sum<kEnsBlock>(dot, red);
    const double s2 = block_sum<kEnsBlock>(nrm, red);
    if (threadIdx.x == 0) {
      dpart[2 * blockIdx.x] = s1;
      dpart[2 * blockIdx.x + 1] = s2;
    }
  }
}

// ---- G_eff + pair-source pack (map) or plain pack of (v, G) (predictor) -------
// scal[r * 8 + {0: S0, 1: S1, 2: lambda}]
__global__ __launch_bounds__(kEnsBlock) void k_ens_pack(MeshView m, EnsView e, const double* __restrict__ x,
                                                        const double* __restrict__ G, const double* __restrict__ dv,
                                                        const double* __restrict__ w, const double* __restrict__ dpart,
                                                        const int* __restrict__ eblk0, const int* __restrict__ eblks,
                                                        const double* __restrict__ scal, int use_gonzalez,
                                                        int map_mode, Src64* __restrict__ s64,
                                                        Src32* __restrict__ s32) {
  const Where wh = where(e, blockIdx.x * kEnsBlock + threadIdx.x);
  const RunCtl& c = e.ctl[wh.run];
  if (map_mode && !c.active) return;
  __shared__ double lam_sh;
  if (threadIdx.x == 0) {
    double lam = 0.0;
    if (map_mode && use_gonzalez) {
      double dot = 0.0, nrm2 = 0.0;
      const int b0 = eblk0[wh.run];
      for (int b = 0; b < eblks[wh.run]; ++b) {
        dot += dpart[2 * (b0 + b)];
        nrm2 += dpart[2 * (b0 + b) + 1];
      }
      lam = nrm2 > 1e-30 ? (scal[wh.run * 8 + 1] - scal[wh.run * 8 + 0] - dot) / nrm2 : 0.0;
    }
    lam_sh = lam;
  }
  __syncthreads();
  if (wh.local < 0) return;
  const int NT = e.NT, i = blockIdx.x * kEnsBlock + threadIdx.x;
  double g1 = G[i], g2 = G[NT + i];
  if (map_mode) {  // G holds dS here: G_eff = -(dS + lambda dv) / w  (solver.jl:107-111)
    const double iw = 1.0 / w[i];
    g1 = -(g1 + lam_sh * dv[i]) * iw;
    g2 = -(g2 + lam_sh * dv[NT + i]) * iw;
  }
  const double a = x[i], b = x[NT + i];
  const bool in = in_box(a, b, m.lo1, m.hi1, m.lo2, m.hi2);
  const double wm = in ? w[i] : 0.0;
  if (!map_mode || c.kind == 0) {
    Src64 s;
    s.v1 = a;
    s.v2 = b;
    s.g1 = g1;
    s.g2 = g2;
    s.w = wm;
    s.in_domain = in ? 1.0 : 0.0;
    s.pad0 = s.pad1 = 0.0;
    s64[i] = s;
  } else {
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

// ---- persistent pair sum over the LPT tile list ------------------------------
// tile t: run tr[t], first target tg[t] (local), slice ts[t] of S_r = slices[r].
// part layout per run: part[poff[r] + s * 2 n_r + comp * n_r + local].
template <int BT>
__global__ __launch_bounds__(BT) void k_ens_pair(EnsView e, const int* __restrict__ tr, const int* __restrict__ tg,
                                                 const int* __restrict__ ts, int ntiles,
                                                 const int* __restrict__ slices, const long long* __restrict__ poff,
                                                 const Src64* __restrict__ s64, const Src32* __restrict__ s32,
                                                 int force_f64, double* __restrict__ part) {
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int r = tr[t];
    const RunCtl& c = e.ctl[r];
    if (!force_f64 && !c.active) continue;
    const int n = e.n[r], o = e.off[r], S = slices[r];
    const int chunk = (n + S - 1) / S;
    const int s = ts[t];
    const int a0 = o + s * chunk, a1 = o + min(n, (s + 1) * chunk);
    const int loc = tg[t] + (int)threadIdx.x;
    const int g = o + loc;
    double* pr = part + poff[r] + (size_t)s * 2 * n;
    if (force_f64 || c.kind == 0) {
      double vg1 = 0, vg2 = 0, Gg1 = 0, Gg2 = 0, acc1 = 0, acc2 = 0;
      bool act = false;
      if (loc < n) {
        const Src64 q = s64[g];
        vg1 = q.v1;
        vg2 = q.v2;
        Gg1 = q.g1;
        Gg2 = q.g2;
        act = q.in_domain != 0.0;
      }
#pragma unroll 2
      for (int a = a0; a < a1; ++a) {
        double d1, d2, tt;
        pair_term_f64<1>(vg1, vg2, Gg1, Gg2, s64[a].v1, s64[a].v2, s64[a].g1, s64[a].g2, s64[a].w, acc1, acc2, d1,
                         d2, tt);
      }
      if (loc < n) {
        pr[loc] = act ? acc1 : 0.0;
        pr[n + loc] = act ? acc2 : 0.0;
      }
    } else {
      float vg1 = 0, vg2 = 0, Gg1 = 0, Gg2 = 0, b1 = 0, b2 = 0, h1 = 0, h2 = 0, l1 = 0, l2 = 0;
      bool act = false;
      if (loc < n) {
        const Src32 q = s32[g];
        vg1 = q.v1;
        vg2 = q.v2;
        Gg1 = q.g1;
        Gg2 = q.g2;
        act = q.in_domain != 0.0f;
      }
      for (int ab = a0; ab < a1; ab += 32) {
        const int ae = min(a1, ab + 32);
#pragma unroll 4
        for (int a = ab; a < ae; ++a)
          pair_term_f32(vg1, vg2, Gg1, Gg2, s32[a].v1, s32[a].v2, s32[a].g1, s32[a].g2, s32[a].w, b1, b2);
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

// ---- map epilogue (solver.jl:225, 339, 391-399) per run ------------------------
struct EnsBufs {
  double* v;        // state v0 [2][NT]
  double* v1;       // iterate
  double* R[2];     // residual ping-pong
  double* GV[2];    // map value ping-pong
  double* dF;       // [kMaxWindow][2][NT]
  double* dG;
  double* best;
  double* statp;    // [eblocks][kStatWidth]
};

__global__ __launch_bounds__(kEnsBlock) void k_ens_epilogue(EnsView e, EnsBufs B, const double* __restrict__ part,
                                                            const long long* __restrict__ poff,
                                                            const int* __restrict__ slices, double dt) {
  __shared__ double red[kEnsBlock / 64];
  const Where wh = where(e, blockIdx.x * kEnsBlock + threadIdx.x);
  const RunCtl& c = e.ctl[wh.run];
  if (!c.active) return;
  const int NT = e.NT, i = blockIdx.x * kEnsBlock + threadIdx.x, n = e.n[wh.run];
  double st[kStatWidth];
#pragma unroll
  for (int j = 0; j < kStatWidth; ++j) st[j] = 0.0;
  if (wh.local >= 0) {
    const double* pr = part + poff[wh.run];
    const int S = slices[wh.run];
    double* R = B.R[c.cur];
    double* GV = B.GV[c.cur];
    const double* RP = B.R[1 - c.cur];
    const double* GP = B.GV[1 - c.cur];
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      double s = 0.0;
      for (int k = 0; k < S; ++k) s += pr[(size_t)k * 2 * n + comp * n + wh.local];
      const int idx = comp * NT + i;
      const double gv = B.v[idx] + dt * s;
      GV[idx] = gv;
      const double x1 = B.v1[idx];
      const double ri = gv - x1;
      R[idx] = ri;
      st[0] += ri * ri;
      st[1] += x1 * x1;
      st[2] += (gv - B.v[idx]) * (gv - B.v[idx]);
      double f = 0.0;
      if (c.col >= 0) {
        f = ri - RP[idx];
        B.dF[(size_t)c.col * 2 * NT + idx] = f;
        B.dG[(size_t)c.col * 2 * NT + idx] = gv - GP[idx];
      }
#pragma unroll
      for (int j = 0; j < kMaxWindow; ++j) {
        if (j < c.hist) {
          const double xj = (j == c.col) ? f : B.dF[(size_t)j * 2 * NT + idx];
          if (c.col >= 0) st[3 + j] += xj * f;
          st[3 + kMaxWindow + j] += xj * ri;
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kStatWidth; ++j) {
    const double t = block_sum<kEnsBlock>(st[j], red);
    if (threadIdx.x == 0) B.statp[(size_t)blockIdx.x * kStatWidth + j] = t;
  }
}

// ---- per-run update of the iterate before the map ----------------------------
__global__ void k_ens_update(EnsView e, EnsBufs B, int all_to_pending) {
  const Where wh = where(e, blockIdx.x * kEnsBlock + threadIdx.x);
  if (wh.local < 0) return;
  const RunCtl& c = e.ctl[wh.run];
  if (c.upd == 0 && !c.copy_best) return;
  const int NT = e.NT, i = blockIdx.x * kEnsBlock + threadIdx.x;
  const int prev = 1 - c.cur;  // buffers of the previous map
  const double* GV = B.GV[prev];
  (void)all_to_pending;
#pragma unroll
  for (int comp = 0; comp < 2; ++comp) {
    const int idx = comp * NT + i;
    const double gv = GV[idx];
    const double x = B.v1[idx];
    if (c.copy_best) B.best[idx] = c.exit_picard ? gv : x;
    double nx = x;
    if (c.upd == 1) {
      double y = gv;
      for (int j = 0; j < c.upd_hist; ++j) y -= B.dG[(size_t)j * 2 * NT + idx] * c.gamma[j];
      nx = c.beta * y + (1.0 - c.beta) * x;
    } else if (c.upd == 2) {
      nx = gv;
    } else if (c.upd == 3) {
      nx = c.copy_best ? (c.exit_picard ? gv : x) : B.best[idx];
    }
    B.v1[idx] = nx;
  }
}

// ---- predictor: v1 = v + dt * (slice sum), keeps dot_v ----------------------
__global__ void k_ens_euler(EnsView e, const double* __restrict__ part, const long long* __restrict__ poff,
                            const int* __restrict__ slices, const double* __restrict__ v, double dt,
                            double* __restrict__ v1, double* __restrict__ dotv) {
  const Where wh = where(e, blockIdx.x * kEnsBlock + threadIdx.x);
  if (wh.local < 0) return;
  const int NT = e.NT, i = blockIdx.x * kEnsBlock + threadIdx.x, n = e.n[wh.run], S = slices[wh.run];
  const double* pr = part + poff[wh.run];
#pragma unroll
  for (int comp = 0; comp < 2; ++comp) {
    double s = 0.0;
    for (int k = 0; k < S; ++k) s += pr[(size_t)k * 2 * n + comp * n + wh.local];
    dotv[comp * NT + i] = s;
    v1[comp * NT + i] = v[comp * NT + i] + dt * s;
  }
}

}  // namespace
}  // namespace collop

namespace collop {
struct EnsembleState {
  int R = 0;
};
void ensemble_free(EnsembleState* e) { delete e; }
}  // namespace collop
