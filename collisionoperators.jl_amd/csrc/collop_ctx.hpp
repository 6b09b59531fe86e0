// collop_ctx.hpp -- the context behind the opaque collop_ctx handle and the
// host-side helpers shared by the C-ABI translation units (capi.hip,
// ensemble.hip).  Internal: not part of the public header.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/collop.h"
#include "collop_launch.hpp"

namespace collop {

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct EnsembleState;  // ensemble.hip
void ensemble_free(EnsembleState* e);

}  // namespace collop

struct collop_ctx {
  using DBuf = collop::DBuf;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // ---- mesh
  bool has_mesh = false;
  collop::MeshView mv{};
  std::vector<double> h_bp1, h_bp2;
  std::vector<DBuf*> mesh_bufs;
  // ---- scratch buffers by name (grow-only)
  DBuf b_v, b_w, b_G, b_out, b_vmid, b_dv, b_dS, b_s64, b_s32, b_part, b_spart_sc, b_mom, b_dpart, b_iota, b_spart_q,
      b_vec, b_scal, b_statp, b_stmp, b_lbpart;
  DBuf nk_V, nk_vec, nk_part, nk_h, nk_A;  // Newton-Krylov: basis, work vectors, dot partials, scalars, dense A
  // ---- device-resident state
  bool has_state = false;
  int64_t N = 0;
  double w_abs_sum = 0.0;  // sum |w| of the state (fixed-point scale of the scatter)
  double w_max = 0.0;      // max w of the state (near-pair prefilter)
  DBuf s_v, s_w, s_v1, s_Gv, s_r, s_rp, s_Gp, s_best, s_dF, s_dG, s_fc, s_dotv;
  DBuf s_Gvo, s_Gk, s_FA, s_D;  // defect correction: G(v_k), G_eff(v_k), [F, A], A (G_eff - G_k)
  DBuf s_Ge, s_vmk;             // G_eff and v_mid of the inner iterate / of v_k (near field)
  DBuf s_xk, s_oU, s_oUp, s_of, s_ofp, s_odF, s_odG;  // outer Anderson window
  DBuf b_near_list, b_near_count, b_ntg, b_noff, b_nal;  // near pairs (list, CSR by target)
  DBuf b_nidx;                  // near-target row of every marker (-1: none)
  DBuf s_v1n;                   // the inner iterate after a scatter-applied Anderson update
  int near_nt = 0;              // near targets of the current step
  // stiff clusters of the current step (cluster.hip): cell buckets, the near
  // links, the clusters in CSR, the member row of every marker, map values
  DBuf b_cell_of, b_cell_cnt, b_cell_start, b_cell_sorted, b_links, b_link_count;
  DBuf b_cl_off, b_cl_idx, b_cl_row, b_cl_x, b_cl_root, b_cl_have, b_cl_fail, b_cl_its;
  int ncl = 0, cl_total = 0;    // clusters and their members
  int64_t cl_row_n = 0;         // markers b_cl_row covers (all -1 outside the members)
  double* h_stat = nullptr;  // pinned, kStatWidth doubles
  double* h_stat_dev = nullptr;  // its device-side address: the statistics reduction writes there
  unsigned* h_seq = nullptr;      // pinned: sequence number published after each h_stat write
  unsigned* h_seq_dev = nullptr;  // its device-side address
  unsigned seq = 0;               // the last sequence number launched
  long stat_syncs = 0;            // statistics polls that ran out (a stream synchronisation instead)
  // ---- device-resident Anderson solve (control block AndCtl in b_actl)
  bool dev_solve = true;               // Landau Anderson steps keep their control on the device
  DBuf b_actl;
  std::vector<unsigned char> h_actl;   // host image of the control block (upload / read-back)
  // ---- warm-start MLP
  DBuf m_params, m_adam_m, m_adam_v, m_data, m_rec, m_y, m_gpart, m_lpart, m_tiles, m_eval, m_grad;
  bool m_has_params = false, m_has_norm = false;
  float m_mu[6] = {0}, m_isd[6] = {0}, m_oscale[2] = {1, 1};
  int64_t m_cap = 0, m_size = 0;
  long m_adam_step = 0;
  // ---- pair tile and profiling
  collop::PairTile tile64{1, 32, 1};
  collop::PairTile tile32{1, 32, 0};
  bool tile_auto = true;
  int manual_slices = 0;
  bool pair_sym = true;  // symmetric pair sum (each unordered pair once)
  int sym_tiles = 0;     // its waves per block (0: automatic)
  bool fused_maps = true;  // fused small-kernel maps where the split mass solves apply
  bool w_pow2 = false;     // the state's markers all carry one weight, a power of two
  bool dc_skip = false;    // hand-over memory (dc_sticky): the next defect-correction step runs Anderson
  bool pow2w_path = true;  // the symmetric sums take their equal-weight form where the weights allow
  bool profile = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
  size_t ev_used = 0;
  std::vector<int> ev_kind;
  std::vector<int> ev_map;  // map of the step each event pair times (device-resident solve)
  int prof_map = 0;         // map the next events belong to (0: not in a device-resident solve)
  double prof_ms[4] = {0.0, 0.0, 0.0, 0.0};  // FP64, FP32c, FP64 + metric, FP32c + metric
  long prof_n[4] = {0, 0, 0, 0};
  // ---- batched ensemble (ensemble.hip)
  collop::EnsembleState* ens = nullptr;
};

namespace collop {

inline int fail(collop_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPC(expr)                                                                                     \
  do {                                                                                                 \
    hipError_t _e = (expr);                                                                            \
    if (_e != hipSuccess)                                                                              \
      return fail(ctx, COLLOP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));             \
  } while (0)

#define CHECK_LAUNCH()                                                                                 \
  do {                                                                                                 \
    hipError_t _e = hipGetLastError();                                                                 \
    if (_e != hipSuccess) return fail(ctx, COLLOP_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(_e)); \
  } while (0)

inline int ensure(collop_ctx* ctx, DBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.bytes >= bytes) return COLLOP_OK;
  if (b.p) {
    hipError_t e = hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    if (e != hipSuccess) return fail(ctx, COLLOP_ERR_HIP, std::string("hipFree: ") + hipGetErrorString(e));
  }
  hipError_t e = hipMalloc(&b.p, bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return fail(ctx, COLLOP_ERR_HIP, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
  }
  b.bytes = bytes;
  return COLLOP_OK;
}

inline void release(DBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

template <class T>
T* P(DBuf& b) {
  return static_cast<T*>(b.p);
}

#define ENSURE(buf, bytes)                             \
  do {                                                 \
    int _r = ensure(ctx, (buf), (bytes));              \
    if (_r != COLLOP_OK) return _r;                    \
  } while (0)

}  // namespace collop
