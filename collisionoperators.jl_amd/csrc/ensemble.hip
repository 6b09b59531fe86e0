// ensemble.hip -- a batched ensemble of independent Landau runs with uneven N
// on one mesh, advanced together (collop_ensemble_*).
//
// Layout in HBM: the runs' markers are concatenated, run r at [off_r, off_r + n_r)
// with off_r a multiple of 256, in SoA arrays whose component stride is NT
// (v1 of every run, then v2 of every run).  Per-run tables (EnsRun) give each
// batched kernel the block count the single-run kernel would use at N = n_r,
// so each run's arithmetic is the single-run arithmetic.
//
// One Anderson iteration of the whole ensemble: upload the per-run control
// blocks (RunCtl), then one replay of a captured hipGraph (per-run update of the
// iterate -> scatter / solve / quadrature / gather / G_eff of every active run ->
// ONE persistent pair-sum launch over the tiles of every active run (FP32c or
// FP64 per run) -> epilogue -> per-run statistics), then download.  The host then
// advances each run's Anderson state machine (the same one the single-run step
// uses) and writes the next control blocks.  Kernel arguments never change
// between replays; everything that varies lives in the control blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/collop.h"
#include "anderson.hpp"
#include "collop_ctx.hpp"

namespace collop {

constexpr int kEnsBlock = 256;
constexpr int kRowsMomR0 = 0, kRowsMomN = 1, kRowsStatR0 = 2, kRowsStatN = 3, kRowsQ0 = 4, kRowsQ1 = 5, kRowsQN = 6;
constexpr int kRowsKinds = 7;

struct EnsembleState {
  int R = 0, NT = 0;
  std::vector<EnsRun> runs;
  int n_sblk = 0, n_gblk = 0, n_pblk = 0, n_eblk = 0;
  long long part_len = 0;
  DBuf d_runs, d_ctl, d_tab, d_rows;
  DBuf v, w, v1, rr[2], gv[2], dF, dG, best, vmid, dv, dS, G, s64, s32, part, spart_sc, mom, dpart, iota, spart_q,
      vec, fc, scal, statp, stat, tiles;
  RunCtl* h_ctl = nullptr;   // pinned [R]
  double* h_stat = nullptr;  // pinned [R][kStatWidth]
  double* h_scal = nullptr;  // pinned [R][8]
  int ntiles = 0, bt = 256, waves = 8, n_cu = 256;
  int sym = 0;  // symmetric pair sum with sym waves per block (0: ordered tiles)
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double g_dt = 0.0;
  int g_gonz = -1, g_logsq = -1;
  std::vector<AndersonRun> an;
};

void ensemble_free(EnsembleState* e) {
  if (!e) return;
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  for (DBuf* b : {&e->d_runs, &e->d_ctl, &e->d_tab, &e->d_rows, &e->v, &e->w, &e->v1, &e->rr[0], &e->rr[1],
                  &e->gv[0], &e->gv[1], &e->dF, &e->dG, &e->best, &e->vmid, &e->dv, &e->dS, &e->G, &e->s64,
                  &e->s32, &e->part, &e->spart_sc, &e->mom, &e->dpart, &e->iota, &e->spart_q, &e->vec, &e->fc,
                  &e->scal, &e->statp, &e->stat, &e->tiles})
    release(*b);
  if (e->h_ctl) (void)hipHostFree(e->h_ctl);
  if (e->h_stat) (void)hipHostFree(e->h_stat);
  if (e->h_scal) (void)hipHostFree(e->h_scal);
  delete e;
}

namespace {

int slices_for(int64_t n) {
  // about 2048 sources per slice (the single-run choice at N = 65536)
  const int64_t s = (n + 2047) / 2048;
  return (int)std::max<int64_t>(1, std::min<int64_t>(64, s));
}

EnsView view(EnsembleState* E, bool with_ctl) {
  EnsView v{};
  v.R = E->R;
  v.NT = E->NT;
  v.n_sblk = E->n_sblk;
  v.n_gblk = E->n_gblk;
  v.n_pblk = E->n_pblk;
  v.n_eblk = E->n_eblk;
  v.runs = P<EnsRun>(E->d_runs);
  v.ctl = with_ctl ? P<RunCtl>(E->d_ctl) : nullptr;
  const int2* tab = P<int2>(E->d_tab);
  v.sblk = tab;
  v.gblk = tab + E->n_sblk;
  v.pblk = tab + E->n_sblk + E->n_gblk;
  v.eblk = tab + E->n_sblk + E->n_gblk + E->n_pblk;
  return v;
}

EnsBufs bufs(EnsembleState* E) {
  EnsBufs b{};
  b.v0 = P<double>(E->v);
  b.v1 = P<double>(E->v1);
  b.R[0] = P<double>(E->rr[0]);
  b.R[1] = P<double>(E->rr[1]);
  b.GV[0] = P<double>(E->gv[0]);
  b.GV[1] = P<double>(E->gv[1]);
  b.dF = P<double>(E->dF);
  b.dG = P<double>(E->dG);
  b.best = P<double>(E->best);
  b.part = P<double>(E->part);
  b.statp = P<double>(E->statp);
  return b;
}

const int* rows(EnsembleState* E, int kind) { return P<int>(E->d_rows) + (size_t)kind * E->R; }

// Tile list of the persistent pair sum.  Ordered form: runs in decreasing
// slice length (longest-processing-time order), each run's tiles in (target
// block, slice) order; all tiles of a run cost the same (bt targets x one
// slice of sources).  Symmetric form: the off-diagonal blocks (run, a, b) of
// every run, largest runs first, then the diagonal blocks; every block but a
// run's last super-row costs the same.
int build_tiles(collop_ctx* ctx, EnsembleState* E, int bt) {
  std::vector<int> order(E->R);
  std::iota(order.begin(), order.end(), 0);
  if (E->sym) {
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return E->runs[a].n > E->runs[b].n; });
    std::vector<int4> t;
    for (int r : order) {
      const int S = sym_rows(E->runs[r].n, E->sym);
      for (int b = 1; b < S; ++b)
        for (int a = 0; a < b; ++a) t.push_back(make_int4(r, a, b, 0));
    }
    for (int r : order)
      for (int a = 0; a < sym_rows(E->runs[r].n, E->sym); ++a) t.push_back(make_int4(r, a, a, 0));
    ENSURE(E->tiles, t.size() * sizeof(int4));
    HIPC(hipMemcpy(E->tiles.p, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
    E->ntiles = (int)t.size();
    E->bt = bt;
    return COLLOP_OK;
  }
  auto chunk = [&](int r) { return (E->runs[r].n + E->runs[r].S - 1) / E->runs[r].S; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return chunk(a) > chunk(b); });
  std::vector<int4> t;
  for (int r : order) {
    const int nb = (E->runs[r].n + bt - 1) / bt;
    for (int g = 0; g < nb; ++g)
      for (int s = 0; s < E->runs[r].S; ++s) t.push_back(make_int4(r, g * bt, s, 0));
  }
  ENSURE(E->tiles, t.size() * sizeof(int4));
  HIPC(hipMemcpy(E->tiles.p, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
  E->ntiles = (int)t.size();
  E->bt = bt;
  return COLLOP_OK;
}

EnsTiles tiles(EnsembleState* E) {
  EnsTiles T;
  T.tiles = P<int4>(E->tiles);
  T.ntiles = E->ntiles;
  T.bt = E->bt;
  T.sym = E->sym;
  T.grid = std::max(1, E->waves * E->n_cu / (E->sym ? E->sym : E->bt / 64));
  return T;
}

void drop_graph(EnsembleState* E) {
  if (E->exec) (void)hipGraphExecDestroy(E->exec);
  if (E->graph) (void)hipGraphDestroy(E->graph);
  E->exec = nullptr;
  E->graph = nullptr;
}

// Symmetric form with R waves per block (super-rows of 128 R markers): every
// run's slot count and partial offset, the partial buffer and the block list
// follow R.
int set_sym_form(collop_ctx* ctx, EnsembleState* E, int R) {
  for (const EnsRun& q : E->runs)
    if (sym_rows(q.n, R) > kSymMaxSlots) return fail(ctx, COLLOP_ERR_INVALID, "too many partial slots for this tile");
  E->sym = R;
  long long poff = 0;
  for (EnsRun& q : E->runs) {
    q.S = sym_rows(q.n, R);
    q.poff = poff;
    poff += (long long)q.S * 2 * q.n;
  }
  E->part_len = poff;
  ENSURE(E->part, (size_t)poff * 8);
  HIPC(hipMemcpy(E->d_runs.p, E->runs.data(), E->runs.size() * sizeof(EnsRun), hipMemcpyHostToDevice));
  drop_graph(E);
  return build_tiles(ctx, E, E->bt);
}

// Projection of every run's state markers, its entropy and moments
// (main.jl:298-303): fc[r], scal[r] = {S0, -, -, P1, P2, sum w|v|^2}.
int ens_project(collop_ctx* ctx, EnsembleState* E, int use_logsq) {
  const MeshView& m = ctx->mv;
  const EnsView V = view(E, false);
  hipStream_t st = ctx->stream;
  double* scal = P<double>(E->scal);
  launch_scatter_b(m, V, nullptr, P<double>(E->v), P<double>(E->w), 1, 1, nullptr, nullptr, P<long long>(E->spart_sc),
                   P<double>(E->mom), st);
  CHECK_LAUNCH();
  launch_solve_b(m, V, P<long long>(E->spart_sc), 1, 0, -1, P<double>(E->vec), P<double>(E->fc), st);
  CHECK_LAUNCH();
  launch_quad1_b(m, V, P<double>(E->fc), m.n_dofs, 1, use_logsq, nullptr, P<double>(E->spart_q), st);
  CHECK_LAUNCH();
  launch_sum_rows_b(V, P<double>(E->spart_q), rows(E, kRowsQ0), rows(E, kRowsQN), 1, 0, 1, scal + 0, 8, st);
  CHECK_LAUNCH();
  launch_sum_rows_b(V, P<double>(E->mom), rows(E, kRowsMomR0), rows(E, kRowsMomN), 3, 1, 3, scal + 3, 8, st);
  CHECK_LAUNCH();
  return COLLOP_OK;
}

// Euler predictor of every run (main.jl:239-245): v1 = v + dt vdot(v, G) with
// G = grad(M^{-1} r(f_s)); best <- v1.
int ens_predict(collop_ctx* ctx, EnsembleState* E, const collop_step_params& p) {
  const MeshView& m = ctx->mv;
  const EnsView V = view(E, false);
  hipStream_t st = ctx->stream;
  launch_quad1_b(m, V, P<double>(E->fc), m.n_dofs, 1, p.use_logsq, P<double>(E->iota), nullptr, st);
  CHECK_LAUNCH();
  launch_quad2_b(m, V, P<double>(E->iota), P<double>(E->vec), st);
  CHECK_LAUNCH();
  launch_solve_b(m, V, nullptr, 1, 2, 3, P<double>(E->vec), nullptr, st);
  CHECK_LAUNCH();
  launch_gather_b(m, V, P<double>(E->v), P<double>(E->vec), 3, nullptr, nullptr, 0, P<double>(E->G), nullptr, st);
  CHECK_LAUNCH();
  if (p.precision == COLLOP_PREC_F32C) {  // as the single-run step: an FP32c predictor
    for (int i = 0; i < E->R; ++i) {
      std::memset(&E->h_ctl[i], 0, sizeof(RunCtl));
      E->h_ctl[i].active = 1;
      E->h_ctl[i].kind = 1;
    }
    HIPC(hipMemcpyAsync(E->d_ctl.p, E->h_ctl, (size_t)E->R * sizeof(RunCtl), hipMemcpyHostToDevice, st));
    // the host rewrites h_ctl right after this function: the upload must have read it
    HIPC(hipStreamSynchronize(st));
    launch_pack_b(V, P<double>(E->v), P<double>(E->G), P<double>(E->w), m.lo1, m.hi1, m.lo2, m.hi2, nullptr,
                  P<Src32>(E->s32), st);
    CHECK_LAUNCH();
    launch_pair_b(view(E, true), tiles(E), nullptr, P<Src32>(E->s32), 0, P<double>(E->part), st);
    CHECK_LAUNCH();
  } else {
    launch_pack_b(V, P<double>(E->v), P<double>(E->G), P<double>(E->w), m.lo1, m.hi1, m.lo2, m.hi2, P<Src64>(E->s64),
                  nullptr, st);
    CHECK_LAUNCH();
    launch_pair_b(V, tiles(E), P<Src64>(E->s64), nullptr, 1, P<double>(E->part), st);
    CHECK_LAUNCH();
  }
  launch_euler_b(V, P<double>(E->part), P<double>(E->v), p.dt, P<double>(E->v1), nullptr, st);
  CHECK_LAUNCH();
  launch_copy_b(V, P<double>(E->v1), P<double>(E->best), st);
  CHECK_LAUNCH();
  return COLLOP_OK;
}

// One ensemble iteration: control upload, per-run update, Picard map of every
// active run (solver.jl:176-227) with its epilogue, statistics download.
int ens_iteration(collop_ctx* ctx, EnsembleState* E, const collop_step_params& p) {
  const MeshView& m = ctx->mv;
  const EnsView V = view(E, true);
  const EnsBufs B = bufs(E);
  hipStream_t st = ctx->stream;
  const int gf = p.use_gonzalez ? 2 : 1;
  double* vec = P<double>(E->vec);
  double* scal = P<double>(E->scal);
  launch_update_b(V, B, st);
  CHECK_LAUNCH();
  launch_scatter_b(m, V, P<double>(E->v), P<double>(E->v1), P<double>(E->w), 2, gf, P<double>(E->vmid),
                   P<double>(E->dv), P<long long>(E->spart_sc), nullptr, st);
  CHECK_LAUNCH();
  launch_solve_b(m, V, P<long long>(E->spart_sc), gf, 0, 0, vec, nullptr, st);
  CHECK_LAUNCH();
  launch_quad1_b(m, V, vec, 4 * m.n_dofs, gf, p.use_logsq, P<double>(E->iota), P<double>(E->spart_q), st);
  CHECK_LAUNCH();
  launch_quad2_b(m, V, P<double>(E->iota), vec, st);
  CHECK_LAUNCH();
  launch_solve_b(m, V, nullptr, 1, 2, 3, vec, nullptr, st);
  CHECK_LAUNCH();
  if (p.use_gonzalez) {
    launch_sum_rows_b(V, P<double>(E->spart_q), rows(E, kRowsQ1), rows(E, kRowsQN), 1, 0, 1, scal + 1, 8, st);
    CHECK_LAUNCH();
  }
  launch_gather_b(m, V, P<double>(E->vmid), vec, 3, P<double>(E->w), P<double>(E->dv), 1, P<double>(E->dS),
                  P<double>(E->dpart), st);
  CHECK_LAUNCH();
  launch_geff_pack_b(V, P<double>(E->vmid), P<double>(E->dS), P<double>(E->dv), P<double>(E->w), P<double>(E->dpart),
                     scal, p.use_gonzalez, m.lo1, m.hi1, m.lo2, m.hi2, P<Src64>(E->s64), P<Src32>(E->s32), st);
  CHECK_LAUNCH();
  launch_pair_b(V, tiles(E), P<Src64>(E->s64), P<Src32>(E->s32), 0, P<double>(E->part), st);
  CHECK_LAUNCH();
  launch_epilogue_b(V, B, p.dt, st);
  CHECK_LAUNCH();
  launch_sum_rows_b(V, P<double>(E->statp), rows(E, kRowsStatR0), rows(E, kRowsStatN), kStatWidth, 1, kStatWidth,
                    P<double>(E->stat), kStatWidth, st);
  CHECK_LAUNCH();
  return COLLOP_OK;
}

// One ensemble iteration: the control blocks go up, the iteration's kernels run
// (captured once in a hipGraph and replayed, or launched one by one), the
// statistics come down.  The graph holds kernel nodes only: the two copies
// stay outside it, on the same stream.
int ens_replay(collop_ctx* ctx, EnsembleState* E, const collop_step_params& p) {
  HIPC(hipMemcpyAsync(E->d_ctl.p, E->h_ctl, (size_t)E->R * sizeof(RunCtl), hipMemcpyHostToDevice, ctx->stream));
  if (!E->use_graph) {
    int r = ens_iteration(ctx, E, p);
    if (r) return r;
  } else {
    if (E->exec && (E->g_dt != p.dt || E->g_gonz != p.use_gonzalez || E->g_logsq != p.use_logsq)) drop_graph(E);
    if (!E->exec) {
      HIPC(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      const int r = ens_iteration(ctx, E, p);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
      if (r) {
        if (g) (void)hipGraphDestroy(g);
        return r;
      }
      if (ec != hipSuccess)
        return fail(ctx, COLLOP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
      E->graph = g;
      HIPC(hipGraphInstantiate(&E->exec, E->graph, nullptr, nullptr, 0));
      E->g_dt = p.dt;
      E->g_gonz = p.use_gonzalez;
      E->g_logsq = p.use_logsq;
    }
    HIPC(hipGraphLaunch(E->exec, ctx->stream));
  }
  HIPC(hipMemcpyAsync(E->h_stat, E->stat.p, (size_t)E->R * kStatWidth * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

void ctl_from_plan(RunCtl& c, const AndersonPlan& pl) {
  c.active = 1;
  c.kind = pl.kind;
  c.col = pl.col;
  c.hist = pl.hist;
}

// One implicit step of every run (main.jl:226-307 per run).
int ens_step(collop_ctx* ctx, EnsembleState* E, const collop_step_params& p, collop_step_stats* out) {
  const auto t0 = std::chrono::steady_clock::now();
  int r;
  if ((r = ens_predict(ctx, E, p))) return r;
  E->an.resize(E->R);
  for (int i = 0; i < E->R; ++i) {
    E->an[i].start(p);
    RunCtl& c = E->h_ctl[i];
    const int cur = c.cur;
    std::memset(&c, 0, sizeof(RunCtl));
    c.cur = cur;
    c.exit_picard = p.exit_picard_step;
    ctl_from_plan(c, E->an[i].plan());
  }
  int active = E->R;
  while (active > 0) {
    if ((r = ens_replay(ctx, E, p))) return r;
    for (int i = 0; i < E->R; ++i) {
      RunCtl& c = E->h_ctl[i];
      if (!c.active) {  // its exit update ran in this replay
        c.upd = 0;
        c.copy_best = 0;
        continue;
      }
      AndersonAct act;
      if (E->an[i].after_map(E->h_stat + (size_t)i * kStatWidth, act) != COLLOP_OK)
        return fail(ctx, COLLOP_ERR_NONFINITE,
                    "non-finite residual in run " + std::to_string(i) + ", Picard map " +
                        std::to_string(E->an[i].maps()));
      c.copy_best = act.copy_best ? 1 : 0;
      if (act.exit != kContinue) {
        c.active = 0;
        c.upd = act.exit == kConverged ? (p.exit_picard_step ? 2 : 0) : 3;
        c.cur = 1 - c.cur;  // the update reads the map just done (buffer 1 - cur)
        active--;
        continue;
      }
      c.upd = act.update ? 1 : 0;
      c.upd_hist = act.h;
      c.beta = act.beta;
      for (int j = 0; j < kMaxWindow; ++j) c.gamma[j] = act.gamma[j];
      if (act.rotate) c.cur = 1 - c.cur;
      ctl_from_plan(c, E->an[i].plan());
    }
  }
  // exit updates of the runs that finished in the last replay, then v <- v1
  const EnsView V = view(E, true);
  HIPC(hipMemcpyAsync(E->d_ctl.p, E->h_ctl, (size_t)E->R * sizeof(RunCtl), hipMemcpyHostToDevice, ctx->stream));
  launch_update_b(V, bufs(E), ctx->stream);
  CHECK_LAUNCH();
  launch_copy_b(V, P<double>(E->v1), P<double>(E->v), ctx->stream);
  CHECK_LAUNCH();
  if ((r = ens_project(ctx, E, p.use_logsq))) return r;
  HIPC(hipMemcpyAsync(E->h_scal, E->scal.p, (size_t)E->R * 8 * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < E->R; ++i) E->h_ctl[i].upd = E->h_ctl[i].copy_best = 0;
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (out) {
    for (int i = 0; i < E->R; ++i) {
      collop_step_stats& s = out[i];
      const AndersonRun& a = E->an[i];
      const double* sc = E->h_scal + 8 * i;
      s.iters = a.k_done;
      s.iters_f32 = a.iters32;
      s.n_restart = a.n_restart;
      s.exit_reason = a.exit_reason;
      s.residual = a.res_out;
      s.r0 = a.nrm_r0;
      s.entropy = sc[0];
      s.energy = 0.5 * sc[5];
      s.momentum[0] = sc[3];
      s.momentum[1] = sc[4];
      s.seconds = sec;
      s.handover = 0;
      s.dc_skipped = 0;
    }
  }
  return COLLOP_OK;
}

int need_ens(collop_ctx* ctx) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!ctx->ens) return fail(ctx, COLLOP_ERR_NO_STATE, "no ensemble (collop_ensemble_set)");
  return COLLOP_OK;
}

// Mean time of one persistent pair launch with every run active, FP32c plus
// FP64 (ms).
int time_pair(collop_ctx* ctx, EnsembleState* E, int reps, double* ms_out) {
  hipEvent_t a, b;
  HIPC(hipEventCreate(&a));
  HIPC(hipEventCreate(&b));
  const EnsView V = view(E, true);
  double tot = 0.0;
  for (int kind = 0; kind < 2; ++kind) {
    for (int i = 0; i < E->R; ++i) {
      std::memset(&E->h_ctl[i], 0, sizeof(RunCtl));
      E->h_ctl[i].active = 1;
      E->h_ctl[i].kind = kind;
    }
    HIPC(hipMemcpyAsync(E->d_ctl.p, E->h_ctl, (size_t)E->R * sizeof(RunCtl), hipMemcpyHostToDevice, ctx->stream));
    launch_pair_b(V, tiles(E), P<Src64>(E->s64), P<Src32>(E->s32), 0, P<double>(E->part), ctx->stream);  // warm-up
    CHECK_LAUNCH();
    HIPC(hipEventRecord(a, ctx->stream));
    for (int k = 0; k < reps; ++k)
      launch_pair_b(V, tiles(E), P<Src64>(E->s64), P<Src32>(E->s32), 0, P<double>(E->part), ctx->stream);
    CHECK_LAUNCH();
    HIPC(hipEventRecord(b, ctx->stream));
    HIPC(hipEventSynchronize(b));
    float ms = 0.0f;
    HIPC(hipEventElapsedTime(&ms, a, b));
    tot += ms / reps;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  for (int i = 0; i < E->R; ++i) std::memset(&E->h_ctl[i], 0, sizeof(RunCtl));
  *ms_out = tot;
  return COLLOP_OK;
}

int check_params(collop_ctx* ctx, const collop_step_params* p) {
  if (!p) return fail(ctx, COLLOP_ERR_INVALID, "null params");
  if (!(p->dt > 0.0) || !std::isfinite(p->dt)) return fail(ctx, COLLOP_ERR_INVALID, "dt must be positive");
  if (p->m_anderson < 1 || p->m_anderson > kMaxWindow)
    return fail(ctx, COLLOP_ERR_INVALID, "m_anderson must be in [1, 8]");
  if (p->max_iter < 1) return fail(ctx, COLLOP_ERR_INVALID, "max_iter must be >= 1");
  if (p->stag_window < 1) return fail(ctx, COLLOP_ERR_INVALID, "stag_window must be >= 1");
  if (!(p->damping > 0.0 && p->damping <= 1.0)) return fail(ctx, COLLOP_ERR_INVALID, "damping must be in (0, 1]");
  if (p->precision != COLLOP_PREC_F64 && p->precision != COLLOP_PREC_F32C)
    return fail(ctx, COLLOP_ERR_INVALID, "unknown precision");
  if (p->warmstart != COLLOP_WARM_EULER)
    return fail(ctx, COLLOP_ERR_UNSUPPORTED, "the ensemble step supports the Euler predictor only");
  if (p->record_training)
    return fail(ctx, COLLOP_ERR_UNSUPPORTED, "the ensemble step does not record training data");
  if (p->solver != COLLOP_SOLVER_ANDERSON)
    return fail(ctx, COLLOP_ERR_UNSUPPORTED, "the ensemble step runs the Anderson solver only");
  return COLLOP_OK;
}

}  // namespace
}  // namespace collop

using namespace collop;

extern "C" {

int collop_ensemble_slices(int64_t n) { return slices_for(n); }

int collop_ensemble_set(collop_ctx* ctx, int n_runs, const int64_t* n, const double* v, const double* w) {
  if (!ctx) return COLLOP_ERR_INVALID;
  if (!ctx->has_mesh) return fail(ctx, COLLOP_ERR_NO_MESH, "no mesh (collop_set_mesh)");
  if (n_runs < 1 || n_runs > 4096) return fail(ctx, COLLOP_ERR_INVALID, "n_runs must be in [1, 4096]");
  if (!n || !v || !w) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  int64_t total = 0, padded = 0;
  std::vector<double> wabs(n_runs, 0.0);
  for (int r = 0; r < n_runs; ++r) {
    if (n[r] < 2 || n[r] > ((int64_t)1 << 24))
      return fail(ctx, COLLOP_ERR_INVALID, "each run needs 2 <= N <= 2^24 markers");
    for (int64_t i = 0; i < n[r]; ++i) {
      const double x = w[total + i];
      if (!(x > 0.0) || !std::isfinite(x))
        return fail(ctx, COLLOP_ERR_INVALID, "weights must be positive and finite (G_eff divides by w)");
      wabs[r] += x;
    }
    total += n[r];
    padded += (n[r] + kEnsBlock - 1) / kEnsBlock * kEnsBlock;
  }
  if (padded > ((int64_t)1 << 27)) return fail(ctx, COLLOP_ERR_INVALID, "ensemble too large (> 2^27 markers)");
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipStreamSynchronize(ctx->stream));
  ensemble_free(ctx->ens);
  ctx->ens = new EnsembleState();
  EnsembleState* E = ctx->ens;
  const MeshView& m = ctx->mv;
  E->R = n_runs;
  E->NT = (int)padded;
  // pair-sum form: the context's; symmetric with its tiles (or 4) when every
  // run fits in kSymMaxSlots partial slots
  if (ctx->pair_sym) {
    E->sym = ctx->sym_tiles ? ctx->sym_tiles : 4;
    for (int r = 0; r < n_runs; ++r)
      if (sym_rows((int)n[r], E->sym) > kSymMaxSlots) E->sym = 0;
  }
  E->runs.resize(n_runs);
  std::vector<int2> sb, gb, pb, eb;
  std::vector<int> rowtab((size_t)kRowsKinds * n_runs);
  const int qb = quad_blocks(m);
  int off = 0;
  long long poff = 0;
  for (int r = 0; r < n_runs; ++r) {
    EnsRun& q = E->runs[r];
    q.n = (int)n[r];
    q.off = off;
    q.S = E->sym ? sym_rows(q.n, E->sym) : slices_for(n[r]);
    q.sblk0 = (int)sb.size();
    q.nsb = scatter_blocks(q.n);
    for (int b = 0; b < q.nsb; ++b) sb.push_back(make_int2(r, b));
    q.gblk0 = (int)gb.size();
    q.ngb = gather_blocks(q.n);
    for (int b = 0; b < q.ngb; ++b) gb.push_back(make_int2(r, b));
    q.pblk0 = (int)pb.size();
    q.npb = (q.n + kEnsBlock - 1) / kEnsBlock;
    for (int b = 0; b < q.npb; ++b) pb.push_back(make_int2(r, b));
    q.eblk0 = (int)eb.size();
    q.neb = (2 * q.n + kEnsBlock - 1) / kEnsBlock;
    for (int b = 0; b < q.neb; ++b) eb.push_back(make_int2(r, b));
    q.poff = poff;
    poff += (long long)q.S * 2 * q.n;
    q.scale = scatter_scale(wabs[r]);
    q.inv_scale = 1.0 / q.scale;
    rowtab[(size_t)kRowsMomR0 * n_runs + r] = 3 * q.sblk0;
    rowtab[(size_t)kRowsMomN * n_runs + r] = q.nsb;
    rowtab[(size_t)kRowsStatR0 * n_runs + r] = kStatWidth * q.eblk0;
    rowtab[(size_t)kRowsStatN * n_runs + r] = q.neb;
    rowtab[(size_t)kRowsQ0 * n_runs + r] = 2 * r * qb;
    rowtab[(size_t)kRowsQ1 * n_runs + r] = (2 * r + 1) * qb;
    rowtab[(size_t)kRowsQN * n_runs + r] = qb;
    off += q.npb * kEnsBlock;
  }
  E->n_sblk = (int)sb.size();
  E->n_gblk = (int)gb.size();
  E->n_pblk = (int)pb.size();
  E->n_eblk = (int)eb.size();
  E->part_len = poff;
  int dev_cu = 0;
  if (hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && dev_cu > 0)
    E->n_cu = dev_cu;
  const size_t NT = (size_t)E->NT, nd = (size_t)m.n_dofs, R = (size_t)n_runs;
  const size_t n2b = 2 * NT * 8;
  ENSURE(E->d_runs, R * sizeof(EnsRun));
  ENSURE(E->d_ctl, R * sizeof(RunCtl));
  ENSURE(E->d_tab, (sb.size() + gb.size() + pb.size() + eb.size()) * sizeof(int2));
  ENSURE(E->d_rows, rowtab.size() * sizeof(int));
  for (DBuf* b : {&E->v, &E->v1, &E->rr[0], &E->rr[1], &E->gv[0], &E->gv[1], &E->best, &E->vmid, &E->dv, &E->dS, &E->G})
    ENSURE(*b, n2b);
  ENSURE(E->w, NT * 8);
  ENSURE(E->dF, n2b * kMaxWindow);
  ENSURE(E->dG, n2b * kMaxWindow);
  ENSURE(E->s64, NT * sizeof(Src64));
  ENSURE(E->s32, NT * sizeof(Src32));
  ENSURE(E->part, (size_t)poff * 8);
  ENSURE(E->spart_sc, (size_t)2 * E->n_sblk * nd * 8);
  ENSURE(E->mom, (size_t)E->n_sblk * 3 * 8);
  ENSURE(E->dpart, (size_t)E->n_gblk * 2 * 8);
  ENSURE(E->iota, R * (size_t)m.ne1 * m.ne2 * m.nq * m.nq * 8);
  ENSURE(E->spart_q, R * 2 * (size_t)qb * 8);
  ENSURE(E->vec, R * 4 * nd * 8);
  ENSURE(E->fc, R * nd * 8);
  ENSURE(E->scal, R * 8 * 8);
  ENSURE(E->statp, (size_t)E->n_eblk * kStatWidth * 8);
  ENSURE(E->stat, R * kStatWidth * 8);
  HIPC(hipHostMalloc((void**)&E->h_ctl, R * sizeof(RunCtl), hipHostMallocDefault));
  HIPC(hipHostMalloc((void**)&E->h_stat, R * kStatWidth * 8, hipHostMallocDefault));
  HIPC(hipHostMalloc((void**)&E->h_scal, R * 8 * 8, hipHostMallocDefault));
  std::memset(E->h_ctl, 0, R * sizeof(RunCtl));
  // host staging with padding (positions 0, weights 1 outside the runs)
  std::vector<double> hv(2 * NT, 0.0), hw(NT, 1.0);
  int64_t src = 0;
  for (int r = 0; r < n_runs; ++r) {
    const EnsRun& q = E->runs[r];
    std::memcpy(&hv[q.off], v + 2 * src, (size_t)q.n * 8);
    std::memcpy(&hv[NT + q.off], v + 2 * src + q.n, (size_t)q.n * 8);
    std::memcpy(&hw[q.off], w + src, (size_t)q.n * 8);
    src += q.n;
  }
  std::vector<int2> tab;
  tab.insert(tab.end(), sb.begin(), sb.end());
  tab.insert(tab.end(), gb.begin(), gb.end());
  tab.insert(tab.end(), pb.begin(), pb.end());
  tab.insert(tab.end(), eb.begin(), eb.end());
  HIPC(hipMemcpy(E->d_runs.p, E->runs.data(), R * sizeof(EnsRun), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(E->d_tab.p, tab.data(), tab.size() * sizeof(int2), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(E->d_rows.p, rowtab.data(), rowtab.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPC(hipMemcpy(E->v.p, hv.data(), n2b, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(E->w.p, hw.data(), NT * 8, hipMemcpyHostToDevice));
  for (DBuf* b : {&E->v1, &E->rr[0], &E->rr[1], &E->gv[0], &E->gv[1], &E->best, &E->vmid, &E->dv, &E->dS, &E->G})
    HIPC(hipMemset(b->p, 0, n2b));
  HIPC(hipMemset(E->dF.p, 0, n2b * kMaxWindow));
  HIPC(hipMemset(E->dG.p, 0, n2b * kMaxWindow));
  int r;
  if ((r = build_tiles(ctx, E, E->bt))) return r;
  if ((r = ens_project(ctx, E, 0))) return r;
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

int collop_ensemble_get(collop_ctx* ctx, double* v) {
  int r = need_ens(ctx);
  if (r) return r;
  if (!v) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  EnsembleState* E = ctx->ens;
  HIPC(hipSetDevice(ctx->device));
  std::vector<double> hv((size_t)2 * E->NT);
  HIPC(hipMemcpyAsync(hv.data(), E->v.p, hv.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  int64_t dst = 0;
  for (const EnsRun& q : E->runs) {
    std::memcpy(v + 2 * dst, &hv[q.off], (size_t)q.n * 8);
    std::memcpy(v + 2 * dst + q.n, &hv[E->NT + q.off], (size_t)q.n * 8);
    dst += q.n;
  }
  return COLLOP_OK;
}

int collop_ensemble_coeffs(collop_ctx* ctx, double* f_coeffs) {
  int r = need_ens(ctx);
  if (r) return r;
  if (!f_coeffs) return fail(ctx, COLLOP_ERR_INVALID, "null array");
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipMemcpyAsync(f_coeffs, ctx->ens->fc.p, (size_t)ctx->ens->R * ctx->mv.n_dofs * 8, hipMemcpyDeviceToHost,
                      ctx->stream));
  HIPC(hipStreamSynchronize(ctx->stream));
  return COLLOP_OK;
}

int collop_ensemble_step(collop_ctx* ctx, const collop_step_params* params, int n_steps, collop_step_stats* stats) {
  int r = need_ens(ctx);
  if (r) return r;
  if ((r = check_params(ctx, params))) return r;
  if (params->collision_model != COLLOP_MODEL_LANDAU)
    return fail(ctx, COLLOP_ERR_UNSUPPORTED, "the batched ensemble steps the Landau model only");
  if (n_steps < 0) return fail(ctx, COLLOP_ERR_INVALID, "n_steps must be >= 0");
  HIPC(hipSetDevice(ctx->device));
  for (int s = 0; s < n_steps; ++s)
    if ((r = ens_step(ctx, ctx->ens, *params, stats ? stats + (size_t)s * ctx->ens->R : nullptr))) return r;
  return COLLOP_OK;
}

int collop_ensemble_set_tile(collop_ctx* ctx, int gamma_tile, int waves_per_cu, int use_graph) {
  int r = need_ens(ctx);
  if (r) return r;
  if (gamma_tile != 64 && gamma_tile != 128 && gamma_tile != 256 && gamma_tile != 512)
    return fail(ctx, COLLOP_ERR_INVALID, "gamma_tile must be 64, 128, 256 or 512");
  if (waves_per_cu < 1 || waves_per_cu > 64) return fail(ctx, COLLOP_ERR_INVALID, "waves_per_cu must be in [1, 64]");
  EnsembleState* E = ctx->ens;
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipStreamSynchronize(ctx->stream));
  drop_graph(E);
  E->waves = waves_per_cu;
  E->use_graph = use_graph != 0;
  if (gamma_tile != E->bt || !E->ntiles) return build_tiles(ctx, E, gamma_tile);
  return COLLOP_OK;
}

int collop_ensemble_tune(collop_ctx* ctx, int* gamma_tile, int* waves_per_cu, double* ms) {
  int r = need_ens(ctx);
  if (r) return r;
  EnsembleState* E = ctx->ens;
  HIPC(hipSetDevice(ctx->device));
  HIPC(hipStreamSynchronize(ctx->stream));
  drop_graph(E);
  // sources of the current state in both precisions (G = 0: the pair loop has
  // no data-dependent work besides the coincidence guard)
  const EnsView V = view(E, false);
  const MeshView& m = ctx->mv;
  HIPC(hipMemsetAsync(E->G.p, 0, (size_t)2 * E->NT * 8, ctx->stream));
  launch_pack_b(V, P<double>(E->v), P<double>(E->G), P<double>(E->w), m.lo1, m.hi1, m.lo2, m.hi2, P<Src64>(E->s64),
                P<Src32>(E->s32), ctx->stream);
  CHECK_LAUNCH();
  const int ws[4] = {4, 8, 16, 32};
  double best = 1e300;
  int best_bt = E->bt, best_w = E->waves, best_R = E->sym, idx = 0;
  if (ms)
    for (int i = 0; i < 16; ++i) ms[i] = 0.0;
  if (E->sym) {
    // symmetric form: the gamma tile is the super-row, 128 R markers for R in {2, 4, 8}
    const int Rs[3] = {2, 4, 8};
    for (int i = 0; i < 3; ++i) {
      bool fits = true;
      for (const EnsRun& q : E->runs) fits = fits && sym_rows(q.n, Rs[i]) <= kSymMaxSlots;
      if (!fits) continue;
      if ((r = set_sym_form(ctx, E, Rs[i]))) return r;
      for (int j = 0; j < 4; ++j) {
        E->waves = ws[j];
        double t = 0.0;
        if ((r = time_pair(ctx, E, 3, &t))) return r;
        if (ms) ms[i * 4 + j] = t;
        if (t < best) {
          best = t;
          best_R = Rs[i];
          best_w = ws[j];
        }
      }
    }
    E->waves = best_w;
    if ((r = set_sym_form(ctx, E, best_R))) return r;
    if (gamma_tile) *gamma_tile = 128 * best_R;
    if (waves_per_cu) *waves_per_cu = best_w;
    return COLLOP_OK;
  }
  const int bts[4] = {64, 128, 256, 512};
  for (int bt : bts) {
    if ((r = build_tiles(ctx, E, bt))) return r;
    for (int wv : ws) {
      E->waves = wv;
      double t = 0.0;
      if ((r = time_pair(ctx, E, 3, &t))) return r;
      if (ms) ms[idx] = t;
      idx++;
      if (t < best) {
        best = t;
        best_bt = bt;
        best_w = wv;
      }
    }
  }
  E->waves = best_w;
  if ((r = build_tiles(ctx, E, best_bt))) return r;
  if (gamma_tile) *gamma_tile = best_bt;
  if (waves_per_cu) *waves_per_cu = best_w;
  return COLLOP_OK;
}

}  // extern "C"
