"""Benchmark: Landau implicit steps/s at N = 65536 on one MI355X.

Workload (bimodal_n65536): N = 65536 markers from two equal-weight Maxwellians
at (+-2, 0) with covariance 0.5 I, weights 1/N, P_DEG = 2 B-splines on 32 x 32
cells over [-6, 6]^2, dt = 0.02, Anderson m = 5, tol 1e-12 (solver knobs of the
reference's Landau presets).  A "step" is one implicit time step: Euler
predictor, the implicit solve of the Gonzalez discrete-gradient Picard map to
the residual tolerance max(1e-12 ||v||, 1e-10), then the projection / entropy /
energy / momentum of the new state.  The markers are resident in HBM before the
timed region starts.

Paths, on identical markers and to the same tolerance:
  * plain FP64 (the reference's GPU path, restated): every Picard map in FP64
    with the ordered pair sum (every ordered pair, collision_gpu.jl:49-107),
    Anderson solver (step_anderson!, solver.jl:300-423) controlled from the
    host map by map;
  * chosen: frozen-metric defect correction (step_defect!, solver.jl:569-634)
    with outer maps in the FP32-compensated pair sum until the FP32 rounding
    floor, then FP64 outer maps (every accepted state comes from an FP64 map),
    handing a step with stiff near pairs over to the Anderson solve
    (step_anderson!, FP64 maps, control on the device); every pair sum in the
    symmetric form (each unordered pair evaluated once, applied to both
    markers); Euler predictor with the learned MLP correction (--no-mlp: Euler
    alone);
  * comparisons: the plain FP64 path with the symmetric pair sum (the kernel's
    share of the gain), the Anderson solve alone (FP32c then FP64 maps, control
    on the device), that with stiff-cluster elimination, the defect correction
    without the hand-over, and the chosen path with the Euler predictor alone.
`value` is the chosen path's steps/s; the plain FP64 figure and the ratio are
reported beside it.  Per step, |E_{n+1} - E_n| / E_0 and |P_{n+1} - P_n| are
checked against 1e-10 and the entropy against non-decrease.

The roofline object covers the dominant kernel of the chosen path's timed region
over its mean launch duration, measured with HIP events on the library's
stream.  Algorithmic work: the ordered sum costs 23 flops per ordered in-domain
pair (functions.jl:221-231), 38 with the metric (functions.jl:276-281); the
symmetric sum costs, per unordered pair, the same term once plus its
application to the second marker, 28 flops (+5: w_g times the term, two
products with d, two sums) and 50 with the metric (+7 more: w_g |d|^-3 and
three multiply-adds).  The CPU baseline is the oracle (a C++ FP64 restatement
of the reference CPU path, OpenMP over targets like functions.jl:210) timing
whole Picard maps at the same N on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
       python bench.py --report precision --config bimaxwellian_n262144 --steps 20
       python bench.py --config ensemble48 --steps 3
       python bench.py --config lb_bimodal_n40000 --steps 100
       python bench.py --report scatter-order --steps 50   (under rocprofv3 --kernel-trace)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "collisionoperators.jl_amd"))

FP64_VECTOR_PEAK_TFLOPS = 78.6     # MI355X published FP64 vector peak
FP32_VECTOR_PEAK_TFLOPS = 157.3    # MI355X published FP32 vector peak (packed)
FLOPS_PER_PAIR = 23


def equal_pow2_weights(w) -> bool:
    """Every weight the same power of two in [2^-60, 2^60]: the symmetric sums
    then run their equal-weight form (collop_set_weight_path)."""
    m, e = np.frexp(w[0])
    return bool(np.all(w == w[0]) and m == 0.5 and -59 <= e <= 61)


def pair_kernels(tiles: int, pow2w: bool = False):
    """Pair kernels by profiling kind (collop_profile): name, peak TFLOP/s, flops
    per pair unit, unit.  tiles = 0: the ordered sum (flops per ordered pair: 23
    for the pair sum, functions.jl:221-231, 38 with the metric,
    functions.jl:276-281); tiles = R: the symmetric sum with R waves per block
    (flops per unordered pair: 28, 50 with the metric); pow2w: its equal-weight
    form (the FP64 pair sum, the FP32c pair sum and metric kernels)."""
    if not tiles:
        return [("k_pair_f64<1,1,false>", FP64_VECTOR_PEAK_TFLOPS, 23, "ordered"),
                ("k_pair_f32c<1,32,false>", FP32_VECTOR_PEAK_TFLOPS, 23, "ordered"),
                ("k_pair_f64<1,1,true> (pair sum + metric)", FP64_VECTOR_PEAK_TFLOPS, 38, "ordered"),
                ("k_pair_f32c<1,32,true> (pair sum + metric)", FP32_VECTOR_PEAK_TFLOPS, 38, "ordered")]
    R, e = tiles, (",true" if pow2w else "")
    return [(f"k_pair_sym_f64<{R},false,false{e}>", FP64_VECTOR_PEAK_TFLOPS, 28, "unordered"),
            (f"k_pair_sym_f32c<{R},false,false{e}>", FP32_VECTOR_PEAK_TFLOPS, 28, "unordered"),
            (f"k_pair_sym_f64<{R},true,false> (pair sum + metric)", FP64_VECTOR_PEAK_TFLOPS, 50, "unordered"),
            (f"k_pair_sym_f32c<{R},true,false{e}> (pair sum + metric)", FP32_VECTOR_PEAK_TFLOPS, 50, "unordered")]
# HBM bytes per pair-sum launch at bimodal_n65536 from rocprofv3 FETCH_SIZE (x2.0)
# and WRITE_SIZE (x1.0), both scaled by the counters of a 256 MiB copy; see
# profiles/pair_sum_bimodal_n65536.md (the general-form metric kernels from an
# earlier run of the same counters).
PAIR_TRAFFIC_BYTES = {"k_pair_sym_f64<4,false,false,true>": (2 * 48711 + 131072) * 1024,
                      "k_pair_sym_f32c<4,false,false,true>": (2 * 19347 + 131072) * 1024,
                      "k_pair_sym_f32c<4,true,false,true> (pair sum + metric)": (2 * 33961 + 327680) * 1024,
                      "k_pair_f64<1,1,false>": (2 * 17939 + 32768) * 1024,
                      "k_pair_f32c<1,32,false>": 50730 * 1024,
                      "k_pair_f64<1,1,true> (pair sum + metric)": 117802 * 1024,
                      "k_pair_f32c<1,32,true> (pair sum + metric)": 99890 * 1024,
                      "k_pair_sym_f64<4,false,false>": (2 * 50966 + 131072) * 1024,
                      "k_pair_sym_f32c<4,false,false>": (2 * 15446 + 131072) * 1024,
                      "k_pair_sym_f64<4,true,false> (pair sum + metric)": 485372 * 1024,
                      "k_pair_sym_f32c<4,true,false> (pair sum + metric)": 397245 * 1024}


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def run_path(ws, p, v, w, precision, warmup, steps, solver=None, warmstart=None, symmetric=True, device_solve=True):
    import collop
    collop.set_pair_mode(ws, symmetric)
    collop.set_device_solve(ws, device_solve)
    tiles = collop.pair_mode(ws, v.shape[0])[1]
    sim = collop.Simulation(ws, v, w)
    prm = p.step_params(precision=precision, solver=solver, warmstart=warmstart)
    log(f"  knobs: solver {prm.solver} precision {prm.precision} dc_outer_m {prm.dc_outer_m} "
        f"dc_near_theta {prm.dc_near_theta} dc_metric_reuse {prm.dc_metric_reuse} dc_eta {prm.dc_eta} "
        f"dc_eta_f64 {prm.dc_eta_f64} cl_theta {prm.cl_theta:g} cl_kmax {prm.cl_kmax} "
        f"warmstart {prm.warmstart} pair sum {'symmetric, tiles %d' % tiles if tiles else 'ordered'}")
    E0 = 0.5 * float(np.sum(w * (v[:, 0] ** 2 + v[:, 1] ** 2)))
    for i in range(warmup):
        st = sim.run(prm, 1)[0]
        log(f"  warmup step {i}: iters {st['iters']} (f32 {st['iters_f32']}) {st['seconds'] * 1e3:.1f} ms")
    # state before the timed region: energy / momentum / entropy of the current markers
    vc = sim.markers()
    E_prev = 0.5 * float(np.sum(w * (vc[:, 0] ** 2 + vc[:, 1] ** 2)))
    P_prev = (w[:, None] * vc).sum(axis=0)
    S_prev = collop.compute_r_and_entropy_gpu(ws, sim.coeffs())[1]
    collop.profile(ws, True)
    t0 = time.perf_counter()
    stats = []
    for i in range(steps):
        st = sim.run(prm, 1)[0]
        stats.append(st)
        log(f"  step {i}: iters {st['iters']} (f32 {st['iters_f32']}, inner {st['inner_maps']}"
            + (f", handed over after {st['handover']}" if st['handover'] else "")
            + f") exit {st['exit_reason']} res {st['residual']:.2e} {st['seconds'] * 1e3:.1f} ms")
    elapsed = time.perf_counter() - t0
    ms, launches = collop.profile(ws, False)
    out = dict(steps_per_s=steps / elapsed, elapsed=elapsed, stats=stats, pair_ms=ms.tolist(), tiles=tiles,
               pair_launches=launches, markers=sim.markers())
    out.update(constraint_summary(stats, E0, E_prev, P_prev, S_prev))
    return out


def constraint_summary(stats, E0, E_prev, P_prev, S_prev):
    """The metric's constraint over the timed steps (per-step statistics of
    collop_run_steps, in order) from the state before them (energy E_prev,
    momentum P_prev, entropy S_prev; E0 the initial energy): every step
    converged (exit 0, solver.jl:349-367), per step |E_n+1 - E_n| / E0 and
    |P_n+1 - P_n| at most 1e-10, and the entropy never falls."""
    dE = dP = 0.0
    dS_min = np.inf
    P_prev = np.asarray(P_prev, dtype=float)
    for st in stats:
        dE = max(dE, abs(st["energy"] - E_prev) / E0)
        dP = max(dP, float(np.max(np.abs(np.array(st["momentum"]) - P_prev))))
        dS_min = min(dS_min, st["entropy"] - S_prev)
        E_prev, P_prev, S_prev = st["energy"], np.array(st["momentum"]), st["entropy"]
    all_conv = all(s["exit_reason"] == 0 for s in stats)
    return dict(max_dE_rel=dE, max_dP=dP, min_dS=float(dS_min),
                mean_iters=float(np.mean([s["iters"] for s in stats])),
                mean_iters_f32=float(np.mean([s["iters_f32"] for s in stats])),
                mean_inner=float(np.mean([s["inner_maps"] for s in stats])),
                mean_r0=float(np.mean([s["r0"] for s in stats])),
                all_converged=all_conv, stagnated_steps=sum(1 for s in stats if s["exit_reason"] == 1),
                max_iter_steps=sum(1 for s in stats if s["exit_reason"] == 2),
                max_exit_residual=float(max(s["residual"] for s in stats)),
                clusters_mean=float(np.mean([s["n_clusters"] for s in stats])),
                cluster_solve_failures=int(sum(s["cl_fail"] for s in stats)),
                handed_over_steps=int(sum(s["handover"] > 0 for s in stats)),
                constraint_met=bool(all_conv and dE <= 1e-10 and dP <= 1e-10 and dS_min >= 0.0))


def path_summary(res, ref=None):
    """The per-path figures of the JSON line: speed, maps, and the metric's constraint."""
    out = {"steps_per_s": round(res["steps_per_s"], 4), "mean_iters": res["mean_iters"],
           "mean_iters_f32": res["mean_iters_f32"], "mean_r0": res["mean_r0"],
           "all_converged": res["all_converged"], "stagnated_steps": res["stagnated_steps"],
           "max_iter_steps": res["max_iter_steps"], "max_exit_residual": res["max_exit_residual"],
           "max_dE_rel": res["max_dE_rel"], "max_dP": res["max_dP"], "min_dS": res["min_dS"],
           "constraint_met": res["constraint_met"]}
    if res["mean_inner"]:
        out["mean_inner"] = res["mean_inner"]
        out["handed_over_steps"] = res["handed_over_steps"]
    if res["clusters_mean"]:
        out["clusters_mean"] = res["clusters_mean"]
        out["cluster_solve_failures"] = res["cluster_solve_failures"]
    if ref is not None:
        out["ratio_vs_fp64"] = round(res["steps_per_s"] / ref["steps_per_s"], 4)
    return out


def cpu_baseline(p, v, w, mean_maps_per_step, budget_s=20.0):
    """Oracle Picard maps (FP64, OpenMP over targets) on this host, extrapolated to steps/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    # every host thread this process may use (OMP_NUM_THREADS, else the affinity mask)
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or nproc
    m = O.Mesh(p.bp1, p.bp2)
    fc = m.l2_project(v, w)
    S0 = m.entropy(fc)
    P = O.default_params(p.DT, m=p.m_anderson, nthreads=threads)
    v1, _ = m.predictor(P, v, w, fc)
    times = []
    t_all = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        v1 = m.picard_map(P, v1, v, w, S0)
        times.append(time.perf_counter() - t0)
        log(f"  cpu map {len(times)}: {times[-1]:.2f} s")
        if time.perf_counter() - t_all + times[-1] > budget_s or len(times) >= 5:
            break
    t_map = float(np.mean(times))
    N = v.shape[0]
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(value=1.0 / (t_map * mean_maps_per_step), unit="steps/s", cores=threads, kind="port",
                sample=f"{len(times)} full FP64 Picard maps of the oracle at N={N} (pair sum + projection chain), "
                       f"{t_map:.2f} s each; steps/s = 1 / (map time x {mean_maps_per_step:.1f} maps per step, the "
                       f"FP64 GPU run's mean incl. the predictor)",
                pairs_per_s=N * (N - 1) / t_map, nproc=os.cpu_count(), affinity_cpus=nproc,
                omp_num_threads=os.environ.get("OMP_NUM_THREADS"), cpu_model=model)


def precision_report(args):
    """FP32-compensated vs FP64 on the same markers: per-field errors and the
    per-step conservation drift of both paths, plus the FP64 pair sum against
    the CPU oracle on a subsample of targets."""
    import collop
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    p = collop.config(args.config)
    v, w = collop.initial_markers(p)
    ws = collop.build_workspace(p, device=0)
    N = v.shape[0]
    steps = args.steps
    out = {"workload": args.config, "N": N, "steps": steps, "dt": p.DT}
    runs = {}
    base = collop.chosen_config(args.config)   # the same solver for both: only the precision differs
    out["solver"] = ("the chosen path's solver with the Euler predictor: defect correction with outer Anderson "
                     "window 3 and near pairs w/|d| > 0.02, stiff steps handed over to the FP64 Anderson solve")
    for name, prec in (("fp64", collop.PREC_F64), ("f32c", collop.PREC_F32C)):
        sim = collop.Simulation(ws, v, w)
        E0 = 0.5 * float(np.sum(w * (v[:, 0] ** 2 + v[:, 1] ** 2)))
        P0 = (w[:, None] * v).sum(axis=0)
        S_prev = collop.compute_r_and_entropy_gpu(ws, sim.coeffs())[1]
        E_prev, P_prev = E0, P0
        drift = []
        t0 = time.perf_counter()
        for i in range(steps):
            st = sim.run(base.step_params(precision=prec), 1)[0]
            drift.append({"dE_rel": abs(st["energy"] - E_prev) / E0,
                          "dP": float(np.max(np.abs(np.array(st["momentum"]) - P_prev))),
                          "dS": st["entropy"] - S_prev, "iters": st["iters"], "iters_f32": st["iters_f32"],
                          "inner": st["inner_maps"], "residual": st["residual"], "exit": st["exit_reason"]})
            E_prev, P_prev, S_prev = st["energy"], np.array(st["momentum"]), st["entropy"]
            drift[-1]["handover"] = st["handover"]
            log(f"  {name} step {i}: iters {st['iters']} (f32 {st['iters_f32']}, handover {st['handover']}) exit "
                f"{st['exit_reason']} dE/E {drift[-1]['dE_rel']:.1e} dS {drift[-1]['dS']:.2e}")
        runs[name] = dict(v=sim.markers(), fc=sim.coeffs(), S=S_prev, E=E_prev, P=P_prev, drift=drift,
                          seconds=time.perf_counter() - t0)
    a, b = runs["fp64"], runs["f32c"]
    rel = lambda x, y: float(np.max(np.abs(x - y)) / np.max(np.abs(y)))  # noqa: E731
    out["f32c_vs_fp64_after_steps"] = {"v_max_abs": float(np.max(np.abs(b["v"] - a["v"]))), "v_rel": rel(b["v"], a["v"]),
                                       "f_coeffs_rel": rel(b["fc"], a["fc"]), "entropy_abs": abs(b["S"] - a["S"]),
                                       "energy_rel": abs(b["E"] - a["E"]) / abs(a["E"]),
                                       "momentum_abs": float(np.max(np.abs(b["P"] - a["P"])))}
    for name in runs:
        d = runs[name]["drift"]
        out[name] = {"max_dE_rel": max(x["dE_rel"] for x in d), "max_dP": max(x["dP"] for x in d),
                     "min_dS": min(x["dS"] for x in d), "all_converged": all(x["exit"] == 0 for x in d),
                     "mean_iters": float(np.mean([x["iters"] for x in d])),
                     "mean_iters_f32": float(np.mean([x["iters_f32"] for x in d])),
                     "steps_per_s": steps / runs[name]["seconds"], "per_step": d}
    # per-map field errors on the initial state: pair sum FP32c and FP64 vs the oracle (targets subsample)
    m = O.Mesh(p.bp1, p.bp2)
    fc = m.l2_project(v, w)
    L = m.mass_solve(m.compute_r(fc))
    G = m.compute_G(v, L)
    sub = min(N, 2048)
    ref = O.collision_rows(m.box, v, w, G, 0, sub)[:sub]
    d64 = np.zeros((N, 2), order="F")
    d32 = np.zeros((N, 2), order="F")
    collop.compute_collision_gpu(ws, d64, v, w, G)
    collop.compute_collision_gpu32(ws, d32, v, w, G)
    fcg = np.zeros(ws.n_dofs)
    collop.l2_project_gpu(ws, fcg, v, w)
    Gg = np.zeros((N, 2), order="F")
    collop.compute_G_gpu(ws, Gg, v, L)
    out["per_map_vs_oracle"] = {"pair_sum_fp64_rel": rel(d64[:sub], ref), "pair_sum_f32c_rel": rel(d32[:sub], ref),
                                "f_coeffs_rel": rel(fcg, fc), "G_rel": rel(Gg, G), "targets_checked": sub}
    print(json.dumps(out), flush=True)


def scatter_order_report(args):
    """Marker order and the scatter / gathers: the projection (k_scatter + the
    mass solve) and the gradient gather (k_gather) of the workload's markers,
    `--steps` times each in the IC's order and then sorted by mesh cell (the
    order projection_gpu.jl's cell-sorted kernels assume), through the hook
    entry points.  The coefficients agree bit for bit (the scatter's
    fixed-point sums do not depend on order).  Run under rocprofv3
    --kernel-trace: the first half of the k_scatter / k_gather launches is the
    IC's order, the second half the sorted one."""
    import collop
    p = collop.config(args.config)
    v, w = collop.initial_markers(p)
    ws = collop.build_workspace(p, device=0)
    N = v.shape[0]
    h1 = (p.bp1[-1] - p.bp1[0]) / (len(p.bp1) - 1)
    h2 = (p.bp2[-1] - p.bp2[0]) / (len(p.bp2) - 1)
    ci = np.clip(((v[:, 0] - p.bp1[0]) / h1).astype(np.int64), 0, len(p.bp1) - 2)
    cj = np.clip(((v[:, 1] - p.bp2[0]) / h2).astype(np.int64), 0, len(p.bp2) - 2)
    perm = np.argsort(ci + (len(p.bp1) - 1) * cj, kind="stable")
    out = {"workload": args.config, "N": N, "calls": args.steps}
    res = {}
    for name, (vv, ww) in (("ic_order", (v, w)), ("cell_sorted", (np.asfortranarray(v[perm]), w[perm]))):
        fc = np.zeros(ws.n_dofs)
        G = np.zeros((N, 2), order="F")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            collop.l2_project_gpu(ws, fc, vv, ww)
        t_proj = (time.perf_counter() - t0) / args.steps
        r, _ = collop.compute_r_and_entropy_gpu(ws, fc)
        L = collop.mass_solve_gpu(ws, r)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            collop.compute_G_gpu(ws, G, vv, L)
        t_gath = (time.perf_counter() - t0) / args.steps
        res[name] = (fc.copy(), G.copy())
        out[name] = {"projection_call_ms": round(t_proj * 1e3, 4), "gather_call_ms": round(t_gath * 1e3, 4)}
        log(f"[scatter-order] {name}: projection call {t_proj * 1e3:.3f} ms, gather call {t_gath * 1e3:.3f} ms")
    out["coeffs_bitwise_equal"] = bool(np.array_equal(res["ic_order"][0], res["cell_sorted"][0]))
    out["G_max_abs_diff"] = float(np.max(np.abs(res["ic_order"][1][perm] - res["cell_sorted"][1])))
    print(json.dumps(out), flush=True)


def warmstart_report(args):
    """Learned warm start: record 8 seeded N = 65536 runs (the three ICs in turn,
    32 x 32 mesh, dt 0.02) for `--train-steps` steps each, train the MLP with
    Adam (lr 1e-3) on the device, then step a held-out seed `--steps` times cold
    (Euler predictor) and warm (Euler + MLP), with the plain Anderson solver and
    with defect correction (FP32c outer maps), to the same tolerance."""
    import collop
    from collop import warmstart as W
    p = collop.config("bimodal_n65536")
    ws = collop.build_workspace(p, device=0)
    runs = W.training_runs(8, 65536, seed0=100)
    rec_params = collop.defect_config("bimodal_n65536").step_params()
    t0 = time.perf_counter()
    n = W.record(ws, runs, args.train_steps, rec_params, log=log)
    t_rec = time.perf_counter() - t0
    log(f"[warmstart] recorded {n} samples in {t_rec:.1f} s")
    ml = W.WarmStart(ws)
    ml.set_params(W.glorot_init(0))
    mu, sd, os_ = ml.fit_norm()
    t0 = time.perf_counter()
    losses = ml.train(args.epochs, batch=65536, lr=1e-3, seed=1)
    t_train = time.perf_counter() - t0
    log(f"[warmstart] {args.epochs} epochs in {t_train:.1f} s, loss per epoch {losses}")
    if args.save_model:
        ml.save(args.save_model)
    v, w = collop.markers("bimodal", 65536, 999, sigma1=0.5 ** 0.5, sigma2=0.5 ** 0.5, peak=2.0)  # held out
    out = {"report": "warmstart", "train_samples": n, "record_s": t_rec, "train_s": t_train,
           "epochs": args.epochs, "loss_per_epoch": losses, "held_out": "bimodal seed 999", "steps": args.steps,
           "norm": {"mu": mu.tolist(), "sd": sd.tolist(), "out_scale": os_.tolist()}}
    for name, base in (("anderson_fp64", collop.config("bimodal_n65536")),
                       ("defect_correction", collop.defect_config("bimodal_n65536"))):
        res = {}
        for warm in (collop.WARM_EULER, collop.WARM_MLP):
            prm = base.step_params(warmstart=warm)
            r = W.evaluate(ws, v, w, prm, args.steps)
            res["mlp" if warm else "euler"] = r
            log(f"  {name} {'mlp' if warm else 'euler'}: maps {r['mean_maps']:.2f} inner {r['mean_inner']:.1f} "
                f"r0 {r['mean_r0']:.3e} {r['steps_per_s']:.2f} steps/s")
        a, b = res["euler"], res["mlp"]
        out[name] = {k: {kk: vv for kk, vv in d.items() if kk != "markers"} for k, d in res.items()}
        out[name]["maps_ratio_mlp_vs_euler"] = b["mean_maps"] / a["mean_maps"]
        out[name]["r0_ratio_mlp_vs_euler"] = b["mean_r0"] / a["mean_r0"]
        out[name]["steps_per_s_ratio"] = b["steps_per_s"] / a["steps_per_s"]
        out[name]["max_abs_dv_mlp_vs_euler"] = float(np.max(np.abs(b["markers"] - a["markers"])))
    print(json.dumps(out), flush=True)


def ensemble_bench(args):
    """Batched ensemble: 48 runs, N log-uniform in [2048, 65536], random IC,
    shared 32x32 mesh, FP32c pair sums (FP64 finish), dt 0.02.  Times the
    batched step (one persistent pair launch per iteration, hipGraph replay)
    with the symmetric pair sum (`value`) and with the ordered one, and, for
    comparison, the same runs stepped one after another by the single-run path
    (symmetric pair sum)."""
    import collop
    runs = collop.ensemble_markers(48, seed=7)
    ws = collop.Workspace(np.linspace(-6, 6, 33), np.linspace(-6, 6, 33), device=0)
    p = collop.config("bimodal_n65536").step_params(precision=collop.PREC_F32C)
    n = np.array([r[3] for r in runs])
    pairs = float((n.astype(float) * (n - 1)).sum())
    log(f"[ensemble] {len(runs)} runs, N in [{n.min()}, {n.max()}], sum N {n.sum()}, {pairs:.3e} pairs per iteration")
    res = {}
    for form, symmetric in (("symmetric", True), ("ordered", False)):
        collop.set_pair_mode(ws, symmetric, 4 if symmetric else 0)
        ens = collop.Ensemble(ws, [(v, w) for v, w, _, _ in runs])
        t0 = time.perf_counter()
        bt, waves, ms = ens.tune()
        log(f"[ensemble {form}] autotuner ({time.perf_counter() - t0:.1f} s): gamma tile {bt}, waves/CU {waves}; "
            f"FP32c+FP64 pair launch ms rows gamma 64..512 (symmetric: super-rows 256, 512, 1024), "
            f"cols waves 4..32:\n"
            f"{np.array2string(ms, precision=3)}")
        for i in range(args.warmup):
            st = ens.step(p, 1)[0]
            log(f"  warmup step {i}: max iters {max(s['iters'] for s in st)} {st[0]['seconds'] * 1e3:.0f} ms")
        t0 = time.perf_counter()
        stats = ens.step(p, args.steps)
        el = time.perf_counter() - t0
        for i, st in enumerate(stats):
            log(f"  step {i}: iters mean {np.mean([s['iters'] for s in st]):.1f} max {max(s['iters'] for s in st)} "
                f"(f32 mean {np.mean([s['iters_f32'] for s in st]):.1f}) {st[0]['seconds'] * 1e3:.0f} ms")
        res[form] = dict(run_steps_per_s=len(runs) * args.steps / el, el=el, gamma_tile=bt, waves_per_cu=waves,
                         tuner_ms=ms.tolist(), all_converged=all(s["exit_reason"] == 0 for st in stats for s in st))
        log(f"[ensemble {form}] {res[form]['run_steps_per_s']:.2f} run-steps/s")
    # the single-run chosen path (FP32c defect correction with the hand-over,
    # learned warm start) on 4 contexts with a HIP stream each, stepped by host
    # threads
    from collop import warmstart as W
    pch = collop.chosen_config("bimodal_n65536").step_params(precision=collop.PREC_F32C,
                                                             warmstart=collop.WARM_MLP)
    sens = collop.StreamEnsemble(np.linspace(-6, 6, 33), np.linspace(-6, 6, 33), [(v, w) for v, w, _, _ in runs],
                                 workers=4, warm_start=W.load_packaged)
    if args.warmup:
        sens.step(pch, args.warmup)
    t0 = time.perf_counter()
    sst = sens.step(pch, args.steps)
    s_el = time.perf_counter() - t0
    sens.close()
    stream = {"run_steps_per_s": round(len(runs) * args.steps / s_el, 4), "workers": 4,
              "mean_outer_maps": float(np.mean([s["iters"] for st in sst for s in st])),
              "all_converged": all(s["exit_reason"] == 0 for st in sst for s in st)}
    log(f"[ensemble streams] chosen single-run path on 4 streams: {stream['run_steps_per_s']:.2f} run-steps/s, "
        f"mean outer maps {stream['mean_outer_maps']:.1f}")
    # the same runs stepped one after another (single-run path, symmetric pair
    # sum), a bounded sample of steps
    collop.set_pair_mode(ws, True)
    seq_steps = max(1, min(args.steps, 2))
    t0 = time.perf_counter()
    for v, w, _, nn in runs:
        sim = collop.Simulation(ws, v, w)
        sim.run(p, seq_steps)
    seq_el = time.perf_counter() - t0
    sequential = len(runs) * seq_steps / seq_el
    sym = res["symmetric"]
    out = {"metric": "ensemble run-steps/s (48 runs, N log-uniform [2048, 65536], FP32c + FP64 finish, dt 0.02)",
           "value": round(sym["run_steps_per_s"], 4), "unit": "run-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * sym["el"] / args.steps, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32c+f64",
           "data": "synthetic (seeded: N, IC and markers from PCG64 seed 7)",
           "config": {"workload": "ensemble48", "runs": len(runs), "sum_N": int(n.sum()), "N_min": int(n.min()),
                      "N_max": int(n.max()), "mesh": "32x32", "dt": p.dt, "pair_sum": "symmetric, 4 waves per block"},
           "pairs_per_iteration": pairs, "waves_per_cu": sym["waves_per_cu"],
           "all_converged": sym["all_converged"] and res["ordered"]["all_converged"],
           "ordered_pair_sum": {"run_steps_per_s": round(res["ordered"]["run_steps_per_s"], 4),
                                "gamma_tile": res["ordered"]["gamma_tile"],
                                "waves_per_cu": res["ordered"]["waves_per_cu"]},
           "sequential_run_steps_per_s": round(sequential, 4),
           "streams_chosen_path": stream,
           "batched_vs_sequential": round(sym["run_steps_per_s"] / sequential, 3),
           "tuner_ms": {"symmetric": sym["tuner_ms"][0], "ordered": res["ordered"]["tuner_ms"]}}
    print(json.dumps(out), flush=True)


def lb_bench(args):
    """Lenard-Bernstein workload (parameters_LB_bimodal_v1.jl: N = 40000,
    N(+-2, 1) x N(0, 1), non-uniform 27x24 cells, nu = 1, dt = 0.001, m = 8):
    device-resident implicit steps/s, with the oracle's LB step (serial C++)
    timed on a bounded sample of the same steps beside it."""
    import collop
    p = collop.config(args.config)
    v, w = collop.initial_markers(p)
    ws = collop.build_workspace(p, device=0)
    N = v.shape[0]
    prm = p.step_params()
    sim = collop.Simulation(ws, v, w)
    log(f"[lb] {args.config}: N={N}, mesh {len(p.bp1) - 1}x{len(p.bp2) - 1}, dt={p.DT}, nu={p.nu}, m={p.m_anderson}")
    for i in range(args.warmup):
        st = sim.run(prm, 1)[0]
        log(f"  warmup step {i}: iters {st['iters']} {st['seconds'] * 1e3:.1f} ms")
    vc = sim.markers()
    E0 = 0.5 * float(np.sum(w * (vc[:, 0] ** 2 + vc[:, 1] ** 2)))
    E_prev, P_prev = E0, (w[:, None] * vc).sum(axis=0)
    S_prev = collop.compute_r_and_entropy_gpu(ws, sim.coeffs())[1]
    t0 = time.perf_counter()
    stats = sim.run(prm, args.steps)
    el = time.perf_counter() - t0
    dE = dP = 0.0
    dS_min = np.inf
    for i, st in enumerate(stats):
        log(f"  step {i}: iters {st['iters']} exit {st['exit_reason']} res {st['residual']:.2e} "
            f"{st['seconds'] * 1e3:.2f} ms")
        dE = max(dE, abs(st["energy"] - E_prev) / E0)
        dP = max(dP, float(np.max(np.abs(np.array(st["momentum"]) - P_prev))))
        dS_min = min(dS_min, st["entropy"] - S_prev)
        E_prev, P_prev, S_prev = st["energy"], np.array(st["momentum"]), st["entropy"]
    maps = float(np.mean([s["iters"] for s in stats]))
    sps = args.steps / el
    log(f"[lb] {sps:.2f} steps/s, {maps:.1f} maps per step, {1e3 * el / args.steps / maps:.3f} ms per map")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        m = O.Mesh(p.bp1, p.bp2)
        fc = m.l2_project(v, w)
        S = m.entropy(fc)
        P = O.default_params(p.DT, m=p.m_anderson, nthreads=1, collision_model=1, nu=p.nu)
        vv, t_all, done, it_cpu = v, time.perf_counter(), 0, 0
        while done < 3 and time.perf_counter() - t_all < 20.0:
            vv, fc, st = m.time_step(P, vv, w, fc, S)
            S, done, it_cpu = st.entropy, done + 1, it_cpu + st.iters
            log(f"  cpu step {done}: iters {st.iters} exit {st.exit_reason}")
        t_cpu = (time.perf_counter() - t_all) / done
        cpu = {"value": round(1.0 / t_cpu, 4), "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"the oracle's first {done} LB steps from the same markers (serial C++, {it_cpu / done:.1f} "
                         f"maps per step), {t_cpu:.3f} s each"}
    out = {"metric": "Lenard-Bernstein implicit steps/s (N = 40000, 27x24 non-uniform cells, nu 1, dt 0.001)",
           "value": round(sps, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(1e3 * el / args.steps, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded PCG64 markers, seed 42)",
           "config": {"workload": args.config, "N": N, "mesh": f"{len(p.bp1) - 1}x{len(p.bp2) - 1}", "dt": p.DT,
                      "nu": p.nu, "m": p.m_anderson, "solver": "anderson"},
           "mean_maps_per_step": maps, "ms_per_map": round(1e3 * el / args.steps / maps, 4),
           "all_converged": all(s["exit_reason"] == 0 for s in stats),
           "conservation": {"max_dE_rel": dE, "max_dP": dP, "min_dS": float(dS_min)},
           "roofline": None, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # default: the workload's own run, 50 steps from the initial markers
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mlp", action="store_true",
                    help="chosen path with the Euler predictor alone (the learned warm start becomes a comparison "
                         "path)")
    ap.add_argument("--no-variants", action="store_true", help="skip the comparison paths")
    ap.add_argument("--config", default="bimodal_n65536",
                    help="bimodal_n65536 (headline), gaussian_n4096, bimaxwellian_n262144, ensemble48 or "
                         "lb_bimodal_n40000 (Lenard-Bernstein)")
    ap.add_argument("--train-steps", type=int, default=200, help="warmstart report: recorded steps per run")
    ap.add_argument("--epochs", type=int, default=3, help="warmstart report: Adam epochs")
    ap.add_argument("--save-model", default=None, help="warmstart report: write the trained model (.npz)")
    ap.add_argument("--report", choices=["precision", "warmstart", "scatter-order"], default=None,
                    help="precision: FP32c vs FP64 field errors and drift (use --config bimaxwellian_n262144)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    if args.report == "precision":
        precision_report(args)
        return
    if args.report == "warmstart":
        warmstart_report(args)
        return
    if args.report == "scatter-order":
        scatter_order_report(args)
        return
    if args.config == "ensemble48":
        ensemble_bench(args)
        return
    if args.config.startswith("lb_"):
        lb_bench(args)
        return

    import collop
    p = collop.config(args.config)
    v, w = collop.initial_markers(p)
    ws = collop.build_workspace(p, device=0)
    N = v.shape[0]
    box = ws.box
    inside = (v[:, 0] > box[0]) & (v[:, 0] < box[1]) & (v[:, 1] > box[2]) & (v[:, 1] < box[3])
    n_in = int(inside.sum())
    W, K = args.warmup, args.steps

    log(f"[bench] {args.config}: N={N} (in domain {n_in}), mesh {len(p.bp1) - 1}x{len(p.bp2) - 1}, dt={p.DT}; "
        f"{W} untimed steps, then {K} timed steps (steps {W}..{W + K - 1}, t = {W * p.DT:.2f}..{(W + K) * p.DT:.2f})")
    log("[bench] plain FP64 path (the reference's GPU path: ordered pair sum, Anderson)")
    r64 = run_path(ws, p, v, w, collop.PREC_F64, W, K, symmetric=False, device_solve=False)
    log(f"[bench] plain FP64: {r64['steps_per_s']:.3f} steps/s, mean iters {r64['mean_iters']:.1f}")
    pch = collop.chosen_config(args.config)
    from collop import warmstart as W_
    W_.load_packaged(ws)  # the learned warm start (trained by --report warmstart on other seeds)
    warm = collop.WARM_EULER if args.no_mlp else collop.WARM_MLP
    log("[bench] chosen path")
    rch = run_path(ws, pch, v, w, collop.PREC_F32C, W, K, warmstart=warm)
    log(f"[bench] chosen: {rch['steps_per_s']:.3f} steps/s, mean iters {rch['mean_iters']:.1f} "
        f"(f32 {rch['mean_iters_f32']:.1f}), constraint met: {rch['constraint_met']}")

    comps = {}
    if not args.no_variants:
        variants = [
            ("fp64_symmetric", "plain FP64 Anderson with the symmetric pair sum", p, collop.PREC_F64, None,
             collop.WARM_EULER),
            ("anderson_f32c", collop.anderson_description(), collop.anderson_config(args.config), collop.PREC_F32C,
             None, warm),
            ("anderson_clusters", "the Anderson path with stiff-cluster elimination (clusters of "
             f"max(w_i, w_j) / |d| > {collop.cluster_config(args.config).cl_theta:g} solved exactly in every FP64 map)",
             collop.cluster_config(args.config), collop.PREC_F32C, None, warm),
            ("defect_correction", "the chosen path without the hand-over to Anderson (step_defect! alone, 60 inner "
             "maps per outer iteration)", collop.defect_config(args.config), collop.PREC_F32C, None, warm),
        ]
        if warm == collop.WARM_MLP:
            variants.append(("chosen_euler_predictor", "the chosen path with the Euler predictor only", pch,
                             collop.PREC_F32C, None, collop.WARM_EULER))
        else:
            variants.append(("chosen_mlp_warm_start", "the chosen path with the learned MLP warm start", pch,
                             collop.PREC_F32C, None, collop.WARM_MLP))
        for key, what, pp, prec, solver, wm in variants:
            log(f"[bench] comparison path: {what}")
            comps[key] = (what, run_path(ws, pp, v, w, prec, W, K, solver=solver, warmstart=wm))
            log(f"[bench] {key}: {comps[key][1]['steps_per_s']:.3f} steps/s, mean iters "
                f"{comps[key][1]['mean_iters']:.1f}, constraint met: {comps[key][1]['constraint_met']}")

    # roofline of the dominant pair kernel of the chosen path's timed region.
    # Algorithmic work: 23 flops per ordered in-domain pair (functions.jl:221-231),
    # N_in (N_in - 1) ordered pairs per launch for either form; the metric kernels
    # add 15 per ordered pair (functions.jl:276-281).
    pairs = n_in * (n_in - 1)
    pow2w = equal_pow2_weights(w)
    kinds = []
    for path, res in [("chosen", rch), ("plain_fp64", r64)] + [(k, r) for k, (_, r) in comps.items()]:
        for k, (kname, peak, fpp_eval, unit) in enumerate(pair_kernels(res["tiles"], pow2w)):
            if res["pair_launches"][k]:
                avg = res["pair_ms"][k] / res["pair_launches"][k]
                fpp = 38 if "metric" in kname else 23
                kinds.append({"path": path, "kernel": kname, "launches": res["pair_launches"][k],
                              "total_ms": round(res["pair_ms"][k], 3), "avg_ms": round(avg, 4),
                              "tflops": round(fpp * pairs / (avg * 1e-3) / 1e12, 3),
                              "frac": round(fpp * pairs / (avg * 1e-3) / 1e12 / peak, 4),
                              "ordered_pairs_per_s": round(pairs / (avg * 1e-3), 1)})
    dom = int(np.argmax(rch["pair_ms"]))
    kname, peak, fpp_eval, unit = pair_kernels(rch["tiles"], pow2w)[dom]
    avg_ms = rch["pair_ms"][dom] / max(rch["pair_launches"][dom], 1)
    fpp = 38 if "metric" in kname else 23
    flops = fpp * pairs
    achieved = flops / (avg_ms * 1e-3) / 1e12
    units = {"ordered": pairs, "unordered": pairs // 2}
    traffic = PAIR_TRAFFIC_BYTES.get(kname) if args.config == "bimodal_n65536" else None
    roofline = {"bound": "valu", "kernel": kname, "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": traffic, "avg_launch_ms": round(avg_ms, 4),
                "flops_per_ordered_pair": fpp, "ordered_pairs_per_launch": pairs, "flops_per_launch": flops,
                "as_evaluated": {"pair_unit": unit, "flops_per_pair": fpp_eval, "pairs_per_launch": units[unit],
                                 "tflops": round(fpp_eval * units[unit] / (avg_ms * 1e-3) / 1e12, 3)},
                "pair_kernels": kinds}

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(p, v, w, r64["mean_iters"] + 1.0)

    ratio = rch["steps_per_s"] / r64["steps_per_s"]
    # like for like: the timed steps that both paths solve to the tolerance
    both = [i for i, (a, b) in enumerate(zip(rch["stats"], r64["stats"]))
            if a["exit_reason"] == 0 and b["exit_reason"] == 0]
    t_both_ch = sum(rch["stats"][i]["seconds"] for i in both)
    t_both_64 = sum(r64["stats"][i]["seconds"] for i in both)
    per_step = {k: {"ms": [round(st["seconds"] * 1e3, 1) for st in r["stats"]],
                    "maps": [st["iters"] for st in r["stats"]],
                    "exit": [st["exit_reason"] for st in r["stats"]],
                    "handover": [st["handover"] for st in r["stats"]]}
                for k, r in [("chosen", rch), ("plain_fp64", r64)] + [(k, r) for k, (_, r) in comps.items()]}
    converged_both = {"steps": [W + i for i in both],
                      "chosen_steps_per_s": round(len(both) / t_both_ch, 4) if both else None,
                      "plain_fp64_steps_per_s": round(len(both) / t_both_64, 4) if both else None,
                      "ratio": round(t_both_64 / t_both_ch, 4) if both else None,
                      "note": "step times (the library's own clock per step) over the timed steps that converged "
                              "on both paths"}
    dvm = np.max(np.abs(rch["markers"] - r64["markers"]), axis=1)
    out = {
        "metric": "Landau implicit steps/s at N=65536, chosen vs plain FP64 GPU, |ΔE|/E≤1e-10",
        "value": round(rch["steps_per_s"], 4) if rch["constraint_met"] else None,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 / rch["steps_per_s"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32c+f64",
        "data": "synthetic (seeded bimodal Maxwellian markers, PCG64 seed 42)",
        "config": {"workload": args.config, "N": N, "mesh": f"{len(p.bp1) - 1}x{len(p.bp2) - 1}", "dt": p.DT,
                   "m_anderson": p.m_anderson, "tol": p.tol, "abs_floor": p.abs_floor,
                   "timed_steps": [W, W + K - 1]},
        "constraint_met": rch["constraint_met"],
        "constraint": "every timed step converged (exit 0), per-step |E_n+1 - E_n| / E_0 <= 1e-10, "
                      "|P_n+1 - P_n| <= 1e-10, S_n+1 >= S_n; value is null when it fails",
        "value_unconstrained": round(rch["steps_per_s"], 4),
        "fp64_steps_per_s": round(r64["steps_per_s"], 4),
        "ratio_vs_fp64": round(ratio, 4),
        "converged_on_both_paths": converged_both,
        "per_step": per_step,
        "chosen_path": collop.chosen_description(args.config, warm == collop.WARM_MLP),
        "plain_path": "FP64 maps, Anderson (step_anderson!) controlled from the host, Euler predictor, ordered pair "
                      "sum",
        "chosen": path_summary(rch, r64),
        "plain_fp64": path_summary(r64),
        "comparisons": {k: dict(path_summary(r, r64), path=what) for k, (what, r) in comps.items()},
        "chosen_vs_fp64_dv": {
            "max_abs": float(np.max(dvm)),
            "markers_differing_above_1e-6": float(np.mean(dvm > 1e-6)),
            "relative_l2": float(np.linalg.norm(rch["markers"] - r64["markers"]) / np.linalg.norm(r64["markers"])),
            "why": "at dt 0.02 the implicit systems have several roots around close pairs; paths that converge end "
                   "each step on a root within the FP64 tolerance, not always the same one, and a stagnated step "
                   "ends on its best iterate instead",
        },
        "pair_interactions_per_s": round(pairs / (r64["pair_ms"][0] / max(r64["pair_launches"][0], 1) * 1e-3), 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
