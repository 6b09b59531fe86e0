"""SimParameters: the fields of the reference's configuration struct
(Parameters.jl:18-140) that the Landau hot path reads, plus the named
workloads this library is measured on.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib


def _uniform(lo: float, hi: float, n: int) -> np.ndarray:
    # lo + (hi - lo) i / n: exact for the power-of-two cell counts used here
    return np.array([lo + (hi - lo) * i / n for i in range(n + 1)], dtype=np.float64)


@dataclass
class SimParameters:
    """Subset of Parameters.jl:18-140 (same names and defaults unless noted)."""

    bp1: np.ndarray = field(default_factory=lambda: _uniform(-6.0, 6.0, 16))
    bp2: np.ndarray = field(default_factory=lambda: _uniform(-6.0, 6.0, 16))
    P_DEG: int = 2
    K_REG: int = 1
    N_QUAD: int = 6
    N_PARTICLES: int = 10_000
    σ1: float = 4 / 3
    σ2: float = 0.5
    v1_peak: float = 0.0
    DT: float = 0.001
    N_STEPS: int = 400
    use_gonzalez: bool = True
    use_logsq: bool = False
    use_anderson: bool = True
    damping: float = 0.7
    m_anderson: int = 8
    tol: float = 1e-12
    max_iter: int = 2000
    abs_floor: float = 1e-7
    stag_window: int = 50
    stag_rel_tol: float = 0.01
    damp_decay_start: int = 200
    damp_decay_factor: float = 0.5
    exit_picard_step: bool = True
    warmstart: str = "euler"        # :euler | :nn  (Parameters.jl:111)
    use_gpu: bool = True
    gpu_fp32: bool = False           # FP32-compensated pair sum for the early maps
    seed: int = 42
    ic: str = "gaussian"             # gaussian | bimodal | bimaxwellian
    weight_jitter: float = 0.0       # w ∝ 1 + jitter u, u ~ U(-1, 1)
    solver: str = "anderson"         # :anderson | :defect | :newton  (Parameters.jl:78-88)
    dc_eta: float = 0.05             # Parameters.jl:89-91
    dc_max_inner: int = 60
    dc_stag_window: int = 5
    dc_outer_m: int = 0              # extension: outer Anderson window (0 = the reference's outer iteration)
    dc_near_theta: float = 0.0       # extension: exact near pairs in the frozen map (0 = none)
    dc_metric_reuse: int = 0         # extension: 0 A at every outer map, 1 FP64 maps reuse the FP32c A, 2 first only
    dc_eta_f64: float = 0.0          # extension: inner tolerance of the FP64 outer iterations (0 = dc_eta)
    collision_model: str = "landau"  # :landau | :lb  (Parameters.jl:50)
    nu: float = 1.0                  # LB collision frequency (Parameters.jl:53)
    nk_krylov_max: int = 30          # Parameters.jl:92-100 (solver = :newton)
    nk_eta_max: float = 0.9
    nk_fd_h: float = 0.0             # 0 = 1e-6 (FP64 maps)
    nk_fd_rel: float = 0.0
    cl_theta: float = 0.0            # extension: stiff clusters max(w_i, w_j) / |d| > cl_theta solved exactly (0 = off)
    cl_kmax: int = 32                # largest stiff cluster (markers)
    dc_fallback: bool = False        # extension: defect correction hands a stiff step to the Anderson solve
    dc_sticky: bool = False          # extension: after an early hand-over, the next steps start with Anderson

    def step_params(self, precision: int | None = None, warmstart: int | None = None,
                    solver: int | None = None) -> _lib.StepParams:
        """The C step parameters; precision / warmstart / solver override the fields."""
        prec = precision if precision is not None else (_lib.PREC_F32C if self.gpu_fp32 else _lib.PREC_F64)
        warm = warmstart if warmstart is not None else (_lib.WARM_MLP if self.warmstart == "nn" else _lib.WARM_EULER)
        if solver is None:
            solvers = {"anderson": _lib.SOLVER_ANDERSON, "defect": _lib.SOLVER_DEFECT, "newton": _lib.SOLVER_NEWTON}
            if self.solver not in solvers:
                raise ValueError(f"unknown solver={self.solver} (anderson | defect | newton)")
            solver = solvers[self.solver]
        if self.collision_model not in ("landau", "lb"):
            raise ValueError(f"unknown collision_model={self.collision_model} (landau | lb)")
        return _lib.StepParams(
            dt=self.DT, m_anderson=self.m_anderson, max_iter=self.max_iter, tol=self.tol,
            abs_floor=self.abs_floor, stag_window=self.stag_window, stag_rel_tol=self.stag_rel_tol,
            damp_decay_start=self.damp_decay_start, damp_decay_factor=self.damp_decay_factor,
            restart_factor=math.inf, damping=self.damping, reg_factor=1e-10,
            use_anderson=int(self.use_anderson), use_gonzalez=int(self.use_gonzalez),
            exit_picard_step=int(self.exit_picard_step), use_logsq=int(self.use_logsq),
            precision=prec, warmstart=warm, f32_floor_factor=16.0, f32_keep_window=1, record_training=0,
            solver=solver, dc_eta=self.dc_eta, dc_max_inner=self.dc_max_inner, dc_stag_window=self.dc_stag_window,
            dc_outer_m=self.dc_outer_m, dc_near_theta=self.dc_near_theta, dc_metric_reuse=self.dc_metric_reuse,
            collision_model=_lib.MODEL_LB if self.collision_model == "lb" else _lib.MODEL_LANDAU, nu=self.nu,
            dc_eta_f64=self.dc_eta_f64, nk_krylov_max=self.nk_krylov_max, nk_eta_max=self.nk_eta_max, nk_fd_h=self.nk_fd_h,
            nk_fd_rel=self.nk_fd_rel, cl_theta=self.cl_theta, cl_kmax=self.cl_kmax,
            dc_fallback=int(self.dc_fallback), dc_sticky=int(self.dc_sticky))


def parse_overrides(p: SimParameters, args: list[str]) -> SimParameters:
    """``--key=value`` scalar overrides (Parameters.jl:157-192): unknown keys and
    vector fields raise, untouched fields survive."""
    kw = {}
    fields = {f.name: f for f in dataclasses.fields(p)}
    for tok in args:
        if not tok.startswith("--") or "=" not in tok:
            raise ValueError(f"override must look like --key=value: {tok!r}")
        k, v = tok[2:].split("=", 1)
        if k not in fields:
            raise ValueError(f"no such field: {k}")
        cur = getattr(p, k)
        if isinstance(cur, np.ndarray):
            raise ValueError(f"vector field {k} cannot be overridden on the command line")
        if isinstance(cur, bool):
            if v.lower() not in ("true", "false"):
                raise ValueError(f"{k} expects true/false")
            kw[k] = v.lower() == "true"
        elif isinstance(cur, int):
            kw[k] = int(v)
        elif isinstance(cur, float):
            kw[k] = float(v)
        else:
            kw[k] = v.lstrip(":")
    return dataclasses.replace(p, **kw)


# Solver knobs pinned for every path (the Landau presets, parameters_sq_d04.jl:16-20):
_KNOBS = dict(damping=0.7, abs_floor=1e-10, stag_window=30, stag_rel_tol=0.1, damp_decay_start=200,
              damp_decay_factor=0.5, max_iter=2000, N_QUAD=6, K_REG=1, use_logsq=False, use_gonzalez=True,
              exit_picard_step=True, m_anderson=5, tol=1e-12)

# The named workloads (N markers, IC, mesh, dt, steps).
CONFIGS = {
    # isotropic N(0, I), 16x16 cells on [-6,6]^2, dt 0.05, 10 steps
    "gaussian_n4096": dict(N_PARTICLES=4096, ic="gaussian", σ1=1.0, σ2=1.0, bp1=_uniform(-6, 6, 16),
                           bp2=_uniform(-6, 6, 16), DT=0.05, N_STEPS=10, seed=42),
    # two Maxwellians at (+-2, 0) with covariance 0.5 I, 32x32 cells on [-6,6]^2, dt 0.02, 50 steps
    "bimodal_n65536": dict(N_PARTICLES=65536, ic="bimodal", v1_peak=2.0, σ1=math.sqrt(0.5), σ2=math.sqrt(0.5),
                           bp1=_uniform(-6, 6, 32), bp2=_uniform(-6, 6, 32), DT=0.02, N_STEPS=50, seed=42),
    # bi-Maxwellian T_x = 2, T_y = 0.5, w ∝ 1 + 0.2 u; 64x64 cells on [-8,8]^2, dt 0.01, 20 steps
    "bimaxwellian_n262144": dict(N_PARTICLES=262144, ic="bimaxwellian", σ1=math.sqrt(2.0), σ2=math.sqrt(0.5),
                                 weight_jitter=0.2, bp1=_uniform(-8, 8, 64), bp2=_uniform(-8, 8, 64), DT=0.01,
                                 N_STEPS=20, seed=42),
    # Lenard-Bernstein on the bimodal IC of parameters_LB_bimodal_v1.jl: N(+-2, 1) x N(0, 1), nu = 1,
    # non-uniform breakpoints (0.4 cells over [-5, 5] in v1, [-4.5, 4.5] in v2, end cells to +-6), dt 0.001, m 8
    "lb_bimodal_n40000": dict(N_PARTICLES=40_000, ic="bimodal", v1_peak=2.0, σ1=1.0, σ2=1.0,
                              bp1=np.r_[-6.0, np.linspace(-5.0, 5.0, 26), 6.0],
                              bp2=np.r_[-6.0, np.arange(-4.5, 4.5 + 1e-9, 0.4), 6.0],
                              collision_model="lb", nu=1.0, DT=0.001, N_STEPS=25000, m_anderson=8, seed=42),
}


# The defect-correction path: FP32-compensated outer maps then FP64 outer maps,
# frozen-metric defect correction (step_defect!) with the extensions of this
# library (an outer Anderson window of 3 and exact near pairs with w / |d| >
# 0.02), the FP64 outer maps reusing the metric A of the last FP32c outer map.
DEFECT = dict(gpu_fp32=True, solver="defect", dc_outer_m=3, dc_near_theta=0.02, dc_metric_reuse=1)

# The Anderson path: FP32c maps until their rounding floor, then FP64 maps, the
# Anderson solve of step_anderson! (the workload's window of 5) with its
# control on the device, symmetric pair sum.
ANDERSON_F32C = dict(gpu_fp32=True, solver="anderson")
# The chosen path: the defect-correction path (DEFECT) with the hand-over: when
# a frozen-metric inner solve ends unconverged, an outer iteration does not
# reduce the residual by 0.4, or the outer iteration stagnates (stiff near
# pairs), the step continues with the Anderson solve in FP64 maps.  After a
# hand-over within the first 2 outer iterations, the next steps start with the
# Anderson solve until one converges (dc_sticky).
CHOSEN = dict(DEFECT, dc_fallback=True, dc_sticky=True)
# Stiff-cluster elimination (an opt-in extension): pairs with
# max(w_i, w_j) / |d| > CLUSTER_TAU / dt at the start of a step form clusters
# whose rows every FP64 map solves exactly.  The threshold keeps dt w / |d|
# fixed, the factor the stiffness of a pair scales with.
CLUSTER_TAU = 0.005
CLUSTER_KMAX = 32


def chosen_config(name: str, **over) -> SimParameters:
    """config(name) with the chosen path's solver knobs (CHOSEN), then `over`."""
    kw = dict(CHOSEN)
    kw.update(over)
    return config(name, **kw)


def anderson_config(name: str, **over) -> SimParameters:
    """config(name) with the Anderson path's knobs (ANDERSON_F32C), then `over`."""
    kw = dict(ANDERSON_F32C)
    kw.update(over)
    return config(name, **kw)


def cluster_config(name: str, **over) -> SimParameters:
    """anderson_config(name) with stiff-cluster elimination, then `over`."""
    p = anderson_config(name, cl_kmax=CLUSTER_KMAX, **over)
    if "cl_theta" not in over:
        p = dataclasses.replace(p, cl_theta=CLUSTER_TAU / p.DT)
    return p


def defect_config(name: str, **over) -> SimParameters:
    """config(name) with the defect-correction path's knobs (DEFECT), then `over`."""
    kw = dict(DEFECT)
    kw.update(over)
    return config(name, **kw)


def chosen_description(name: str, mlp: bool) -> str:
    return ("frozen-metric defect correction (step_defect!): FP32c outer maps (F and metric A) until the FP32 "
            "rounding floor, then FP64 outer maps, an outer Anderson window of 3 and exact near pairs; an inner solve "
            "that misses its tolerance in 60 maps, an outer iteration that does not reduce the residual by 0.4, or "
            "outer stagnation hands the step to the Anderson solve (step_anderson!, FP64 maps, control on the "
            "device); after a hand-over within 2 outer iterations the next steps run that Anderson solve from the "
            "start until one converges; "
            "FP32c Euler predictor" + (" + learned MLP warm start" if mlp else "")
            + "; symmetric pair sum (each unordered pair once)")


def anderson_description() -> str:
    return ("FP32c maps until the FP32 rounding floor, then FP64 maps, Anderson (step_anderson!) with its control "
            "(least squares, exit and stagnation tests) on the device, FP32c Euler predictor; symmetric pair sum")


def config(name: str, **over) -> SimParameters:
    kw = dict(_KNOBS)
    kw.update(CONFIGS[name])
    kw.update(over)
    return SimParameters(**kw)
