"""ctypes binding of ``libcollop.so`` (the C-ABI declared in include/collop.h).

The shared library is the product; this module only loads it and mirrors its
structs.  There is no fallback: if the library is missing, or no GPU is
visible when a context is created, the call raises :class:`CollopError`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcollop.so")

OK = 0
ERR_INVALID = 1
ERR_UNSUPPORTED = 2
ERR_HIP = 3
ERR_NO_MESH = 4
ERR_NONFINITE = 5
ERR_NO_STATE = 6
ERR_NO_MODEL = 7

PREC_F64 = 0
PREC_F32C = 1
WARM_EULER = 0
WARM_MLP = 1
SOLVER_ANDERSON = 0
SOLVER_DEFECT = 1
SOLVER_NEWTON = 2
MODEL_LANDAU = 0
MODEL_LB = 1
PAIR_ORDERED = 0
PAIR_SYMMETRIC = 1


class CollopError(RuntimeError):
    """A non-zero collop_status, with the library's message."""

    def __init__(self, status: int, message: str):
        super().__init__(f"collop status {status}: {message}")
        self.status = status
        self.message = message


class StepParams(C.Structure):
    """collop_step_params (include/collop.h); fields as Parameters.jl:65-106."""

    _fields_ = [
        ("dt", C.c_double),
        ("m_anderson", C.c_int32),
        ("max_iter", C.c_int32),
        ("tol", C.c_double),
        ("abs_floor", C.c_double),
        ("stag_window", C.c_int32),
        ("stag_rel_tol", C.c_double),
        ("damp_decay_start", C.c_int32),
        ("damp_decay_factor", C.c_double),
        ("restart_factor", C.c_double),
        ("damping", C.c_double),
        ("reg_factor", C.c_double),
        ("use_anderson", C.c_int32),
        ("use_gonzalez", C.c_int32),
        ("exit_picard_step", C.c_int32),
        ("use_logsq", C.c_int32),
        ("precision", C.c_int32),
        ("warmstart", C.c_int32),
        ("f32_floor_factor", C.c_double),
        ("f32_keep_window", C.c_int32),
        ("record_training", C.c_int32),
        ("solver", C.c_int32),
        ("dc_eta", C.c_double),
        ("dc_max_inner", C.c_int32),
        ("dc_stag_window", C.c_int32),
        ("dc_outer_m", C.c_int32),
        ("dc_near_theta", C.c_double),
        ("dc_metric_reuse", C.c_int32),
        ("collision_model", C.c_int32),
        ("nu", C.c_double),
        ("dc_eta_f64", C.c_double),
        ("nk_krylov_max", C.c_int32),
        ("nk_eta_max", C.c_double),
        ("nk_fd_h", C.c_double),
        ("nk_fd_rel", C.c_double),
        ("cl_theta", C.c_double),
        ("cl_kmax", C.c_int32),
        ("dc_fallback", C.c_int32),
        ("dc_sticky", C.c_int32),
    ]


class StepStats(C.Structure):
    """collop_step_stats (include/collop.h)."""

    _fields_ = [
        ("iters", C.c_int32),
        ("iters_f32", C.c_int32),
        ("n_restart", C.c_int32),
        ("exit_reason", C.c_int32),
        ("residual", C.c_double),
        ("r0", C.c_double),
        ("entropy", C.c_double),
        ("energy", C.c_double),
        ("momentum", C.c_double * 2),
        ("seconds", C.c_double),
        ("inner_maps", C.c_int32),
        ("n_clusters", C.c_int32),
        ("cl_fail", C.c_int32),
        ("stat_syncs", C.c_int32),
        ("handover", C.c_int32),
        ("dc_skipped", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {
            "iters": self.iters, "iters_f32": self.iters_f32, "n_restart": self.n_restart,
            "exit_reason": self.exit_reason, "residual": self.residual, "r0": self.r0,
            "entropy": self.entropy, "energy": self.energy,
            "momentum": (self.momentum[0], self.momentum[1]), "seconds": self.seconds,
            "inner_maps": self.inner_maps, "n_clusters": self.n_clusters, "cl_fail": self.cl_fail,
            "stat_syncs": self.stat_syncs, "handover": self.handover, "dc_skipped": self.dc_skipped,
        }


_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_vp = C.c_void_p

# (name, restype, argtypes) for every entry point of include/collop.h
SIGNATURES = [
    ("collop_version", C.c_int, []),
    ("collop_create", C.c_int, [C.c_int, C.POINTER(_vp)]),
    ("collop_destroy", C.c_int, [_vp]),
    ("collop_last_error", C.c_char_p, [_vp]),
    ("collop_set_mesh", C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, _dp, _dp, _i32p, _i32p, C.c_int, C.c_int,
                                  C.c_int, C.c_int]),
    ("collop_collision_f64", C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_int64]),
    ("collop_collision_f32c", C.c_int, [_vp, _dp, _dp, _dp, _dp, C.c_int64]),
    ("collop_collision_metric_f64", C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, C.c_int64]),
    ("collop_collision_metric_f32c", C.c_int, [_vp, _dp, _dp, _dp, _dp, _dp, C.c_int64]),
    ("collop_collision_f64_noalloc", C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int64, _vp]),
    ("collop_l2_rhs", C.c_int, [_vp, _dp, _dp, C.c_int64, _dp]),
    ("collop_l2_project", C.c_int, [_vp, _dp, _dp, C.c_int64, _dp]),
    ("collop_grad_gather", C.c_int, [_vp, _dp, C.c_int64, _dp, _dp]),
    ("collop_loggrad_gather", C.c_int, [_vp, _dp, C.c_int64, _dp, C.c_double, C.c_double, _dp]),
    ("collop_quadrature", C.c_int, [_vp, _dp, C.c_int, _dp, _dp]),
    ("collop_mass_solve", C.c_int, [_vp, _dp, _dp]),
    ("collop_lb_rhs", C.c_int, [_vp, _dp, _dp, _dp, C.c_int64, C.c_double, _dp, _dp]),
    ("collop_gmres_dense", C.c_int, [_vp, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int, C.c_int, _dp,
                                     C.POINTER(C.c_int), _dp]),
    ("collop_negative_part_l1", C.c_int, [_vp, _dp, _dp]),
    ("collop_fs_minus_fp_l2", C.c_int, [_vp, _dp, _dp, C.c_int64, _dp, _dp]),
    ("collop_state_set", C.c_int, [_vp, _dp, _dp, C.c_int64]),
    ("collop_state_get", C.c_int, [_vp, _dp]),
    ("collop_state_coeffs", C.c_int, [_vp, _dp]),
    ("collop_state_diagnostics", C.c_int, [_vp, _dp]),
    ("collop_run_steps", C.c_int, [_vp, C.POINTER(StepParams), C.c_int, C.POINTER(StepStats)]),
    ("collop_step", C.c_int, [_vp, C.POINTER(StepParams), _dp, _dp, C.c_int64, C.POINTER(StepStats)]),
    ("collop_mlp_set_params", C.c_int, [_vp, _fp, C.c_int]),
    ("collop_mlp_get_params", C.c_int, [_vp, _fp, C.c_int]),
    ("collop_mlp_set_norm", C.c_int, [_vp, _fp, _fp, _fp]),
    ("collop_mlp_dataset_alloc", C.c_int, [_vp, C.c_int64]),
    ("collop_mlp_dataset_size", C.c_int, [_vp, C.POINTER(C.c_int64)]),
    ("collop_mlp_dataset_read", C.c_int, [_vp, C.c_int64, C.c_int64, _fp]),
    ("collop_mlp_train", C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_uint64, _dp]),
    ("collop_mlp_eval", C.c_int, [_vp, _fp, C.c_int64, C.c_int, _fp, _dp, _fp]),
    ("collop_ensemble_set", C.c_int, [_vp, C.c_int, C.POINTER(C.c_int64), _dp, _dp]),
    ("collop_ensemble_get", C.c_int, [_vp, _dp]),
    ("collop_ensemble_coeffs", C.c_int, [_vp, _dp]),
    ("collop_ensemble_step", C.c_int, [_vp, C.POINTER(StepParams), C.c_int, C.POINTER(StepStats)]),
    ("collop_ensemble_slices", C.c_int, [C.c_int64]),
    ("collop_ensemble_set_tile", C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    ("collop_ensemble_tune", C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]),
    ("collop_set_pair_tile", C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    ("collop_set_pair_mode", C.c_int, [_vp, C.c_int, C.c_int]),
    ("collop_set_fused_maps", C.c_int, [_vp, C.c_int]),
    ("collop_set_device_solve", C.c_int, [_vp, C.c_int]),
    ("collop_set_weight_path", C.c_int, [_vp, C.c_int]),
    ("collop_get_pair_mode", C.c_int, [_vp, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    ("collop_profile", C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_long)]),
]

_LIB = None


def load(path: str | None = None):
    """Load libcollop.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise CollopError(ERR_HIP, f"{p} not found: build it with collisionoperators.jl_amd/build_lib.py")
    lib = C.CDLL(p)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32
    return a.ctypes.data_as(_fp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_i32p)


class Context:
    """Owns one collop_ctx on one device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = _vp()
        st = self.lib.collop_create(device, C.byref(h))
        if st != OK:
            raise CollopError(st, f"collop_create(device={device}) failed: no usable HIP device")
        self.h = h
        self.device = device

    def check(self, st: int):
        if st != OK:
            msg = self.lib.collop_last_error(self.h)
            raise CollopError(st, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.collop_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
