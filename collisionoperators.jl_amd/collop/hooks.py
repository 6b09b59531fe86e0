"""The reference's GPU hook functions, served by libcollop.so.

Each function has the name and argument order of the CUDA.jl routine it
replaces (without Julia's ``!``) and writes its result in place, like the
reference hooks of solver.jl:30-34 that ``enable_gpu!`` (main.jl:390-423)
installs.  Marker arrays are ``(N, 2)`` float64; any memory order is accepted
(column-major, Julia's layout, avoids a copy).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .mesh import Workspace


def _colmajor(a, shape=None) -> np.ndarray:
    a = np.asfortranarray(np.asarray(a, dtype=np.float64))
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _vec(a, n=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if n is not None and a.shape != (n,):
        raise ValueError(f"expected shape ({n},), got {a.shape}")
    return a


def _markers(v) -> tuple[np.ndarray, int]:
    v = _colmajor(v)
    if v.ndim != 2 or v.shape[1] != 2:
        raise ValueError(f"markers must be (N, 2), got {v.shape}")
    return v, v.shape[0]


def _out(dst: np.ndarray, ncol: int, N: int):
    if dst.shape != (N, ncol) or dst.dtype != np.float64:
        raise ValueError(f"output must be float64 ({N}, {ncol}), got {dst.dtype} {dst.shape}")
    return dst if dst.flags.f_contiguous else np.empty((N, ncol), order="F")


def _collision(fn, ws: Workspace, dot_v, v_parts, w_parts, G):
    ws.activate()
    v, N = _markers(v_parts)
    w = _vec(w_parts, N)
    G = _colmajor(G, (N, 2))
    tmp = _out(dot_v, 2, N)
    ws.ctx.check(fn(ws.ctx.h, _lib.dptr(v), _lib.dptr(w), _lib.dptr(G), _lib.dptr(tmp), N))
    if tmp is not dot_v:
        dot_v[...] = tmp
    return None


def compute_collision_gpu(ws: Workspace, dot_v, v_parts, w_parts, G):
    """COLLISION_FN, FP64 (collision_gpu.jl:258-280; semantics functions.jl:205-237)."""
    return _collision(ws.ctx.lib.collop_collision_f64, ws, dot_v, v_parts, w_parts, G)


def compute_collision_gpu32(ws: Workspace, dot_v, v_parts, w_parts, G):
    """COLLISION_FN with gpu_fp32 (collision_gpu.jl:233-256): FP32 pair math,
    two-sum compensated per-marker sums, FP64 result."""
    return _collision(ws.ctx.lib.collop_collision_f32c, ws, dot_v, v_parts, w_parts, G)


def compute_collision_metric_gpu(ws: Workspace, dot_v, A, v_parts, w_parts, G, fp32c: bool = False):
    """COLLMETRIC_FN (collision_gpu.jl:377-402; semantics functions.jl:249-290).
    ``fp32c``: FP32 pair terms with compensated sums (collop_collision_metric_f32c)."""
    ws.activate()
    v, N = _markers(v_parts)
    w = _vec(w_parts, N)
    G = _colmajor(G, (N, 2))
    t1 = _out(dot_v, 2, N)
    t2 = _out(A, 3, N)
    fn = ws.ctx.lib.collop_collision_metric_f32c if fp32c else ws.ctx.lib.collop_collision_metric_f64
    ws.ctx.check(fn(ws.ctx.h, _lib.dptr(v), _lib.dptr(w), _lib.dptr(G), _lib.dptr(t1), _lib.dptr(t2), N))
    if t1 is not dot_v:
        dot_v[...] = t1
    if t2 is not A:
        A[...] = t2


def l2_rhs_gpu(ws: Workspace, rhs, v_parts, w_parts):
    """Scatter only: rhs_k = sum w phi_k(v) over in-domain markers (functions.jl:121-131)."""
    ws.activate()
    v, N = _markers(v_parts)
    w = _vec(w_parts, N)
    out = np.empty(ws.n_dofs)
    ws.ctx.check(ws.ctx.lib.collop_l2_rhs(ws.ctx.h, _lib.dptr(v), _lib.dptr(w), N, _lib.dptr(out)))
    rhs[...] = out


def l2_project_gpu(ws: Workspace, f_coeffs, v_parts, w_parts):
    """L2PROJ_FN (projection_gpu.jl:213-226): scatter plus the mass solve, both on the device."""
    ws.activate()
    v, N = _markers(v_parts)
    w = _vec(w_parts, N)
    out = np.empty(ws.n_dofs)
    ws.ctx.check(ws.ctx.lib.collop_l2_project(ws.ctx.h, _lib.dptr(v), _lib.dptr(w), N, _lib.dptr(out)))
    f_coeffs[...] = out


def compute_G_gpu(ws: Workspace, G, v_parts, L_vec):
    """COMPG_FN (projection_gpu.jl:228-246; semantics functions.jl:182-202)."""
    ws.activate()
    v, N = _markers(v_parts)
    L = _vec(L_vec, ws.n_dofs)
    tmp = _out(G, 2, N)
    ws.ctx.check(ws.ctx.lib.collop_grad_gather(ws.ctx.h, _lib.dptr(v), N, _lib.dptr(L), _lib.dptr(tmp)))
    if tmp is not G:
        G[...] = tmp


def eval_loggrad_gpu(ws: Workspace, g, v_parts, f_coeffs, fs_floor: float = 1e-30, g_max: float = 100.0):
    """LOGGRAD_FN (projection_gpu.jl:251-269; semantics functions.jl:37-64)."""
    ws.activate()
    v, N = _markers(v_parts)
    fc = _vec(f_coeffs, ws.n_dofs)
    tmp = _out(g, 2, N)
    ws.ctx.check(ws.ctx.lib.collop_loggrad_gather(ws.ctx.h, _lib.dptr(v), N, _lib.dptr(fc), fs_floor, g_max,
                                                  _lib.dptr(tmp)))
    if tmp is not g:
        g[...] = tmp


def compute_r_and_entropy_gpu(ws: Workspace, f_coeffs, use_logsq: bool = False):
    """compute_r! and compute_entropy (functions.jl:137-179) on the device; returns (r, S)."""
    ws.activate()
    fc = _vec(f_coeffs, ws.n_dofs)
    r = np.empty(ws.n_dofs)
    S = np.zeros(1)
    ws.ctx.check(ws.ctx.lib.collop_quadrature(ws.ctx.h, _lib.dptr(fc), int(use_logsq), _lib.dptr(r), _lib.dptr(S)))
    return r, float(S[0])


def mass_solve_gpu(ws: Workspace, b):
    """x = M^{-1} b (ws.M_lu \\ b in the reference)."""
    ws.activate()
    b = _vec(b, ws.n_dofs)
    x = np.empty(ws.n_dofs)
    ws.ctx.check(ws.ctx.lib.collop_mass_solve(ws.ctx.h, _lib.dptr(b), _lib.dptr(x)))
    return x


def compute_lb_rhs_gpu(ws: Workspace, dot_v, v_parts, w_parts, g, nu: float = 1.0):
    """compute_moments + compute_drift_multipliers + compute_LB_velocity!
    (functions.jl:66-115) on the device: writes dot_v = -nu (g + A + B v) and
    returns (A1, A2, B)."""
    ws.activate()
    v, N = _markers(v_parts)
    w = _vec(w_parts, N)
    g = _colmajor(g, (N, 2))
    tmp = _out(dot_v, 2, N)
    mult = np.zeros(3)
    ws.ctx.check(ws.ctx.lib.collop_lb_rhs(ws.ctx.h, _lib.dptr(v), _lib.dptr(w), _lib.dptr(g), N, float(nu),
                                          _lib.dptr(tmp), _lib.dptr(mult)))
    if tmp is not dot_v:
        dot_v[...] = tmp
    return float(mult[0]), float(mult[1]), float(mult[2])


def gmres_dense_gpu(ctx, A, b, rtol: float, atol: float = 0.0, itmax: int | None = None, mem: int | None = None):
    """gmres! (newton_krylov.jl:60-131) of the device Newton-Krylov engine on a
    dense operator: (x, products, GMRES residual norm)."""
    import ctypes as C
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = _vec(b)
    n = b.shape[0]
    mem = min(n, 64) if mem is None else mem
    itmax = mem if itmax is None else itmax
    x = np.zeros(n)
    k = C.c_int(0)
    r = np.zeros(1)
    ctx.check(ctx.lib.collop_gmres_dense(ctx.h, n, _lib.dptr(A), _lib.dptr(b), float(rtol), float(atol), int(itmax),
                                         int(mem), _lib.dptr(x), C.byref(k), _lib.dptr(r)))
    return x, int(k.value), float(r[0])


def compute_negative_part_l1_gpu(ws: Workspace, f_coeffs) -> float:
    """compute_negative_part_l1 (functions.jl:300-313) of f_s = sum_k f_coeffs[k] phi_k."""
    ws.activate()
    fc = _vec(f_coeffs, ws.n_dofs)
    out = np.zeros(1)
    ws.ctx.check(ws.ctx.lib.collop_negative_part_l1(ws.ctx.h, _lib.dptr(fc), _lib.dptr(out)))
    return float(out[0])


def compute_fs_minus_fp_l2_gpu(ws: Workspace, f_coeffs, v_parts, w_parts) -> float:
    """compute_fs_minus_fp_l2 (functions.jl:320-343): ||f_s - f_p||_2 with f_p the
    markers' cellwise histogram density."""
    ws.activate()
    fc = _vec(f_coeffs, ws.n_dofs)
    v, N = _markers(v_parts)
    w = _vec(w_parts, N)
    out = np.zeros(1)
    ws.ctx.check(ws.ctx.lib.collop_fs_minus_fp_l2(ws.ctx.h, _lib.dptr(v), _lib.dptr(w), N, _lib.dptr(fc),
                                                  _lib.dptr(out)))
    return float(out[0])


def enable_gpu(hooks: dict, gpu_fp32: bool = False) -> dict:
    """Mirror of enable_gpu! (main.jl:390-423): repoint a hook table at the device routines."""
    hooks["L2PROJ_FN"] = l2_project_gpu
    hooks["COMPG_FN"] = compute_G_gpu
    hooks["LOGGRAD_FN"] = eval_loggrad_gpu
    hooks["COLLISION_FN"] = compute_collision_gpu32 if gpu_fp32 else compute_collision_gpu
    hooks["COLLMETRIC_FN"] = compute_collision_metric_gpu
    return hooks
