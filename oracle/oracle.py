"""ctypes front end of the CPU oracle (``liboracle.so`` built from collop_oracle.cpp).

TEST INFRASTRUCTURE ONLY: the oracle is the checker the HIP library is compared
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg
of ``bench.py`` import this module.  The shipped package never does.

Array conventions follow the reference (junyixu/CollisionOperators.jl): a marker
array ``v`` is an ``(N, 2)`` float64 array stored column-major in memory
(``v[:, 0]`` contiguous, then ``v[:, 1]``), exactly Julia's ``Matrix{Float64}``.
Use :func:`colmajor` to get such an array from numpy.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class SolverParams(C.Structure):
    """Mirror of ``orc_solver_params``; field meaning as in Parameters.jl:65-106."""

    _fields_ = [
        ("dt", C.c_double),
        ("m_anderson", C.c_int),
        ("max_iter", C.c_int),
        ("tol", C.c_double),
        ("abs_floor", C.c_double),
        ("stag_window", C.c_int),
        ("stag_rel_tol", C.c_double),
        ("damp_decay_start", C.c_int),
        ("damp_decay_factor", C.c_double),
        ("restart_factor", C.c_double),
        ("damping", C.c_double),
        ("reg_factor", C.c_double),
        ("use_anderson", C.c_int),
        ("use_gonzalez", C.c_int),
        ("exit_picard_step", C.c_int),
        ("use_logsq", C.c_int),
        ("nthreads", C.c_int),
        ("solver", C.c_int),
        ("dc_eta", C.c_double),
        ("dc_max_inner", C.c_int),
        ("dc_stag_window", C.c_int),
        ("dc_outer_m", C.c_int),
        ("dc_near_theta", C.c_double),
        ("collision_model", C.c_int),
        ("nu", C.c_double),
        ("nk_krylov_max", C.c_int),
        ("nk_eta_max", C.c_double),
        ("nk_fd_h", C.c_double),
        ("nk_fd_rel", C.c_double),
        ("cl_theta", C.c_double),
        ("cl_kmax", C.c_int),
        ("dc_fallback", C.c_int),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("iters", C.c_int),
        ("n_restart", C.c_int),
        ("residual", C.c_double),
        ("r0", C.c_double),
        ("exit_reason", C.c_int),
        ("entropy", C.c_double),
        ("energy", C.c_double),
        ("p1", C.c_double),
        ("p2", C.c_double),
        ("inner", C.c_int),
        ("n_clusters", C.c_int),
        ("cl_fail", C.c_int),
        ("handover", C.c_int),
    ]


def default_params(dt: float, m: int = 5, nthreads: int = 0, **kw) -> SolverParams:
    """Solver knobs pinned for every path (Landau presets, parameters_sq_d04.jl:16-20)."""
    p = SolverParams(
        dt=dt, m_anderson=m, max_iter=2000, tol=1e-12, abs_floor=1e-10,
        stag_window=30, stag_rel_tol=0.1, damp_decay_start=200,
        damp_decay_factor=0.5, restart_factor=float("inf"), damping=0.7,
        reg_factor=1e-10, use_anderson=1, use_gonzalez=1, exit_picard_step=1,
        use_logsq=0, nthreads=nthreads, solver=0, dc_eta=0.05, dc_max_inner=60, dc_stag_window=5,
        dc_outer_m=0, dc_near_theta=0.0, collision_model=0, nu=1.0,
        nk_krylov_max=30, nk_eta_max=0.9, nk_fd_h=0.0, nk_fd_rel=0.0, cl_theta=0.0, cl_kmax=32, dc_fallback=0,
    )
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_mesh_build.restype = C.c_void_p
        L.orc_mesh_build.argtypes = [_dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_mesh_free.argtypes = [C.c_void_p]
        L.orc_entropy.restype = C.c_double
        L.orc_negative_part_l1.restype = C.c_double
        L.orc_fs_minus_fp_l2.restype = C.c_double
        L.orc_landau_geff.restype = C.c_double
        for name in ("orc_mesh_info", "orc_mesh_tables", "orc_mass_dense", "orc_mass_solve",
                     "orc_eval_particle", "orc_l2_rhs", "orc_l2_project", "orc_entropy",
                     "orc_negative_part_l1", "orc_fs_minus_fp_l2", "orc_lb_rhs", "orc_lb_geff", "orc_gmres_dense", "orc_newton_midpoint",
                     "orc_step_newton",
                     "orc_compute_r", "orc_compute_G", "orc_loggrad", "orc_landau_geff",
                     "orc_picard_map", "orc_predictor", "orc_step_anderson", "orc_step_defect", "orc_time_step",
                     "orc_clusters", "orc_cluster_root", "orc_cluster_residual", "orc_grad_hess"):
            getattr(L, name).argtypes = None
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


def colmajor(v) -> np.ndarray:
    """(N,2) float64 array laid out like Julia's column-major Matrix."""
    return np.asfortranarray(np.asarray(v, dtype=np.float64))


def gauss_legendre01(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre01(C.c_int(n), _p(x), _p(w))
    return x, w


def extraction_1d(bp, p: int = 2, k: int = 1):
    bp = np.ascontiguousarray(bp, dtype=np.float64)
    ne = len(bp) - 1
    q = p + 1
    Cm = np.zeros(q * q * ne)
    bs = np.zeros(ne, dtype=np.int32)
    nd = C.c_int(0)
    lib().orc_extraction_1d(_p(bp), C.c_int(len(bp)), C.c_int(p), C.c_int(k), _p(Cm),
                            bs.ctypes.data_as(_ip), C.byref(nd))
    # C[i, j, e] column-major (Julia layout) -> numpy [e, j, i] row-major view; return [e, i, j]
    Cm = Cm.reshape(ne, q, q).transpose(0, 2, 1).copy()
    return Cm, bs, nd.value


@dataclass
class Mesh:
    """Oracle workspace: MantisWrappers.jl:90-158 restated."""

    bp1: np.ndarray
    bp2: np.ndarray
    p_deg: int = 2
    k_reg: int = 1
    n_quad: int = 6
    handle: int = field(default=0, repr=False)

    def __post_init__(self):
        self.bp1 = np.ascontiguousarray(self.bp1, dtype=np.float64)
        self.bp2 = np.ascontiguousarray(self.bp2, dtype=np.float64)
        h = lib().orc_mesh_build(_p(self.bp1), len(self.bp1), _p(self.bp2), len(self.bp2),
                                 self.p_deg, self.k_reg, self.n_quad)
        if not h:
            raise ValueError("oracle mesh build failed")
        self.handle = h
        info = np.zeros(8, dtype=np.int32)
        lib().orc_mesh_info(C.c_void_p(h), info.ctypes.data_as(_ip))
        (self.n_dofs, self.n_elem, self.nd1, self.nd2, self.ne1, self.ne2, _, _) = (int(x) for x in info)

    def __del__(self):
        if getattr(self, "handle", 0) and _LIB is not None:
            _LIB.orc_mesh_free(C.c_void_p(self.handle))
            self.handle = 0

    @property
    def h(self):
        return C.c_void_p(self.handle)

    @property
    def box(self):
        return np.array([self.bp1[0], self.bp1[-1], self.bp2[0], self.bp2[-1]])

    def tables(self):
        """(C1, C2, bs1, bs2) in the reference GPU layout: C[i,j,cell] column-major flat, 1-based starts."""
        q = self.p_deg + 1
        C1 = np.zeros(q * q * self.ne1)
        C2 = np.zeros(q * q * self.ne2)
        bs1 = np.zeros(self.ne1, dtype=np.int32)
        bs2 = np.zeros(self.ne2, dtype=np.int32)
        lib().orc_mesh_tables(self.h, _p(C1), _p(C2), bs1.ctypes.data_as(_ip), bs2.ctypes.data_as(_ip))
        return C1, C2, bs1, bs2

    def mass_dense(self):
        M = np.zeros((self.n_dofs, self.n_dofs))
        if lib().orc_mass_dense(self.h, _p(M)) != 0:
            raise ValueError("dense mass matrix only kept for n_dofs <= 1600")
        return M

    def mass_solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.n_dofs)
        lib().orc_mass_solve(self.h, _p(b), _p(x))
        return x

    def eval_particle(self, v1: float, v2: float):
        q2 = (self.p_deg + 1) ** 2
        vals, d1, d2, hh = np.zeros(q2), np.zeros(q2), np.zeros(q2), np.zeros(2)
        gids = np.zeros(q2, dtype=np.int32)
        n = lib().orc_eval_particle(self.h, C.c_double(v1), C.c_double(v2), _p(vals), _p(d1), _p(d2),
                                    gids.ctypes.data_as(_ip), _p(hh))
        if n == 0:
            return None
        return vals, d1, d2, gids, hh

    def l2_rhs(self, v, w):
        v = colmajor(v)
        w = np.ascontiguousarray(w, dtype=np.float64)
        rhs = np.zeros(self.n_dofs)
        lib().orc_l2_rhs(self.h, _p(v), _p(w), C.c_int64(v.shape[0]), _p(rhs))
        return rhs

    def l2_project(self, v, w):
        v = colmajor(v)
        w = np.ascontiguousarray(w, dtype=np.float64)
        fc = np.zeros(self.n_dofs)
        lib().orc_l2_project(self.h, _p(v), _p(w), C.c_int64(v.shape[0]), _p(fc))
        return fc

    def entropy(self, fc, use_logsq: bool = False) -> float:
        fc = np.ascontiguousarray(fc, dtype=np.float64)
        return float(lib().orc_entropy(self.h, _p(fc), C.c_int(int(use_logsq))))

    def negative_part_l1(self, fc) -> float:
        """compute_negative_part_l1 (functions.jl:300-313)."""
        fc = np.ascontiguousarray(fc, dtype=np.float64)
        return float(lib().orc_negative_part_l1(self.h, _p(fc)))

    def fs_minus_fp_l2(self, fc, v, w) -> float:
        """compute_fs_minus_fp_l2 (functions.jl:320-343)."""
        fc = np.ascontiguousarray(fc, dtype=np.float64)
        v = colmajor(v)
        w = np.ascontiguousarray(w, dtype=np.float64)
        return float(lib().orc_fs_minus_fp_l2(self.h, _p(fc), _p(v), _p(w), C.c_int64(v.shape[0])))

    def compute_r(self, fc, use_logsq: bool = False):
        fc = np.ascontiguousarray(fc, dtype=np.float64)
        r = np.zeros(self.n_dofs)
        lib().orc_compute_r(self.h, _p(fc), C.c_int(int(use_logsq)), _p(r))
        return r

    def compute_G(self, v, L):
        v = colmajor(v)
        L = np.ascontiguousarray(L, dtype=np.float64)
        G = np.zeros((v.shape[0], 2), order="F")
        lib().orc_compute_G(self.h, _p(v), C.c_int64(v.shape[0]), _p(L), _p(G))
        return G

    def loggrad(self, v, fc, fs_floor: float = 1e-30, g_max: float = 100.0):
        v = colmajor(v)
        fc = np.ascontiguousarray(fc, dtype=np.float64)
        g = np.zeros((v.shape[0], 2), order="F")
        lib().orc_loggrad(self.h, _p(v), C.c_int64(v.shape[0]), _p(fc), C.c_double(fs_floor),
                          C.c_double(g_max), _p(g))
        return g

    def collision(self, v, w, G, nthreads: int = 0):
        return collision(self.box, v, w, G, nthreads)

    def lb_geff(self, vin, v0, w, S0, use_gonzalez=True, use_logsq=False):
        """G_eff of the LB Picard map (solver.jl:185-210)."""
        vin, v0 = colmajor(vin), colmajor(v0)
        w = np.ascontiguousarray(w, dtype=np.float64)
        G = np.zeros((vin.shape[0], 2), order="F")
        lib().orc_lb_geff(self.h, _p(vin), _p(v0), _p(w), C.c_int64(vin.shape[0]), C.c_double(S0),
                          C.c_int(int(use_gonzalez)), C.c_int(int(use_logsq)), _p(G))
        return G

    def landau_geff(self, vin, v0, w, S0, use_gonzalez=True, use_logsq=False):
        vin, v0 = colmajor(vin), colmajor(v0)
        w = np.ascontiguousarray(w, dtype=np.float64)
        N = v0.shape[0]
        Geff = np.zeros((N, 2), order="F")
        vmid = np.zeros((N, 2), order="F")
        lam = lib().orc_landau_geff(self.h, _p(vin), _p(v0), _p(w), C.c_int64(N), C.c_double(S0),
                                    C.c_int(int(use_gonzalez)), C.c_int(int(use_logsq)), _p(Geff), _p(vmid))
        return Geff, vmid, float(lam)

    def picard_map(self, params: SolverParams, vin, v0, w, S0):
        vin, v0 = colmajor(vin), colmajor(v0)
        w = np.ascontiguousarray(w, dtype=np.float64)
        vout = np.zeros_like(v0, order="F")
        lib().orc_picard_map(self.h, C.byref(params), _p(vin), _p(v0), _p(w), C.c_int64(v0.shape[0]),
                             C.c_double(S0), _p(vout))
        return vout

    def predictor(self, params: SolverParams, v, w, fc):
        v = colmajor(v)
        w = np.ascontiguousarray(w, dtype=np.float64)
        fc = np.ascontiguousarray(fc, dtype=np.float64)
        v1 = np.zeros_like(v, order="F")
        dv = np.zeros_like(v, order="F")
        lib().orc_predictor(self.h, C.byref(params), _p(v), _p(w), C.c_int64(v.shape[0]), _p(fc), _p(v1), _p(dv))
        return v1, dv

    def step_anderson(self, params: SolverParams, v1, v0, w, S0):
        v1 = colmajor(v1).copy(order="F")
        v0 = colmajor(v0)
        w = np.ascontiguousarray(w, dtype=np.float64)
        st = StepStats()
        res = np.zeros(params.max_iter)
        lib().orc_step_anderson(self.h, C.byref(params), _p(v1), _p(v0), _p(w), C.c_int64(v0.shape[0]),
                                C.c_double(S0), C.byref(st), _p(res))
        return v1, st, res[: st.iters]

    def clusters(self, v, w, theta: float, kmax: int = 32):
        """Stiff clusters of markers v (extension): list of sorted index arrays."""
        v = colmajor(v)
        w = np.ascontiguousarray(w, dtype=np.float64)
        N = v.shape[0]
        off = np.zeros(N + 1, dtype=np.int32)
        idx = np.zeros(N, dtype=np.int32)
        n = lib().orc_clusters(self.h, _p(v), _p(w), C.c_int64(N), C.c_double(theta), C.c_int(kmax),
                               off.ctypes.data_as(_ip), C.c_int(N + 1), idx.ctypes.data_as(_ip), C.c_int(N))
        if n < 0:
            raise ValueError("cluster arrays too small")
        return [idx[off[c]:off[c + 1]].copy() for c in range(n)]

    def grad_hess(self, L, x: float, y: float):
        """(grad L, (H11, H12, H22)) at (x, y), or None outside the box."""
        L = np.ascontiguousarray(L, dtype=np.float64)
        g, H = np.zeros(2), np.zeros(3)
        ok = lib().orc_grad_hess(self.h, _p(L), C.c_double(x), C.c_double(y), _p(g), _p(H))
        return (g, H) if ok else None

    def cluster_residual(self, L, lam, dt, v0c, wc, ext, m, t=1.0, jacobian=False):
        """The cluster equations Phi(m; t) (interleaved (k, 2) arrays) and optionally their Jacobian."""
        v0c, ext, m = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (v0c, ext, m))
        wc = np.ascontiguousarray(wc, dtype=np.float64)
        k = wc.shape[0]
        F = np.zeros(2 * k)
        J = np.zeros((2 * k, 2 * k)) if jacobian else None
        ok = lib().orc_cluster_residual(self.h, _p(np.ascontiguousarray(L, dtype=np.float64)), C.c_double(lam),
                                        C.c_double(dt), C.c_int(k), _p(v0c), _p(wc), _p(ext), _p(m), C.c_double(t),
                                        _p(F), _p(J) if jacobian else None)
        if not ok:
            return None
        return (F, J) if jacobian else F

    def cluster_root(self, L, lam, dt, v0c, wc, ext, mprev=None):
        """One cluster's midpoint root (k, 2), or None; and the Newton iteration count."""
        v0c, ext = (np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (v0c, ext))
        wc = np.ascontiguousarray(wc, dtype=np.float64)
        k = wc.shape[0]
        m = np.zeros(2 * k)
        its = C.c_int(0)
        mp = None if mprev is None else np.ascontiguousarray(mprev, dtype=np.float64).ravel()
        ok = lib().orc_cluster_root(self.h, _p(np.ascontiguousarray(L, dtype=np.float64)), C.c_double(lam),
                                    C.c_double(dt), C.c_int(k), _p(v0c), _p(wc), _p(ext),
                                    _p(mp) if mp is not None else None, _p(m), C.byref(its))
        return (m.reshape(k, 2) if ok else None), its.value

    def step_defect(self, params: SolverParams, v1, v0, w, S0):
        """step_defect! (solver.jl:569-634); returns (v1, stats, outer residuals)."""
        v1 = colmajor(v1).copy(order="F")
        v0 = colmajor(v0)
        w = np.ascontiguousarray(w, dtype=np.float64)
        st = StepStats()
        res = np.zeros(params.max_iter)
        lib().orc_step_defect(self.h, C.byref(params), _p(v1), _p(v0), _p(w), C.c_int64(v0.shape[0]),
                              C.c_double(S0), C.byref(st), _p(res))
        return v1, st, res[: st.iters]

    def time_step(self, params: SolverParams, v, w, fc, S0):
        """One step of main.jl:226-307; returns (v_new, fc_new, stats)."""
        v = colmajor(v).copy(order="F")
        fc = np.ascontiguousarray(fc, dtype=np.float64).copy()
        w = np.ascontiguousarray(w, dtype=np.float64)
        st = StepStats()
        lib().orc_time_step(self.h, C.byref(params), _p(v), _p(w), C.c_int64(v.shape[0]), _p(fc),
                            C.c_double(S0), C.byref(st))
        return v, fc, st


def collision(box, v, w, G, nthreads: int = 0):
    """functions.jl:205-237 restated (OpenMP over targets gamma)."""
    v, G = colmajor(v), colmajor(G)
    w = np.ascontiguousarray(w, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    out = np.zeros((v.shape[0], 2), order="F")
    lib().orc_collision(_p(box), _p(v), _p(w), _p(G), C.c_int64(v.shape[0]), _p(out), C.c_int(nthreads))
    return out


def collision_rows(box, v, w, G, g0: int, g1: int, nthreads: int = 0):
    """The pair sum for targets gamma in [g0, g1) only (bounded CPU-baseline sample)."""
    v, G = colmajor(v), colmajor(G)
    w = np.ascontiguousarray(w, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    out = np.zeros((v.shape[0], 2), order="F")
    lib().orc_collision_rows(_p(box), _p(v), _p(w), _p(G), C.c_int64(v.shape[0]), C.c_int64(g0),
                             C.c_int64(g1), _p(out), C.c_int(nthreads))
    return out


def collision_metric(box, v, w, G, nthreads: int = 0):
    """functions.jl:249-290 restated; returns (dot_v, A) with A[:, (11, 12, 22)]."""
    v, G = colmajor(v), colmajor(G)
    w = np.ascontiguousarray(w, dtype=np.float64)
    box = np.ascontiguousarray(box, dtype=np.float64)
    N = v.shape[0]
    out = np.zeros((N, 2), order="F")
    A = np.zeros((N, 3), order="F")
    lib().orc_collision_metric(_p(box), _p(v), _p(w), _p(G), C.c_int64(N), _p(out), _p(A), C.c_int(nthreads))
    return out, A


def lb_rhs(v, w, g, nu: float = 1.0):
    """(dot_v, (A1, A2, B)): compute_moments, compute_drift_multipliers and
    compute_LB_velocity! (functions.jl:66-115)."""
    v, g = colmajor(v), colmajor(g)
    w = np.ascontiguousarray(w, dtype=np.float64)
    dv = np.zeros((v.shape[0], 2), order="F")
    AB = np.zeros(3)
    lib().orc_lb_rhs(_p(v), _p(w), _p(g), C.c_int64(v.shape[0]), C.c_double(nu), _p(dv), _p(AB))
    return dv, AB


def gmres_dense(A, b, rtol: float, atol: float = 0.0, itmax: int | None = None, mem: int | None = None):
    """gmres! (newton_krylov.jl:60-131) on y = A u from x0 = 0: (x, products, residual norm)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = b.shape[0]
    mem = n if mem is None else mem
    itmax = mem if itmax is None else itmax
    x = np.zeros(n)
    r = C.c_double(0.0)
    k = lib().orc_gmres_dense(C.c_int(n), _p(A), _p(b), C.c_double(rtol), C.c_double(atol), C.c_int(itmax),
                              C.c_int(mem), _p(x), C.byref(r))
    return x, int(k), float(r.value)


def newton_midpoint(B, x0, dt: float, x, tol: float, max_evals: int, fd_h: float, krylov_max: int = 20):
    """newton_krylov! (newton_krylov.jl:189-303) on F(x) = x - x0 - dt f((x0 + x)/2),
    f(y) = tanh(B y) - y^3 / 2.  Returns (x, F, status, n_evals, rnorm, n_newton, rnorm0);
    status 0 converged, 1 stalled, 2 budget."""
    B = np.ascontiguousarray(B, dtype=np.float64)
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    x = np.array(x, dtype=np.float64)
    n = x0.shape[0]
    F = np.zeros(n)
    out = np.zeros(4)
    st = lib().orc_newton_midpoint(C.c_int(n), _p(B), _p(x0), C.c_double(dt), C.c_double(tol), C.c_int(max_evals),
                                   C.c_double(fd_h), C.c_int(krylov_max), _p(x), _p(F), _p(out))
    return x, F, int(st), int(out[0]), float(out[1]), int(out[2]), float(out[3])


def momentum_energy(v, w):
    v = colmajor(v)
    w = np.ascontiguousarray(w, dtype=np.float64)
    o = np.zeros(3)
    lib().orc_momentum_energy(_p(v), _p(w), C.c_int64(v.shape[0]), _p(o))
    return (o[0], o[1]), o[2]
