"""Workspace: the mesh tables the reference's GPU path reads from its Mantis
workspace (MantisWrappers.jl:39-80, built in :90-158), built here without
Mantis and handed to the device library once.

Only the space the device kernels support is built: tensor-product quadratic
B-splines (P_DEG = 2) with C^1 continuity (K_REG = 1) on an open knot vector
over the breakpoints ``bp1`` x ``bp2`` (the configuration of every Landau
preset, e.g. parameters_sq_d04.jl:8).  The reference raises for P_DEG != 2 on
its GPU path (projection_gpu.jl:189); so does this module.
"""
from __future__ import annotations

import numpy as np

from . import _lib


def extraction_p2c1(bp) -> tuple[np.ndarray, np.ndarray, int]:
    """Bezier extraction of the C^1 quadratic B-splines over breakpoints ``bp``.

    Returns ``(C, bs, nd)``: ``C[e, i, j]`` maps Bernstein polynomial ``i`` to
    local basis function ``j`` of cell ``e`` (phi_j = sum_i C[e,i,j] B_i, the
    convention of MantisWrappers.jl:237), ``bs[e]`` the 1-based index of the
    first basis function on cell ``e``, ``nd`` the number of basis functions.

    Derived by blossoming: on the span [t_i, t_{i+1}] the Bernstein
    coefficients of sum_k c_k N_k are (s(t_i,t_i), s(t_i,t_{i+1}),
    s(t_{i+1},t_{i+1})), and the blossom s is affine in each argument between
    the de Boor points c_k = s(t_{k+1}, t_{k+2}).
    """
    bp = np.asarray(bp, dtype=np.float64)
    ne = len(bp) - 1
    if ne < 1 or not np.all(np.diff(bp) > 0):
        raise ValueError("breakpoints must be strictly increasing with at least one cell")
    C = np.zeros((ne, 3, 3))
    for e in range(ne):
        left = bp[e - 1] if e > 0 else bp[e]       # t_{i-1} (repeated end knot at e = 0)
        right = bp[e + 2] if e + 2 <= ne else bp[e + 1]  # t_{i+2}
        h = bp[e + 1] - bp[e]
        alpha = h / (bp[e + 1] - left)
        beta = h / (right - bp[e])
        C[e, 0, 0] = alpha
        C[e, 0, 1] = 1.0 - alpha
        C[e, 1, 1] = 1.0
        C[e, 2, 1] = 1.0 - beta
        C[e, 2, 2] = beta
    bs = np.arange(1, ne + 1, dtype=np.int32)
    return C, bs, ne + 2


def flat_tables(C: np.ndarray) -> np.ndarray:
    """C[e, i, j] -> the flat column-major layout C[i + 3 j + 9 e] (projection_gpu.jl:191-197)."""
    return np.ascontiguousarray(C.transpose(0, 2, 1).reshape(-1))


class Workspace:
    """Mesh tables plus the device context they live on.

    Field names follow MantisWrappers.Workspace: ``bp1``, ``bp2``, ``n_dofs``,
    ``n_elements``, ``ext1d_1``/``ext1d_2`` (C[e, i, j]), ``basis_start_1d_1``/
    ``basis_start_1d_2`` (1-based) and ``n_dofs_1d_1``/``n_dofs_1d_2``.
    """

    def __init__(self, bp1, bp2, p_deg: int = 2, k_reg: int = 1, n_quad: int = 6, device: int = 0,
                 ctx: _lib.Context | None = None):
        if p_deg != 2:
            raise _lib.CollopError(_lib.ERR_UNSUPPORTED, f"the device kernels are specialised to P_DEG=2 (got {p_deg})")
        if k_reg != 1:
            raise _lib.CollopError(_lib.ERR_UNSUPPORTED, f"only K_REG=1 (C^1 quadratics) is built (got {k_reg})")
        self.P_DEG, self.K_REG, self.N_QUAD = p_deg, k_reg, n_quad
        self.bp1 = np.ascontiguousarray(bp1, dtype=np.float64)
        self.bp2 = np.ascontiguousarray(bp2, dtype=np.float64)
        self.ext1d_1, self.basis_start_1d_1, self.n_dofs_1d_1 = extraction_p2c1(self.bp1)
        self.ext1d_2, self.basis_start_1d_2, self.n_dofs_1d_2 = extraction_p2c1(self.bp2)
        self.n_dofs = self.n_dofs_1d_1 * self.n_dofs_1d_2
        self.n_elements = (len(self.bp1) - 1) * (len(self.bp2) - 1)
        self.ctx = ctx if ctx is not None else _lib.Context(device)
        self._C1 = flat_tables(self.ext1d_1)
        self._C2 = flat_tables(self.ext1d_2)
        self._token = object()
        self.activate()

    def activate(self):
        """Make this mesh the context's current one.  A context holds one mesh at
        a time (collop_set_mesh replaces it and drops any device-resident state),
        so workspaces sharing a context re-upload their tables when they take
        turns."""
        if getattr(self.ctx, "mesh_owner", None) is self._token:
            return
        L = self.ctx.lib
        self.ctx.check(L.collop_set_mesh(
            self.ctx.h, _lib.dptr(self.bp1), len(self.bp1), _lib.dptr(self.bp2), len(self.bp2),
            _lib.dptr(self._C1), _lib.dptr(self._C2), _lib.iptr(self.basis_start_1d_1),
            _lib.iptr(self.basis_start_1d_2), self.n_dofs_1d_1, self.n_dofs_1d_2, self.P_DEG, self.N_QUAD))
        self.ctx.mesh_owner = self._token
        self.ctx.state_owner = None

    @property
    def box(self):
        return (self.bp1[0], self.bp1[-1], self.bp2[0], self.bp2[-1])


def build_workspace(p, device: int = 0, ctx: _lib.Context | None = None) -> Workspace:
    """build_workspace(p) of MantisWrappers.jl:90 for a :class:`SimParameters`."""
    return Workspace(p.bp1, p.bp2, p.P_DEG, p.K_REG, p.N_QUAD, device=device, ctx=ctx)
