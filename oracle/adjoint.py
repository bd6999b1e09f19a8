"""Discrete adjoint of the FIM step on the oracle, FP64 on the CPU. TEST INFRASTRUCTURE ONLY.

The objective is the surface oil of the producers over the accepted substeps,
J = sum_n dt_n sum_e (fo/Bo)(x_n) q_e,n (orc_producer_oil); the parameters are a per-cell
log-multiplier of kx, ky and kz together and the per-cell porosity. Per substep n, with R_n
the residual of assemble_fim_system_ad (rs/fim/assembly_ad.rs:782-983) at its converged
iterate x_n against the state before it:

    J_n^T lam_n = -dj_n/dx_n + (d inv(x_n)/dx_n)^T lam_{n+1}      (R_{n+1} holds -inv(x_n))
    dJ/dtheta   = sum_n lam_n^T dR_n/dtheta

J_n is the oracle's full AD Jacobian (cells and the well tail) and is solved directly:
scipy's SuperLU on J_n^T with the unknowns in a geometric nested-dissection order of the
grid (recursive bisection of the longest axis, separator planes last, the well tail at the
end), which roughly halves the fill of SuperLU's own orderings on these grids. The forward
run replays a given substep schedule, so that a finite difference of J with the same
schedule checks the gradient.
"""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import _DP, _IP, OracleSimulator, lib


def _sigs():
    L = lib()
    P = ctypes.c_void_p
    I, D = ctypes.c_int, ctypes.c_double
    L.orc_last_iterate.restype = I
    L.orc_last_iterate.argtypes = [P, _DP, _IP, _DP, _DP]
    L.orc_assemble_prev.restype = I
    L.orc_assemble_prev.argtypes = [P, D, _DP, _IP, _DP, _IP, _DP, _DP, ctypes.POINTER(I), ctypes.POINTER(I)]
    L.orc_adjoint_state_terms.restype = I
    L.orc_adjoint_state_terms.argtypes = [P, D, _DP, _IP, _DP, _DP, _DP, _DP, _DP]
    L.orc_adjoint_param_terms.restype = I
    L.orc_adjoint_param_terms.argtypes = [P, D, _DP, _IP, _DP, _IP, _DP, _DP, _DP, _DP, _DP]
    return L


def _p(a, t=_DP):
    return a.ctypes.data_as(t)


def nested_dissection(nx, ny, nz, leaf=64):
    """Cell order: both halves of the longest axis recursively, then the separating plane."""
    out = []

    def rec(b):
        if b.size <= leaf:
            out.append(b.ravel())
            return
        ax = int(np.argmax(b.shape))
        m = b.shape[ax] // 2
        rec(np.take(b, range(0, m), axis=ax))
        rec(np.take(b, range(m + 1, b.shape[ax]), axis=ax))
        out.append(np.take(b, [m], axis=ax).ravel())

    rec(np.arange(nx * ny * nz).reshape(nz, ny, nx))
    return np.concatenate(out)


def solve_transposed(Jn, b, cells):
    """x with Jn^T x = b; unknown 3 i + c of cell i, the well tail after the cells."""
    nu = Jn.shape[0]
    perm = np.concatenate([(3 * cells[:, None] + np.arange(3)).ravel(), np.arange(3 * len(cells), nu)])
    A = sp.csc_matrix(Jn.T)[perm][:, perm]
    lu = spla.splu(sp.csc_matrix(A), permc_spec="MMD_AT_PLUS_A")
    x = np.empty(nu)
    x[perm] = lu.solve(b[perm])
    return x


class OracleAdjoint:
    """forward(dts) records the substeps of a replayed schedule; gradient() returns
    (dJ/d ln k, dJ/d phi) per cell."""

    def __init__(self, deck, newton_options=None):
        self.deck = deck
        self.newton_options = dict(newton_options or {})
        self.sim = None
        self.records = []
        self.J = 0.0

    def forward(self, dts):
        L = _sigs()
        o = self.deck.build(OracleSimulator)
        if self.newton_options:
            o.set_newton_options(**self.newton_options)
        n, W, P = o.n, L.orc_num_wells(o._h), L.orc_num_perfs(o._h)
        self.records = []
        for dt in dts:
            prev, prev_reg = o.initial_state()
            rc, _, _ = o.run_substep(dt)
            if rc != 0:
                raise RuntimeError(f"oracle substep failed: {o.last_warning()}")
            x = np.empty(3 * n)
            reg = np.empty(n, dtype=np.int32)
            bhp, q = np.empty(max(W, 1)), np.empty(max(P, 1))
            L.orc_last_iterate(o._h, _p(x), _p(reg, _IP), _p(bhp), _p(q))
            self.records.append((float(dt), np.ascontiguousarray(prev.ravel()), np.ascontiguousarray(prev_reg, np.int32),
                                 x, reg, bhp[:W].copy(), q[:P].copy()))
        self.sim = o
        self.J = o.producer_oil()
        return self.J

    def gradient(self):
        L = _sigs()
        o = self.sim
        n = o.n
        g_lnk, g_phi = np.zeros(n), np.zeros(n)
        cells = nested_dissection(self.deck.nx, self.deck.ny, self.deck.nz)
        lam_next = None
        for dt, prev, prev_reg, x, reg, bhp, q in reversed(self.records):
            nu, nnz = ctypes.c_int(0), ctypes.c_int(0)
            L.orc_assemble_prev(o._h, dt, _p(prev), _p(prev_reg, _IP), _p(x), _p(reg, _IP), _p(bhp), _p(q),
                                ctypes.byref(nu), ctypes.byref(nnz))
            R = np.empty(nu.value)
            ptr = np.empty(nu.value + 1, dtype=np.int32)
            col = np.empty(nnz.value, dtype=np.int32)
            val = np.empty(nnz.value)
            L.orc_get_assembly(_p(R), _p(ptr, _IP), _p(col, _IP), _p(val))
            Jn = sp.csr_matrix((val, col, ptr), shape=(nu.value, nu.value))
            j = ctypes.c_double(0.0)
            djdx = np.empty(nu.value)
            dinv = np.empty(9 * n)
            L.orc_adjoint_state_terms(o._h, dt, _p(x), _p(reg, _IP), _p(bhp), _p(q), ctypes.byref(j), _p(djdx),
                                      _p(dinv))
            b = -djdx
            if lam_next is not None:
                # (d inv(x_n)/dx_n)^T lam_{n+1}, cell by cell (rows: components, columns: unknowns)
                D = dinv.reshape(n, 3, 3)
                b[:3 * n] += np.einsum("irv,ir->iv", D, lam_next[:3 * n].reshape(n, 3)).ravel()
            lam = solve_transposed(Jn, b, cells)
            L.orc_adjoint_param_terms(o._h, dt, _p(prev), _p(prev_reg, _IP), _p(x), _p(reg, _IP), _p(bhp), _p(q),
                                      _p(np.ascontiguousarray(lam)), _p(g_lnk), _p(g_phi))
            lam_next = lam
        return g_lnk, g_phi
