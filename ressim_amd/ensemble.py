"""Batched FIM ensemble over the C ABI of include/ressim_ensemble.h (an extension: the
reference steps one model per simulator, rs/fim/timestep.rs:988-1890).

Members are ordinary ReservoirSimulator handles configured through the reference's
setters; BatchedEnsemble copies their models and stored states into one packed index
space on the device and advances all of them together, one captured HIP graph per
Newton iteration of every running member:

    sims = build_members([decks.two_phase_member(m) for m in range(256)])
    ens = BatchedEnsemble(sims)           # the tile choices are timed here
    ens.run(report_dt=2.0, n_reports=40)  # each member with its own substep control
    p, sw = ens.member_state(3)
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from .simulator import RSF_ERR_HIP, RSF_ERR_INVALID, RSF_ERR_NOT_CONVERGED, RSF_OK, ReservoirSimulator, lib

_D, _I, _P, _L = ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_long
_DP, _IP = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)


class EnsembleOptions(ctypes.Structure):
    _fields_ = [("max_newton_iterations", _I), ("min_newton_iteration_index", _I), ("newton_residual_tol", _D),
                ("newton_mb_tol", _D), ("linear_rel_tol", _D), ("krylov_dimension", _I), ("pressure_max_iterations", _I),
                ("pressure_rel_tol", _D), ("max_retries", _I),
                ("autotune", _I), ("threads_per_block", _I), ("cells_per_thread", _I * 3), ("use_graph", _I)]


class MemberReport(ctypes.Structure):
    _fields_ = [("status", _I), ("substeps", _I), ("retries", _I), ("newton_iterations", _I),
                ("linear_iterations", _I), ("n_cells", _I), ("time_days", _D)]


class Tuning(ctypes.Structure):
    _fields_ = [("threads_per_block", _I), ("cells_per_thread", _I * 3), ("class_bounds", _I * 2),
                ("class_members", _I * 3), ("probe_ms", ((_D * 3) * 3) * 3), ("ticks", _L), ("graph_nodes", _L),
                ("n_members", _I), ("n_cells", _L)]


_SIGS = {
    "rsf_ensemble_default_options": (None, [ctypes.POINTER(EnsembleOptions)]),
    "rsf_ensemble_create": (_I, [ctypes.POINTER(_P), _I, ctypes.POINTER(EnsembleOptions), ctypes.POINTER(_P)]),
    "rsf_ensemble_destroy": (None, [_P]),
    "rsf_ensemble_run": (_I, [_P, _D, _I]),
    "rsf_ensemble_replay": (_I, [_P, _DP, _IP]),
    "rsf_ensemble_member_state": (_I, [_P, _I, _DP, _DP, _I]),
    "rsf_ensemble_member_wells": (_I, [_P, _I, _DP, _I, _DP, _I]),
    "rsf_ensemble_member_stats": (_I, [_P, _I, ctypes.POINTER(MemberReport)]),
    "rsf_ensemble_tuning_info": (_I, [_P, ctypes.POINTER(Tuning)]),
    "rsf_ensemble_last_error": (ctypes.c_char_p, []),
}


def _lib():
    L = lib()
    if not getattr(L, "_ens_sigs", False):
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._ens_sigs = True
    return L


def default_options() -> EnsembleOptions:
    o = EnsembleOptions()
    _lib().rsf_ensemble_default_options(ctypes.byref(o))
    return o


def build_members(member_decks):
    """Simulator handles for the decks, configured by their setter calls (not stepped)."""
    return [d.build(ReservoirSimulator) for d in member_decks]


class BatchedEnsemble:
    STATUS = ("running", "finished", "stopped")

    def __init__(self, sims, **options):
        self._l = _lib()
        o = default_options()
        for k, v in options.items():
            if k == "cells_per_thread":
                for i, c in enumerate(v):
                    o.cells_per_thread[i] = int(c)
            elif hasattr(o, k):
                setattr(o, k, type(getattr(o, k))(v))
            else:
                raise ValueError(f"unknown ensemble option {k}")
        self.options = o
        self.n_members = len(sims)
        self._cells = [s.n for s in sims]
        self._wells = [s._l.rsf_num_wells(s._h) for s in sims]
        self._perfs = [s._l.rsf_num_perforations(s._h) for s in sims]
        arr = (_P * len(sims))(*[s._h for s in sims])
        h = _P()
        self._check(self._l.rsf_ensemble_create(arr, len(sims), ctypes.byref(o), ctypes.byref(h)), "create")
        self._h = h

    def _check(self, rc, what):
        if rc == RSF_OK:
            return
        msg = self._l.rsf_ensemble_last_error().decode()
        if rc == RSF_ERR_INVALID:
            raise ValueError(f"ensemble {what}: {msg}")
        if rc == RSF_ERR_HIP:
            raise RuntimeError(f"ensemble {what}: HIP error {msg}")
        raise RuntimeError(f"ensemble {what}: rc {rc} {msg}")

    def close(self):
        if getattr(self, "_h", None):
            self._l.rsf_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, report_dt, n_reports):
        """Every member through n_reports report steps; returns (rc, wall seconds)."""
        t0 = time.perf_counter()
        rc = self._l.rsf_ensemble_run(self._h, float(report_dt), int(n_reports))
        el = time.perf_counter() - t0
        if rc not in (RSF_OK, RSF_ERR_NOT_CONVERGED):
            self._check(rc, "run")
        return rc, el

    def replay(self, dts_per_member):
        """Member m takes exactly the substeps dts_per_member[m]; returns rc."""
        offs = np.zeros(self.n_members + 1, dtype=np.int32)
        for m, d in enumerate(dts_per_member):
            offs[m + 1] = offs[m] + len(d)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(d, dtype=np.float64) for d in dts_per_member])
                                    if offs[-1] else np.zeros(1))
        rc = self._l.rsf_ensemble_replay(self._h, flat.ctypes.data_as(_DP), offs.ctypes.data_as(_IP))
        if rc not in (RSF_OK, RSF_ERR_NOT_CONVERGED):
            self._check(rc, "replay")
        return rc

    def member_state(self, m):
        n = self._cells[m]
        p, sw = np.empty(n), np.empty(n)
        self._check(self._l.rsf_ensemble_member_state(self._h, int(m), p.ctypes.data_as(_DP), sw.ctypes.data_as(_DP),
                                                      n), "member_state")
        return p, sw

    def member_wells(self, m):
        """(BHP per well, connection rate per completion), the member's own orders."""
        nw, npf = self._wells[m], self._perfs[m]
        b, q = np.empty(max(nw, 1)), np.empty(max(npf, 1))
        self._check(self._l.rsf_ensemble_member_wells(self._h, int(m), b.ctypes.data_as(_DP), nw,
                                                      q.ctypes.data_as(_DP), npf), "member_wells")
        return b[:nw], q[:npf]

    def member_stats(self, m):
        r = MemberReport()
        self._check(self._l.rsf_ensemble_member_stats(self._h, int(m), ctypes.byref(r)), "member_stats")
        return dict(status=self.STATUS[r.status], substeps=r.substeps, retries=r.retries,
                    newton_iterations=r.newton_iterations, linear_iterations=r.linear_iterations,
                    n_cells=r.n_cells, time_days=r.time_days)

    def tuning(self):
        t = Tuning()
        self._check(self._l.rsf_ensemble_tuning_info(self._h, ctypes.byref(t)), "tuning_info")
        probe = [[[round(t.probe_ms[a][b][c], 4) for c in range(3)] for b in range(3)] for a in range(3)]
        return dict(threads_per_block=t.threads_per_block, cells_per_thread=list(t.cells_per_thread),
                    class_bounds=list(t.class_bounds), class_members=list(t.class_members),
                    probe_ms_by_threads_class_cpt=probe, probe_threads=[128, 256, 512], probe_cpt=[1, 2, 4],
                    ticks=t.ticks, graph_nodes=t.graph_nodes, n_members=t.n_members, n_cells=t.n_cells)
