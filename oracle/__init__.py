"""Oracle: ctypes binding of the CPU FP64 restatement of the reference FIM step.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker. The product library (``ressim_amd``) never loads it.

``OracleSimulator`` exposes the same reference-named setters as
``ressim_amd.ReservoirSimulator`` (rs/frontend.rs) so one deck description
drives both.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_fim.so")
_lib = None

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_LP = ctypes.POINTER(ctypes.c_long)
_S = ctypes.c_char_p

class StepStats(ctypes.Structure):  # FimStepStats scalars, rs/reporting.rs:191-243; -1 / NaN: None
    _fields_ = [("time_days", _D), ("target_dt_days", _D), ("advanced_dt_days", _D), ("accepted_substeps", _I),
                ("linear_bad_retries", _I), ("nonlinear_bad_retries", _I), ("mixed_retries", _I),
                ("min_accepted_dt_days", _D), ("max_accepted_dt_days", _D), ("last_accepted_dt_days", _D),
                ("last_growth_factor", _D), ("growth_limiter", _I), ("last_retry_class", _I),
                ("last_retry_dominant_family", _I), ("n_rungs", _I), ("solver_ms", _D), ("accepted_solver_ms", _D),
                ("retry_solver_ms", _D), ("assembly_ms", _D), ("linear_solve_ms", _D),
                ("last_retry_dominant_row", ctypes.c_long)]


class StepRung(ctypes.Structure):  # FimAcceptedRungStats / FimRetryRungStats, rs/reporting.rs:141-188
    _fields_ = [("substep", _I), ("retried", _I), ("newton_iterations", _I), ("linear_iterations", _I),
                ("linear_solve_count", _I), ("retry_class", _I), ("dominant_family", _I), ("last_linear_converged", _I),
                ("linear_iteration_count", ctypes.c_long), ("dt_days", _D), ("linear_solve_ms", _D),
                ("dominant_row", ctypes.c_long)]


# labels of the codes (rs/fim/timestep.rs:502-528, rs/fim/newton.rs:174-182,
# rs/fim/newton/convergence.rs:294-304)
GROWTH_LIMITERS = ("opm-max-growth", "opm-pid", "opm-iter", "opm-restart-growth")
RETRY_CLASSES = ("linear-bad", "nonlinear-bad", "mixed")
RESIDUAL_FAMILIES = ("water", "oil", "gas", "well", "perf")


def decode_step_stats(s, rungs):
    """One FimStepStats record as the reference serialises it: None for absent values,
    labels for the codes, and the rungs split into accepted_rungs and retry_rungs."""
    def opt(x):
        return None if x != x else x

    def label(table, code):
        return None if code < 0 else table[code]

    out = {f: getattr(s, f) for f, _ in StepStats._fields_ if f != "n_rungs"}
    for f in ("min_accepted_dt_days", "max_accepted_dt_days", "last_accepted_dt_days", "last_growth_factor"):
        out[f] = opt(out[f])
    out["growth_limiter"] = label(GROWTH_LIMITERS, s.growth_limiter)
    out["last_retry_class"] = label(RETRY_CLASSES, s.last_retry_class)
    out["last_retry_dominant_family"] = label(RESIDUAL_FAMILIES, s.last_retry_dominant_family)
    out["last_retry_dominant_row"] = None if s.last_retry_dominant_row < 0 else s.last_retry_dominant_row
    acc, ret = [], []
    for r in rungs:
        d = dict(substep=r.substep, dt_days=r.dt_days, newton_iterations=r.newton_iterations,
                 linear_iterations=None if r.linear_iterations < 0 else r.linear_iterations,
                 linear_solve_count=r.linear_solve_count, linear_iteration_count=r.linear_iteration_count,
                 linear_solve_ms=r.linear_solve_ms,
                 last_linear_converged=None if r.last_linear_converged < 0 else bool(r.last_linear_converged))
        if r.retried:
            d["retry_class"] = label(RETRY_CLASSES, r.retry_class)
            d["dominant_family"] = label(RESIDUAL_FAMILIES, r.dominant_family)
            d["dominant_row"] = r.dominant_row
            ret.append(d)
        else:
            acc.append(d)
    out["accepted_rungs"] = acc
    out["retry_rungs"] = ret
    return out


def _csr_args(A):
    import scipy.sparse as sp

    A = sp.csr_matrix(A)
    A.sort_indices()
    ptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
    col = np.ascontiguousarray(A.indices, dtype=np.int32)
    val = np.ascontiguousarray(A.data, dtype=np.float64)
    return A.shape[0], (ptr, col, val)


def scalar_ilu0(A, rhs):
    """factorize_pressure_ilu0 / apply (rs/fim/linear/gmres_block_jacobi.rs:746-808, :634-658)
    on a scalar matrix: (the product L U as a dense array, the apply to rhs)."""
    n, (ptr, col, val) = _csr_args(A)
    lu = np.empty((n, n))
    x = np.empty(n)
    r = np.ascontiguousarray(rhs, dtype=np.float64)
    lib().orc_scalar_ilu0(n, ptr.ctypes.data_as(_IP), col.ctypes.data_as(_IP), val.ctypes.data_as(_DP),
                          r.ctypes.data_as(_DP), lu.ctypes.data_as(_DP), x.ctypes.data_as(_DP))
    return lu, x


def block_ilu0_apply(A, rhs):
    """factorize_block_ilu0 / apply (:239-354, :190-237) on a cell system of 3x3 blocks."""
    n, (ptr, col, val) = _csr_args(A)
    assert n % 3 == 0
    x = np.empty(n)
    r = np.ascontiguousarray(rhs, dtype=np.float64)
    lib().orc_block_ilu0_apply(n // 3, ptr.ctypes.data_as(_IP), col.ctypes.data_as(_IP), val.ctypes.data_as(_DP),
                               r.ctypes.data_as(_DP), x.ctypes.data_as(_DP))
    return x


def cpr_probe(A, rp, mode=0, max_iterations=50, tol=0.0):
    """The CPR of a cell system of 3x3 blocks (:1022-1363): quasi-IMPES weights, the coarse
    pressure matrix, and mode 0 the coarse solve of rp as the preconditioner does it (dense
    inverse up to 299 rows, else BiCGStab), mode 1 BiCGStab with the given budget and
    absolute tolerance, mode 2 the full two-stage apply of the cell vector rp."""
    n, (ptr, col, val) = _csr_args(A)
    nb = n // 3
    w = np.empty((nb, 3))
    P = np.empty((nb, nb))
    r = np.ascontiguousarray(rp, dtype=np.float64)
    out = np.empty(n if mode == 2 else nb)
    info = np.zeros(2, dtype=np.int32)
    lib().orc_cpr_probe(nb, ptr.ctypes.data_as(_IP), col.ctypes.data_as(_IP), val.ctypes.data_as(_DP),
                        r.ctypes.data_as(_DP), int(mode), int(max_iterations), float(tol), w.ctypes.data_as(_DP),
                        P.ctypes.data_as(_DP), out.ctypes.data_as(_DP), info.ctypes.data_as(_IP))
    return dict(weights=w, pressure_matrix=P, x=out, dense=bool(info[0]), bicgstab_iterations=int(info[1]))


_SIGS = {
    "orc_set_threads": (_I, [_I]),
    "orc_set_scal_tables": (_I, [_P, _DP, _I, _DP, _I]),
    "orc_scalar_ilu0": (_I, [_I, _IP, _IP, _DP, _DP, _DP, _DP]),
    "orc_block_ilu0_apply": (_I, [_I, _IP, _IP, _DP, _DP, _DP]),
    "orc_cpr_probe": (_I, [_I, _IP, _IP, _DP, _DP, _I, _I, _D, _DP, _DP, _DP, _IP]),
    "orc_create": (_P, [_I, _I, _I, _D]),
    "orc_destroy": (None, [_P]),
    "orc_set_cell_dimensions_per_layer": (_I, [_P, _D, _D, _DP]),
    "orc_set_fields": (_I, [_P, _DP, _DP, _DP, _DP]),
    "orc_set_pc_scale": (_I, [_P, _DP]),
    "orc_set_initial_state": (_I, [_P, _DP, _DP, _DP, _DP]),
    "orc_set_fluid": (_I, [_P, _D, _D, _D, _D, _D, _D]),
    "orc_set_rock_properties": (_I, [_P, _D, _D, _D, _D]),
    "orc_set_gas_fluid_properties": (_I, [_P, _D, _D, _D]),
    "orc_set_rel_perm_props": (_I, [_P, _D, _D, _D, _D, _D, _D]),
    "orc_set_three_phase_rel_perm_props": (_I, [_P] + [_D] * 11),
    "orc_set_capillary_params": (_I, [_P, _D, _D]),
    "orc_set_gas_oil_capillary_params": (_I, [_P, _D, _D]),
    "orc_set_flags": (_I, [_P, _I, _I, _I, _I]),
    "orc_set_gas_redissolution": (_I, [_P, _I]),
    "orc_set_reference_pressure": (_I, [_P, _D]),
    "orc_set_pvt_table": (_I, [_P, _DP, _I]),
    "orc_set_well_controls": (_I, [_P, _I, _I, _I, _D, _D, _D, _D, _D, _D]),
    "orc_add_well": (_I, [_P, _I, _I, _I, _D, _D, _D, _I, _S]),
    "orc_set_well_schedule": (_I, [_P, _S, _I, _D, _D, _D, _I]),
    "orc_set_well_datum": (_I, [_P, _S, _D, _D]),
    "orc_set_well_fluid": (_I, [_P, _S, _I]),
    "orc_set_newton_options": (_I, [_P, _I, _I, _D, _D, _I, _I]),
    "orc_step": (_I, [_P, _D]),
    "orc_run_substep": (_I, [_P, _D, _IP, _LP]),
    "orc_get_field": (_I, [_P, _I, _DP]),
    "orc_num_wells": (_I, [_P]),
    "orc_num_perfs": (_I, [_P]),
    "orc_get_well_state": (_I, [_P, _DP, _DP]),
    "orc_get_time": (_D, [_P]),
    "orc_get_stats": (_I, [_P, _LP]),
    "orc_get_accepted_dts": (_I, [_P, _DP, _I]),
    "orc_step_stats_count": (_I, [_P]),
    "orc_classify_retry": (_I, [_I, _I, _I, _I, _D, _D]),
    "orc_dominant_family": (_I, [_DP, _I]),
    "orc_family_peaks": (_I, [_I, _I, _I, _DP, _DP, _DP, _LP]),
    "orc_get_step_stats": (_I, [_P, _I, ctypes.POINTER(StepStats)]),
    "orc_get_step_rungs": (_I, [_P, _I, ctypes.POINTER(StepRung), _I]),
    "orc_assemble": (_I, [_P, _D, _DP, _IP, _DP, _DP, _IP, _IP]),
    "orc_get_assembly": (_I, [_DP, _IP, _IP, _DP]),
    "orc_residual_check": (_I, [_P, _D, _DP, _IP, _DP, _DP, _DP]),
    "orc_scaled_residual": (_I, [_P, _D, _DP, _IP, _DP, _DP, _DP]),
    "orc_producer_oil": (_D, [_P]),
    "orc_rate_history_len": (_I, [_P]),
    "orc_get_rate_history": (_I, [_P, _DP, _I]),
    "orc_inventory_totals": (_I, [_P, _DP]),
    "orc_get_initial_state": (_I, [_P, _DP, _IP]),
    "orc_probe": (_I, [_P, _I, _D, _D, _DP]),
    "orc_last_warning": (_S, [_P]),
    "orc_chop": (_I, [_I, _DP, _IP, _I, _DP, _DP, _D, _DP]),
    "orc_cnv_mb": (_I, [_I, _DP, _DP, _DP, _I, _DP]),
    "orc_ad_selftest": (_I, [_DP]),
    "orc_newton_direction": (_I, [_P, _D, _DP, _IP, _DP, _DP, _D, _I, _DP, _IP]),
    "orc_set_nested_well_solve": (_I, [_P, _I]),
    "orc_set_sweep_config": (_I, [_P, _S, _D, _D, _D]),
    "orc_get_initial_well_state": (_I, [_P, _DP, _DP]),
    "orc_well_local_system": (_I, [_P, _DP, _IP, _DP, _DP, _I, _DP, _DP]),
    "orc_solve_wells_locally": (_I, [_P, _DP, _IP, _DP, _DP, _I, _IP, _IP, _DP]),
    "orc_chop_bhp_update": (_D, [_D, _D, _D]),
    "orc_relax_wells": (_I, [_P, _DP, _IP, _DP, _DP]),
    "orc_pid_dt": (_D, [_D, _DP]),
    "orc_iteration_count_dt": (_D, [_D, _I]),
    "orc_growth_factor": (_D, [_D, _I, _DP, _I, _IP]),
    "orc_detect_oscillation": (_I, [_DP, _DP, _DP]),
    "orc_next_relaxation_factor": (_D, [_D, _I]),
    "orc_relative_change": (_D, [_I, _DP, _IP, _DP, _IP]),
}


def build(force: bool = False) -> str:
    """Compile liboracle_fim.so with the committed Makefile (g++, -O3)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def scal_rows(rows, keys):
    """SWOF/SGOF rows as an (n, 3) array of (saturation, kr, kro) from the reference's row
    objects (dicts with those keys) or from 3- or 4-sequences."""
    out = []
    for r in rows:
        if isinstance(r, dict):
            out.append([float(r[k]) for k in keys])
        else:
            out.append([float(v) for v in list(r)[:3]])
    return np.ascontiguousarray(np.array(out, dtype=np.float64).reshape(-1, 3))


def set_threads(n):
    """Threads of the oracle's parallel loops (n <= 0: the OpenMP default); returns the count
    in effect. Results are the same for every count."""
    return lib().orc_set_threads(int(n))


def classify_retry(have_linear, converged, direct=False, coarse_applications=0, average_reduction=1.0,
                   last_reduction=1.0):
    """rs/fim/newton.rs:439-481 classify_retry_failure_with_site: the class label."""
    return RETRY_CLASSES[lib().orc_classify_retry(int(have_linear), int(converged), int(direct),
                                                  int(coarse_applications), float(average_reduction),
                                                  float(last_reduction))]


def dominant_family(family_maxima, with_wells):
    """rs/fim/newton/convergence.rs:793-806: label of the dominant residual family."""
    f = np.ascontiguousarray(family_maxima, dtype=np.float64)
    assert f.shape == (5,)
    return RESIDUAL_FAMILIES[lib().orc_dominant_family(f.ctypes.data_as(_DP), int(with_wells))]


def family_peaks(n_cells, n_wells, n_perf, residual, scale):
    """rs/fim/newton/convergence.rs:734-790 residual_family_diagnostics over a residual in
    the reference's unknown numbering with one scale per row: (peak values, peak rows,
    global family label); rows of a family with no rows are -1."""
    r = np.ascontiguousarray(residual, dtype=np.float64)
    s = np.ascontiguousarray(scale, dtype=np.float64)
    assert r.size == s.size == 3 * n_cells + n_wells + n_perf
    vals = np.empty(5)
    rows = np.empty(5, dtype=np.int64)
    g = lib().orc_family_peaks(int(n_cells), int(n_wells), int(n_perf), r.ctypes.data_as(_DP),
                               s.ctypes.data_as(_DP), vals.ctypes.data_as(_DP), rows.ctypes.data_as(_LP))
    if g < 0:
        raise ValueError("orc_family_peaks: invalid arguments")
    return vals, rows, RESIDUAL_FAMILIES[g]


def _dp(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_DP)


def chop(cells, regimes, bhp, update, relax=1.0):
    """Per-cell OPM chop of a raw Newton update (cells: n x [p, sw, hc])."""
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 3)
    rg = np.ascontiguousarray(regimes, dtype=np.int32)
    b = np.ascontiguousarray(bhp, dtype=np.float64).reshape(-1)
    up = np.ascontiguousarray(update, dtype=np.float64)
    out = np.empty_like(up)
    lib().orc_chop(len(cells), cells.ctypes.data_as(_DP), rg.ctypes.data_as(_IP), len(b),
                   b.ctypes.data_as(_DP) if len(b) else None, up.ctypes.data_as(_DP), relax, out.ctypes.data_as(_DP))
    return out


def cnv_mb(residual, pv, fvf, relax_final=False):
    r = np.ascontiguousarray(residual, dtype=np.float64)
    p = np.ascontiguousarray(pv, dtype=np.float64)
    f = np.ascontiguousarray(fvf, dtype=np.float64).reshape(-1)
    out = np.empty(9)
    lib().orc_cnv_mb(len(p), r.ctypes.data_as(_DP), p.ctypes.data_as(_DP), f.ctypes.data_as(_DP), int(relax_final),
                     out.ctypes.data_as(_DP))
    return dict(cnv=out[0:3], mb=out[3:6], violating_pv_fraction=out[6], pv_rule_relaxes=bool(out[7]),
                would_accept=bool(out[8]))


def chop_bhp_update(bhp, raw_delta, dbhp_max_rel=1.0):
    """rs/fim/wells_inner.rs:432-436: relative BHP chop with the 1 bar - 1 Pa floor."""
    return lib().orc_chop_bhp_update(bhp, raw_delta, dbhp_max_rel)


def pid_dt(dt, errors):
    """rs/fim/timestep.rs:469-482 opm_pid_dt."""
    e = np.ascontiguousarray(errors, dtype=np.float64)
    return lib().orc_pid_dt(dt, e.ctypes.data_as(_DP))


def iteration_count_dt(dt, newton_iterations):
    """rs/fim/timestep.rs:484-494 opm_iteration_count_dt."""
    return lib().orc_iteration_count_dt(dt, int(newton_iterations))


GROWTH_LIMITERS = ("opm-max-growth", "opm-pid", "opm-iter", "opm-restart-growth")


def growth_decision(dt, newton_iterations, errors, retries):
    """rs/fim/timestep.rs:502-528: (factor, limiter label)."""
    e = np.ascontiguousarray(errors, dtype=np.float64)
    lim = ctypes.c_int(0)
    f = lib().orc_growth_factor(dt, int(newton_iterations), e.ctypes.data_as(_DP), int(retries), ctypes.byref(lim))
    return f, GROWTH_LIMITERS[lim.value]


FAMILIES = ("water", "oil_component", "gas_component", "well_constraint", "perforation_flow")


def detect_oscillation(current, prev1, prev2):
    """rs/fim/newton/damping.rs:724-757 over {family: norm} dicts."""
    arrs = [np.array([d[f] for f in FAMILIES], dtype=np.float64) for d in (current, prev1, prev2)]
    return lib().orc_detect_oscillation(*(a.ctypes.data_as(_DP) for a in arrs))


def next_relaxation_factor(relax, oscillating):
    return lib().orc_next_relaxation_factor(relax, int(oscillating))


def relative_change(cells_a, regimes_a, cells_b, regimes_b):
    """rs/fim/timestep.rs:436-466 opm_relative_change between two cell states."""
    a = np.ascontiguousarray(np.ravel(cells_a), dtype=np.float64)
    b = np.ascontiguousarray(np.ravel(cells_b), dtype=np.float64)
    ra = np.ascontiguousarray(regimes_a, dtype=np.int32)
    rb = np.ascontiguousarray(regimes_b, dtype=np.int32)
    return lib().orc_relative_change(len(ra), a.ctypes.data_as(_DP), ra.ctypes.data_as(_IP), b.ctypes.data_as(_DP),
                                     rb.ctypes.data_as(_IP))


def ad_selftest():
    out = np.empty(16)
    lib().orc_ad_selftest(out.ctypes.data_as(_DP))
    return out


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


class OracleSimulator:
    """CPU reference-algorithm simulator with the reference's setter names."""

    FIELD = {"pressure": 0, "sw": 1, "so": 2, "sg": 3, "rs": 4}

    def __init__(self, nx, ny, nz, porosity):
        self._l = lib()
        self._h = self._l.orc_create(nx, ny, nz, porosity)
        if not self._h:
            raise ValueError("invalid grid dimensions")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz
        # Accumulated scalar settings (the reference splits these over several setters).
        self._fluid = dict(mu_o=1.0, mu_w=0.5, c_o=1e-5, c_w=3e-6, rho_o=800.0, rho_w=1000.0)
        self._flags = dict(three_phase=0, gravity=0, corey=21, fluid=1)
        self._ctl = dict(inj_rate=0, prod_rate=0, inj_enabled=1, tgt_inj=0.0, tgt_prod=0.0,
                         tgt_inj_surf=0.0, tgt_prod_surf=0.0, bhp_min=-100.0, bhp_max=2000.0)
        self._p = np.full(self.n, 300.0)
        self._sw = np.full(self.n, 0.3)
        self._sg = np.zeros(self.n)
        self._rs = np.zeros(self.n)
        self._phi = np.full(self.n, float(porosity))
        self._kx = np.full(self.n, 100.0)
        self._ky = np.full(self.n, 100.0)
        self._kz = np.full(self.n, 10.0)

    def __del__(self):
        try:
            if self._h:
                self._l.orc_destroy(self._h)
                self._h = None
        except Exception:
            pass

    # ---- setters (reference names, rs/frontend.rs) ----
    def _push_fluid(self):
        f = self._fluid
        _chk(self._l.orc_set_fluid(self._h, f["mu_o"], f["mu_w"], f["c_o"], f["c_w"], f["rho_o"], f["rho_w"]), "fluid")

    def _push_flags(self):
        f = self._flags
        _chk(self._l.orc_set_flags(self._h, f["three_phase"], f["gravity"], f["corey"], f["fluid"]), "flags")

    def _push_ctl(self):
        c = self._ctl
        _chk(self._l.orc_set_well_controls(self._h, c["inj_rate"], c["prod_rate"], c["inj_enabled"], c["tgt_inj"],
                                           c["tgt_prod"], c["tgt_inj_surf"], c["tgt_prod_surf"], c["bhp_min"],
                                           c["bhp_max"]), "controls")

    def _push_state(self):
        a, pa = _dp(self._p); b, pb = _dp(self._sw); c, pc = _dp(self._sg); d, pd = _dp(self._rs)
        _chk(self._l.orc_set_initial_state(self._h, pa, pb, pc, pd), "state")

    def _push_fields(self):
        a, pa = _dp(self._phi); b, pb = _dp(self._kx); c, pc = _dp(self._ky); d, pd = _dp(self._kz)
        _chk(self._l.orc_set_fields(self._h, pa, pb, pc, pd), "fields")

    def set_cell_dimensions_per_layer(self, dx, dy, dz):
        a, pa = _dp(dz)
        _chk(self._l.orc_set_cell_dimensions_per_layer(self._h, dx, dy, pa), "dims")

    def set_cell_dimensions(self, dx, dy, dz):
        self.set_cell_dimensions_per_layer(dx, dy, np.full(self.nz, float(dz)))

    def set_permeability_field(self, kx, ky, kz):
        self._kx, self._ky, self._kz = (np.asarray(v, dtype=np.float64).copy() for v in (kx, ky, kz))
        self._push_fields()

    def set_porosity_field(self, phi):
        self._phi = np.asarray(phi, dtype=np.float64).copy()
        self._push_fields()

    def set_capillary_scale_field(self, scale):
        a, pa = _dp(scale)
        _chk(self._l.orc_set_pc_scale(self._h, pa), "pc scale")

    def set_initial_pressure(self, p):
        self._p[:] = p
        _chk(self._l.orc_set_reference_pressure(self._h, float(p)), "reference pressure")
        self._push_state()

    def set_initial_saturation(self, sw):
        self._sw[:] = sw
        self._push_state()

    def set_initial_gas_saturation(self, sg):
        self._sg[:] = np.minimum(np.clip(sg, 0.0, 1.0), 1.0 - self._sw)
        self._push_state()

    def set_initial_rs(self, rs):
        self._rs[:] = rs
        self._push_state()

    def set_initial_state_arrays(self, p, sw, sg, rs):
        self._p[:] = p
        self._sw[:] = sw
        self._sg[:] = sg
        self._rs[:] = rs
        self._push_state()

    def set_fluid_properties(self, mu_o, mu_w):
        self._fluid.update(mu_o=mu_o, mu_w=mu_w)
        self._push_fluid()

    def set_fluid_compressibilities(self, c_o, c_w):
        self._fluid.update(c_o=c_o, c_w=c_w)
        self._push_fluid()

    def set_fluid_densities(self, rho_o, rho_w):
        self._fluid.update(rho_o=rho_o, rho_w=rho_w)
        self._push_fluid()

    def set_rock_properties(self, c_r, depth_reference_m, b_o, b_w):
        _chk(self._l.orc_set_rock_properties(self._h, c_r, depth_reference_m, b_o, b_w), "rock")

    def set_gas_fluid_properties(self, mu_g, c_g, rho_g):
        _chk(self._l.orc_set_gas_fluid_properties(self._h, mu_g, c_g, rho_g), "gas")

    def set_rel_perm_props(self, s_wc, s_or, n_w, n_o, k_rw_max, k_ro_max):
        _chk(self._l.orc_set_rel_perm_props(self._h, s_wc, s_or, n_w, n_o, k_rw_max, k_ro_max), "relperm")

    def set_three_phase_rel_perm_props(self, *args):
        _chk(self._l.orc_set_three_phase_rel_perm_props(self._h, *[float(a) for a in args]), "relperm3")

    def set_three_phase_scal_tables(self, tables):
        """rs/frontend.rs:1145-1158: {"swof": [{sw, krw, krow, pcow}], "sgof": [{sg, krg, krog, pcog}]}
        (rows may also be 3- or 4-sequences); the capillary columns are not used."""
        sw = scal_rows(tables["swof"], ("sw", "krw", "krow"))
        sg = scal_rows(tables["sgof"], ("sg", "krg", "krog"))
        _chk(self._l.orc_set_scal_tables(self._h, sw.ctypes.data_as(_DP), len(sw), sg.ctypes.data_as(_DP), len(sg)),
             "scal tables")

    def set_permeability_per_layer(self, kx, ky, kz):
        """rs/frontend.rs:1243-1285: one permeability per layer, copied to every cell of it."""
        nxy = self.n // len(kx)
        self.set_permeability_field(*(np.repeat(np.asarray(v, dtype=np.float64), nxy) for v in (kx, ky, kz)))

    def set_capillary_params(self, p_entry, lam):
        _chk(self._l.orc_set_capillary_params(self._h, p_entry, lam), "pc")

    def set_gas_oil_capillary_params(self, p_entry, lam):
        _chk(self._l.orc_set_gas_oil_capillary_params(self._h, p_entry, lam), "pcog")

    def set_gas_redissolution_enabled(self, on):
        _chk(self._l.orc_set_gas_redissolution(self._h, int(bool(on))), "gas_redissolution")

    def set_three_phase_mode_enabled(self, on):
        self._flags["three_phase"] = int(bool(on))
        self._push_flags()

    def set_gravity_enabled(self, on):
        self._flags["gravity"] = int(bool(on))
        self._push_flags()

    def set_fim_corey_table_points(self, pts):
        self._flags["corey"] = int(pts)
        self._push_flags()

    def set_injected_fluid(self, fluid):
        self._flags["fluid"] = {"water": 0, "gas": 1}[fluid.lower()]
        self._push_flags()

    def set_pvt_table(self, rows):
        rows = np.ascontiguousarray(np.asarray(rows, dtype=np.float64).reshape(-1, 6))
        _chk(self._l.orc_set_pvt_table(self._h, rows.ctypes.data_as(_DP), rows.shape[0]), "pvt")
        out = np.empty(self.n)
        self._l.orc_get_field(self._h, 4, out.ctypes.data_as(_DP))
        self._rs[:] = out

    def set_well_control_modes(self, inj, prod):
        self._ctl.update(inj_rate=int(inj.lower() == "rate"), prod_rate=int(prod.lower() == "rate"))
        self._push_ctl()

    def set_injector_enabled(self, on):
        self._ctl["inj_enabled"] = int(bool(on))
        self._push_ctl()

    def set_target_well_rates(self, inj, prod):
        self._ctl.update(tgt_inj=inj, tgt_prod=prod)
        self._push_ctl()

    def set_target_well_surface_rates(self, inj, prod):
        self._ctl.update(tgt_inj_surf=inj, tgt_prod_surf=prod)
        self._push_ctl()

    def set_well_bhp_limits(self, bhp_min, bhp_max):
        self._ctl.update(bhp_min=bhp_min, bhp_max=bhp_max)
        self._push_ctl()

    def add_well(self, i, j, k, bhp, rw, skin, injector):
        _chk(self._l.orc_add_well(self._h, i, j, k, bhp, rw, skin, int(injector), None), "add_well")

    def add_well_with_id(self, i, j, k, bhp, rw, skin, injector, well_id):
        _chk(self._l.orc_add_well(self._h, i, j, k, bhp, rw, skin, int(injector), well_id.encode()), "add_well")

    def set_well_schedule(self, well_id, mode, target, surface, limit, enabled):
        _chk(self._l.orc_set_well_schedule(self._h, well_id.encode(), int(mode.lower() == "rate"), target, surface,
                                           limit, int(enabled)), "schedule")

    def set_well_datum(self, well_id, datum, density):
        _chk(self._l.orc_set_well_datum(self._h, well_id.encode(), datum, density), "datum")

    def set_well_injected_fluid(self, well_id, fluid):
        _chk(self._l.orc_set_well_fluid(self._h, well_id.encode(), {"water": 0, "gas": 1}[fluid.lower()]), "fluid")

    def set_newton_options(self, max_iterations=20, min_iteration_index=2, residual_tol=0.0, linear_rel_tol=5e-3,
                           linear_max_iterations=20, allow_direct=True):
        _chk(self._l.orc_set_newton_options(self._h, max_iterations, min_iteration_index, residual_tol,
                                            linear_rel_tol, linear_max_iterations, int(allow_direct)), "options")

    # ---- stepping ----
    def step(self, dt):
        return self._l.orc_step(self._h, dt)

    def run_substep(self, dt):
        ni = ctypes.c_int(0)
        li = ctypes.c_long(0)
        rc = self._l.orc_run_substep(self._h, dt, ctypes.byref(ni), ctypes.byref(li))
        return rc, ni.value, li.value

    # ---- getters ----
    def get_field(self, name):
        out = np.empty(self.n)
        self._l.orc_get_field(self._h, self.FIELD[name], out.ctypes.data_as(_DP))
        return out

    def get_pressures(self):
        return self.get_field("pressure")

    def get_well_state(self):
        nw = self._l.orc_num_wells(self._h)
        npf = self._l.orc_num_perfs(self._h)
        bhp = np.empty(max(nw, 1))
        q = np.empty(max(npf, 1))
        self._l.orc_get_well_state(self._h, bhp.ctypes.data_as(_DP), q.ctypes.data_as(_DP))
        return bhp[:nw], q[:npf]

    def get_time(self):
        return self._l.orc_get_time(self._h)

    def stats(self):
        out = (ctypes.c_long * 5)()
        self._l.orc_get_stats(self._h, out)
        return dict(newton=out[0], linear=out[1], substeps=out[2], retries=out[3], bicgstab=out[4])

    def step_stats_history(self):
        """FimStepStats per step() call (rs/reporting.rs:141-243)."""
        out = []
        for i in range(self._l.orc_step_stats_count(self._h)):
            s = StepStats()
            assert self._l.orc_get_step_stats(self._h, i, ctypes.byref(s)) == 0
            rungs = (StepRung * max(s.n_rungs, 1))()
            k = self._l.orc_get_step_rungs(self._h, i, rungs, s.n_rungs)
            out.append(decode_step_stats(s, rungs[:k]))
        return out

    def accepted_dts(self):
        n = self._l.orc_get_accepted_dts(self._h, None, 0)
        out = np.empty(max(n, 1))
        self._l.orc_get_accepted_dts(self._h, out.ctypes.data_as(_DP), n)
        return out[:n]

    def last_warning(self):
        return self._l.orc_last_warning(self._h).decode()

    def initial_state(self):
        cs = np.empty(3 * self.n)
        rg = np.empty(self.n, dtype=np.int32)
        self._l.orc_get_initial_state(self._h, cs.ctypes.data_as(_DP), rg.ctypes.data_as(_IP))
        return cs.reshape(-1, 3), rg

    def initial_well_state(self):
        """(bhp per well, q per completion) of the initial iterate."""
        b = np.zeros(max(self._l.orc_num_wells(self._h), 1))
        q = np.zeros(max(self._l.orc_num_perfs(self._h), 1))
        self._l.orc_get_initial_well_state(self._h, b.ctypes.data_as(_DP), q.ctypes.data_as(_DP))
        return b[:self._l.orc_num_wells(self._h)].copy(), q[:self._l.orc_num_perfs(self._h)].copy()

    def assemble(self, dt, cell_state=None, regimes=None, bhp=None, q=None):
        """Residual and CSR Jacobian at (state, stored previous state)."""
        nu = ctypes.c_int(0)
        nnz = ctypes.c_int(0)
        keep = []

        def opt(a, ctype):
            if a is None:
                return None
            arr = np.ascontiguousarray(a, dtype=np.float64 if ctype is _DP else np.int32)
            keep.append(arr)
            return arr.ctypes.data_as(ctype)

        self._l.orc_assemble(self._h, dt, opt(None if cell_state is None else np.ravel(cell_state), _DP),
                             opt(regimes, _IP), opt(bhp, _DP), opt(q, _DP), ctypes.byref(nu), ctypes.byref(nnz))
        R = np.empty(nu.value)
        ptr = np.empty(nu.value + 1, dtype=np.int32)
        col = np.empty(max(nnz.value, 1), dtype=np.int32)
        val = np.empty(max(nnz.value, 1))
        self._l.orc_get_assembly(R.ctypes.data_as(_DP), ptr.ctypes.data_as(_IP), col.ctypes.data_as(_IP),
                                 val.ctypes.data_as(_DP))
        return R, (ptr, col[:nnz.value], val[:nnz.value])

    def producer_oil(self):
        """Surface oil produced by the producers over the accepted substeps."""
        return self._l.orc_producer_oil(self._h)

    # TimePointRates fields in order (rs/reporting.rs:246-292); sweep metrics NaN where None
    RATE_FIELDS = ("time", "total_production_oil", "total_production_liquid", "total_production_liquid_reservoir",
                   "total_injection", "total_injection_reservoir", "material_balance_error_m3",
                   "material_balance_error_oil_m3", "avg_reservoir_pressure", "avg_water_saturation",
                   "total_production_gas", "avg_gas_saturation", "material_balance_error_gas_m3", "producing_gor",
                   "producer_bhp_limited_fraction", "injector_bhp_limited_fraction",
                   "sweep_e_a", "sweep_e_v", "sweep_e_vol", "sweep_mobile_oil_recovered")

    def rate_history(self):
        """One report per accepted substep: {field: array over the entries}."""
        k = self._l.orc_rate_history_len(self._h)
        out = np.empty((max(k, 1), 20))
        self._l.orc_get_rate_history(self._h, out.ctypes.data_as(_DP), k)
        return {f: out[:k, j].copy() for j, f in enumerate(self.RATE_FIELDS)}

    def inventory_totals(self):
        """[water, oil, gas] in place (sc m3) over the stored state."""
        out = np.empty(3)
        self._l.orc_inventory_totals(self._h, out.ctypes.data_as(_DP))
        return out

    def residual_check(self, dt, cell_state, regimes, bhp, q):
        """The reference's acceptance measures at an iterate against the stored state:
        dict of scaled family peaks (water, oil, gas, well, perf), CNV[3] and MB[3], the
        violating pore-volume fraction and its 3 % rule, the OPM verdict with the strict and
        with the final-iteration tiers, whether every well passes its own test, and the row
        of each family's peak (-1 for a family with no rows)."""
        cs = np.ascontiguousarray(np.ravel(cell_state), dtype=np.float64)
        rg = np.ascontiguousarray(regimes, dtype=np.int32)
        b = np.ascontiguousarray(bhp, dtype=np.float64)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        out = np.empty(21)
        self._l.orc_residual_check(self._h, dt, cs.ctypes.data_as(_DP), rg.ctypes.data_as(_IP),
                                   b.ctypes.data_as(_DP), qq.ctypes.data_as(_DP), out.ctypes.data_as(_DP))
        return dict(families=out[:5].copy(), cnv=out[5:8].copy(), mb=out[8:11].copy(),
                    violating_pv_fraction=float(out[11]), pv_rule_relaxes=bool(out[12]),
                    accept=bool(out[13]), accept_final=bool(out[14]), wells_converged=bool(out[15]),
                    peak_rows=out[16:21].astype(np.int64))

    def scaled_residual(self, dt, cell_state, regimes, bhp, q):
        """|R| / scale per row at an iterate, in the reference's unknown numbering."""
        cs = np.ascontiguousarray(np.ravel(cell_state), dtype=np.float64)
        rg = np.ascontiguousarray(regimes, dtype=np.int32)
        b = np.ascontiguousarray(bhp, dtype=np.float64)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        out = np.empty(3 * self.n + self._l.orc_num_wells(self._h) + self._l.orc_num_perfs(self._h))
        self._l.orc_scaled_residual(self._h, dt, cs.ctypes.data_as(_DP), rg.ctypes.data_as(_IP),
                                    b.ctypes.data_as(_DP), qq.ctypes.data_as(_DP), out.ctypes.data_as(_DP))
        return out

    def newton_direction(self, dt, cell_state, regimes, bhp, q, rel_tol=1e-10, allow_direct=False):
        nw = self._l.orc_num_wells(self._h)
        npf = self._l.orc_num_perfs(self._h)
        dx = np.empty(3 * self.n + nw + npf)
        its = ctypes.c_int(0)
        cs = np.ascontiguousarray(np.ravel(cell_state), dtype=np.float64)
        rg = np.ascontiguousarray(regimes, dtype=np.int32)
        b = np.ascontiguousarray(bhp, dtype=np.float64)
        qq = np.ascontiguousarray(q, dtype=np.float64)
        rc = self._l.orc_newton_direction(self._h, dt, cs.ctypes.data_as(_DP), rg.ctypes.data_as(_IP),
                                          b.ctypes.data_as(_DP), qq.ctypes.data_as(_DP), rel_tol, int(allow_direct),
                                          dx.ctypes.data_as(_DP), ctypes.byref(its))
        return rc, dx, its.value

    def set_sweep_config(self, config):
        if config is None:
            _chk(self._l.orc_set_sweep_config(self._h, None, 0.0, 0.0, 0.0), "sweep")
            return
        _chk(self._l.orc_set_sweep_config(self._h, str(config["geometry"]).encode(), float(config["swept_threshold"]),
                                          float(config["initial_oil_saturation"]),
                                          float(config["residual_oil_saturation"])), "sweep")

    def relax_wells(self, cell_state, regimes, bhp, q):
        """The well relaxation at an iterate: (bhp, q) after it."""
        cs, rg, b, qq = self._iterate(cell_state, regimes, bhp, q)
        self._l.orc_relax_wells(self._h, cs.ctypes.data_as(_DP), rg.ctypes.data_as(_IP), b.ctypes.data_as(_DP),
                                qq.ctypes.data_as(_DP))
        return b, qq

    def set_fim_nested_well_solve(self, on):
        _chk(self._l.orc_set_nested_well_solve(self._h, int(bool(on))), "nested")

    @staticmethod
    def _iterate(cell_state, regimes, bhp, q):
        return (np.ascontiguousarray(np.ravel(cell_state), dtype=np.float64),
                np.ascontiguousarray(regimes, dtype=np.int32),
                np.array(bhp, dtype=np.float64), np.array(q, dtype=np.float64))

    def well_local_system(self, cell_state, regimes, bhp, q, well):
        """One well's local residual and Jacobian (unknowns: bhp, then q of each of its
        perforations) at an iterate, cells frozen."""
        cs, rg, b, qq = self._iterate(cell_state, regimes, bhp, q)
        npf = self._l.orc_num_perfs(self._h)
        r = np.zeros(npf + 1)
        J = np.zeros((npf + 1) ** 2)
        dim = self._l.orc_well_local_system(self._h, cs.ctypes.data_as(_DP), rg.ctypes.data_as(_IP),
                                            b.ctypes.data_as(_DP), qq.ctypes.data_as(_DP), well,
                                            r.ctypes.data_as(_DP), J.ctypes.data_as(_DP))
        return r[:dim].copy(), J[:dim * dim].reshape(dim, dim).copy()

    def solve_wells_locally(self, cell_state, regimes, bhp, q, max_iterations=-1):
        """Nested per-well Newton at an iterate: (bhp, q, converged, iterations, peak)."""
        cs, rg, b, qq = self._iterate(cell_state, regimes, bhp, q)
        nw = len(b)
        conv = np.zeros(nw, dtype=np.int32)
        its = np.zeros(nw, dtype=np.int32)
        peak = np.zeros(nw)
        self._l.orc_solve_wells_locally(self._h, cs.ctypes.data_as(_DP), rg.ctypes.data_as(_IP),
                                        b.ctypes.data_as(_DP), qq.ctypes.data_as(_DP), max_iterations,
                                        conv.ctypes.data_as(_IP), its.ctypes.data_as(_IP), peak.ctypes.data_as(_DP))
        return b, qq, conv.astype(bool), its, peak

    def probe(self, what, x, y=0.0, init=None):
        out = np.zeros(16)
        if init is not None:
            out[0] = init
        _chk(self._l.orc_probe(self._h, what, x, y, out.ctypes.data_as(_DP)), "probe")
        return out
