"""Python host mirror of the reference ``ReservoirSimulator`` over the HIP C ABI.

Method names, argument meaning and units follow the reference's Rust API
(src/lib/ressim/src/frontend.rs); a setter that the reference rejects with
``Err(String)`` raises ``ValueError`` here. All compute goes through
``libressim_fim.so`` (include/ressim_fim.h); loading fails loudly when the
library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import build as _build

_lib = None

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_S = ctypes.c_char_p

RSF_OK, RSF_ERR_INVALID, RSF_ERR_NOT_CONVERGED, RSF_ERR_HIP = 0, 1, 2, 3


class SolverOptions(ctypes.Structure):
    _fields_ = [("max_newton_iterations", ctypes.c_int), ("min_newton_iteration_index", ctypes.c_int),
                ("newton_residual_tol", ctypes.c_double), ("newton_mb_tol", ctypes.c_double),
                ("linear_rel_tol", ctypes.c_double), ("linear_abs_tol", ctypes.c_double),
                ("linear_max_iterations", ctypes.c_int), ("linear_restart", ctypes.c_int),
                ("precision", ctypes.c_int), ("ilu_ordering", ctypes.c_int),
                ("pressure_max_iterations", ctypes.c_int), ("pressure_rel_tol", ctypes.c_double)]


class StepReport(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("newton_iterations", ctypes.c_int),
                ("linear_iterations", ctypes.c_long), ("final_scaled_residual", ctypes.c_double)]


class FimStepReport(ctypes.Structure):
    """FimStepReport (fim/newton.rs:921-944) of one rsf_run_fim_timestep attempt."""
    _fields_ = [("converged", ctypes.c_int), ("newton_iterations", ctypes.c_int),
                ("final_residual_inf_norm", ctypes.c_double), ("final_material_balance_inf_norm", ctypes.c_double),
                ("final_update_inf_norm", ctypes.c_double), ("has_last_linear_report", ctypes.c_int),
                ("last_linear_iterations", ctypes.c_int), ("last_linear_converged", ctypes.c_int),
                ("linear_solve_count", ctypes.c_long), ("linear_iteration_count", ctypes.c_long),
                ("applied_update_count", ctypes.c_int), ("retry_factor", ctypes.c_double),
                ("failure_class", ctypes.c_int), ("failure_dominant_family", ctypes.c_int),
                ("total_time_ms", ctypes.c_double), ("assembly_ms", ctypes.c_double),
                ("linear_solve_time_ms", ctypes.c_double), ("failure_dominant_row", ctypes.c_long)]


class Stats(ctypes.Structure):
    _fields_ = [("newton_iterations", ctypes.c_long), ("linear_iterations", ctypes.c_long),
                ("substeps", ctypes.c_long), ("retries", ctypes.c_long), ("pressure_iterations", ctypes.c_long),
                ("pressure_solves", ctypes.c_long), ("assembly_ms", ctypes.c_double),
                ("linear_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


class KernelTiming(ctypes.Structure):
    _fields_ = [("spmv_ms", ctypes.c_double), ("spmv_bytes", ctypes.c_double), ("bilu_ms", ctypes.c_double),
                ("bilu_bytes", ctypes.c_double), ("bilu_launches", ctypes.c_int), ("pilu_ms", ctypes.c_double),
                ("pilu_bytes", ctypes.c_double), ("pspmv_ms", ctypes.c_double), ("pspmv_bytes", ctypes.c_double),
                ("assembly_ms", ctypes.c_double), ("assembly_bytes", ctypes.c_double), ("copy_ms", ctypes.c_double),
                ("copy_bytes", ctypes.c_double)]


class AdjointReport(ctypes.Structure):
    _fields_ = [("substeps", ctypes.c_int), ("linear_iterations", ctypes.c_long),
                ("pressure_iterations", ctypes.c_long), ("replay_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]


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


class TimePoint(ctypes.Structure):  # TimePointRates, rs/reporting.rs:246-292; sweep: NaN where None
    _fields_ = [(f, ctypes.c_double) for f in (
        "time", "total_production_oil", "total_production_liquid", "total_production_liquid_reservoir",
        "total_injection", "total_injection_reservoir", "material_balance_error_m3", "material_balance_error_oil_m3",
        "avg_reservoir_pressure", "avg_water_saturation", "total_production_gas", "avg_gas_saturation",
        "material_balance_error_gas_m3", "producing_gor", "producer_bhp_limited_fraction",
        "injector_bhp_limited_fraction", "sweep_e_a", "sweep_e_v", "sweep_e_vol", "sweep_mobile_oil_recovered")]


class LinearReport(ctypes.Structure):
    _fields_ = [("converged", ctypes.c_int), ("iterations", ctypes.c_int), ("residual_norm", ctypes.c_double),
                ("rhs_norm", ctypes.c_double)]


class KernelProbe(ctypes.Structure):
    _fields_ = [("launches", ctypes.c_long), ("total_ms", ctypes.c_double), ("avg_ms", ctypes.c_double),
                ("bytes_per_launch", ctypes.c_double), ("kernel", ctypes.c_char * 64)]


# every function the header declares, with its ctypes signature
SIGNATURES = {
    "rsf_create": (_I, [_I, _I, _I, _D, ctypes.POINTER(_P)]),
    "rsf_destroy": (None, [_P]),
    "rsf_set_cell_dimensions": (_I, [_P, _D, _D, _D]),
    "rsf_set_cell_dimensions_per_layer": (_I, [_P, _D, _D, _DP, _I]),
    "rsf_set_permeability_field": (_I, [_P, _DP, _DP, _DP, _I]),
    "rsf_set_perm_field_log": (_I, [_P, _DP, _D, _I]),
    "rsf_set_permeability_per_layer": (_I, [_P, _DP, _DP, _DP, _I]),
    "rsf_set_three_phase_scal_tables": (_I, [_P, _DP, _I, _DP, _I]),
    "rsf_set_porosity_field": (_I, [_P, _DP, _I]),
    "rsf_set_capillary_scale_field": (_I, [_P, _DP, _I]),
    "rsf_set_initial_pressure": (_I, [_P, _D]),
    "rsf_set_initial_saturation": (_I, [_P, _D]),
    "rsf_set_initial_gas_saturation": (_I, [_P, _D]),
    "rsf_set_initial_rs": (_I, [_P, _D]),
    "rsf_set_initial_state_arrays": (_I, [_P, _DP, _DP, _DP, _DP, _I]),
    "rsf_set_fluid_properties": (_I, [_P, _D, _D]),
    "rsf_set_fluid_compressibilities": (_I, [_P, _D, _D]),
    "rsf_set_fluid_densities": (_I, [_P, _D, _D]),
    "rsf_set_rock_properties": (_I, [_P, _D, _D, _D, _D]),
    "rsf_set_gas_fluid_properties": (_I, [_P, _D, _D, _D]),
    "rsf_set_rel_perm_props": (_I, [_P, _D, _D, _D, _D, _D, _D]),
    "rsf_set_three_phase_rel_perm_props": (_I, [_P] + [_D] * 11),
    "rsf_set_capillary_params": (_I, [_P, _D, _D]),
    "rsf_set_gas_oil_capillary_params": (_I, [_P, _D, _D]),
    "rsf_set_three_phase_mode_enabled": (_I, [_P, _I]),
    "rsf_set_gravity_enabled": (_I, [_P, _I]),
    "rsf_set_gas_redissolution_enabled": (_I, [_P, _I]),
    "rsf_set_fim_corey_table_points": (_I, [_P, _I]),
    "rsf_set_fim_nested_well_solve": (_I, [_P, _I]),
    "rsf_set_injected_fluid": (_I, [_P, _S]),
    "rsf_set_pvt_table": (_I, [_P, _DP, _I]),
    "rsf_set_well_control_modes": (_I, [_P, _S, _S]),
    "rsf_set_injector_enabled": (_I, [_P, _I]),
    "rsf_set_target_well_rates": (_I, [_P, _D, _D]),
    "rsf_set_target_well_surface_rates": (_I, [_P, _D, _D]),
    "rsf_set_well_bhp_limits": (_I, [_P, _D, _D]),
    "rsf_add_well": (_I, [_P, _I, _I, _I, _D, _D, _D, _I]),
    "rsf_add_well_with_id": (_I, [_P, _I, _I, _I, _D, _D, _D, _I, _S]),
    "rsf_set_well_schedule": (_I, [_P, _S, _S, _D, _D, _D, _I]),
    "rsf_set_well_datum": (_I, [_P, _S, _D, _D]),
    "rsf_set_well_injected_fluid": (_I, [_P, _S, _S]),
    "rsf_default_solver_options": (_I, [ctypes.POINTER(SolverOptions)]),
    "rsf_set_solver_options": (_I, [_P, ctypes.POINTER(SolverOptions)]),
    "rsf_step": (_I, [_P, _D]),
    "rsf_run_timestep": (_I, [_P, _D, ctypes.POINTER(StepReport)]),
    "rsf_run_timestep_from": (_I, [_P, _D, _DP, _IP, _DP, _DP, ctypes.POINTER(StepReport)]),
    "rsf_run_fim_timestep": (_I, [_P, _D, _DP, _IP, _DP, _IP, _DP, _DP, ctypes.POINTER(SolverOptions),
                                  ctypes.POINTER(FimStepReport), _DP, _IP, _DP, _DP]),
    "rsf_get_pressures": (_I, [_P, _DP, _I]),
    "rsf_get_sat_water": (_I, [_P, _DP, _I]),
    "rsf_get_sat_oil": (_I, [_P, _DP, _I]),
    "rsf_get_sat_gas": (_I, [_P, _DP, _I]),
    "rsf_get_rs": (_I, [_P, _DP, _I]),
    "rsf_num_wells": (_I, [_P]),
    "rsf_num_perforations": (_I, [_P]),
    "rsf_get_well_state": (_I, [_P, _DP, _I, _DP, _I]),
    "rsf_get_time": (_D, [_P]),
    "rsf_get_stats": (_I, [_P, ctypes.POINTER(Stats)]),
    "rsf_get_accepted_dts": (_I, [_P, _DP, _I]),
    "rsf_step_stats_count": (_I, [_P]),
    "rsf_debug_classify_retry": (_I, [_I, _I, _I, _I, _D, _D]),
    "rsf_debug_dominant_family": (_I, [_DP, _I]),
    "rsf_get_step_stats": (_I, [_P, _I, ctypes.POINTER(StepStats)]),
    "rsf_get_step_rungs": (_I, [_P, _I, ctypes.POINTER(StepRung), _I]),
    "rsf_last_warning": (_I, [_P, ctypes.c_char_p, _I]),
    "rsf_debug_assemble": (_I, [_P, _D, _DP, _IP, _DP, _DP, _I, _DP, _DP]),
    "rsf_debug_solve_wells_locally": (_I, [_P, _DP, _IP, _DP, _DP, _I, _IP, _IP, _DP]),
    "rsf_debug_newton_direction": (_I, [_P, _D, _DP, _IP, _DP, _DP, _DP, _IP]),
    "rsf_debug_newton_record": (_I, [_P, _D, _DP, _IP, _DP, _DP, _DP]),
    "rsf_debug_initial_state": (_I, [_P, _DP, _IP]),
    "rsf_debug_last_iterate": (_I, [_P, _DP, _IP, _DP, _I, _DP, _I]),
    "rsf_profile_kernels": (_I, [_P, _I, ctypes.POINTER(KernelTiming)]),
    "rsf_set_kernel_probe": (_I, [_P, _I]),
    "rsf_set_ilu_sweep": (_I, [_P, _I, _I, _I]),
    "rsf_adjoint_record": (_I, [_P, _I]),
    "rsf_adjoint_objective": (_I, [_P, _DP, ctypes.POINTER(ctypes.c_int)]),
    "rsf_adjoint_gradient": (_I, [_P, _DP, _DP, _I, ctypes.POINTER(AdjointReport)]),
    "rsf_get_kernel_probe": (_I, [_P, ctypes.POINTER(KernelProbe)]),
    "rsf_set_capture_dir": (_I, [_P, _S, _I]),
    "rsf_rate_history_len": (_I, [_P]),
    "rsf_get_rate_history": (_I, [_P, ctypes.POINTER(TimePoint), _I]),
    "rsf_get_rate_history_since": (_I, [_P, _I, ctypes.POINTER(TimePoint), _I]),
    "rsf_set_sweep_config": (_I, [_P, _S, _D, _D, _D]),
    "rsf_load_state": (_I, [_P, _D, _DP, _DP, _DP, _DP, _DP, _I, _DP, _I, ctypes.POINTER(TimePoint), _I]),
    "rsf_solve_captured_system": (_I, [_P, _I, _IP, _IP, _DP, _DP, _DP, ctypes.POINTER(LinearReport)]),
    "rsf_cnv_mb": (_I, [_P, _P, _P, _I, _I, _P, _P]),
    "rsf_cnv_mb_noalloc": (_I, [_P, _P, _P, _I, _I, _P, _P]),
    "rsf_cnv_mb_last_error": (_I, [ctypes.c_char_p, _I]),
    "rsf_version": (_S, []),
}

CNV_MB_FIELDS = ("cnv", "mb", "violating_pv_fraction", "pv_rule_relaxes", "would_accept", "would_accept_strict")


def _cnv_mb_call(fn, residual, pore_volumes, fvf, relax_final, stream):
    """Shared body of cnv_mb: device tensors in, the CnvMbDiagnostics scalars out."""
    import torch

    n = int(pore_volumes.numel())
    for name, t, size in (("residual", residual, 3 * n), ("fvf", fvf, 3 * n), ("pore_volumes", pore_volumes, n)):
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == size):
            raise ValueError(f"{name}: expected a contiguous float64 device tensor of {size} values")
    out = torch.empty(10, dtype=torch.float64, device=pore_volumes.device)
    stream = stream if stream is not None else torch.cuda.current_stream()
    st = ctypes.c_void_p(stream.cuda_stream)
    rc = fn(ctypes.c_void_p(residual.data_ptr()), ctypes.c_void_p(pore_volumes.data_ptr()),
            ctypes.c_void_p(fvf.data_ptr()), n, int(bool(relax_final)), ctypes.c_void_p(out.data_ptr()), st)
    if rc != 0:
        buf = ctypes.create_string_buffer(512)
        lib().rsf_cnv_mb_last_error(buf, 512)
        raise RuntimeError(f"{fn.__name__} failed ({rc}): {buf.value.decode()}")
    stream.synchronize()
    o = out.cpu().numpy()
    return dict(cnv=o[0:3].copy(), mb=o[3:6].copy(), violating_pv_fraction=float(o[6]),
                pv_rule_relaxes=bool(o[7]), would_accept=bool(o[8]), would_accept_strict=bool(o[9]))


def cnv_mb(residual, pore_volumes, fvf, relax_final=False, stream=None):
    """cnv_mb_from_parts (rs/fim/newton/convergence.rs:478-612) through rsf_cnv_mb.
    residual: 3n cell rows (index 3 i + c), pore_volumes: n, fvf: 3n [B_w, B_o, B_g] per
    cell, all float64 device tensors; stream: a torch.cuda.Stream (the current one if None)."""
    return _cnv_mb_call(lib().rsf_cnv_mb, residual, pore_volumes, fvf, relax_final, stream)


def cnv_mb_noalloc(residual, pore_volumes, fvf, relax_final=False, stream=None):
    """cnv_mb through rsf_cnv_mb_noalloc: no device allocation by the library (its block
    records are a fixed device array shared by every call on the device)."""
    return _cnv_mb_call(lib().rsf_cnv_mb_noalloc, residual, pore_volumes, fvf, relax_final, stream)


def debug_classify_retry(have_linear, converged, direct=False, coarse_applications=0, average_reduction=1.0,
                         last_reduction=1.0):
    """rsf_debug_classify_retry: the retry class label of a rejected attempt's last linear
    solve (rs/fim/newton.rs:439-481)."""
    return RETRY_CLASSES[lib().rsf_debug_classify_retry(int(have_linear), int(converged), int(direct),
                                                        int(coarse_applications), float(average_reduction),
                                                        float(last_reduction))]


def debug_dominant_family(family_maxima, with_wells):
    """rsf_debug_dominant_family: label of the dominant residual family
    (rs/fim/newton/convergence.rs:793-806)."""
    f = np.ascontiguousarray(family_maxima, dtype=np.float64)
    assert f.shape == (5,)
    return RESIDUAL_FAMILIES[lib().rsf_debug_dominant_family(f.ctypes.data_as(_DP), int(with_wells))]


def library_path() -> str:
    return _build.LIB


def lib():
    """Load libressim_fim.so, building it first when it is absent or was built from other
    sources than the tree's (content hash). Raises if it cannot be built or loaded."""
    global _lib
    if _lib is None:
        path = _build.LIB
        if _build.needs_build():
            _build.build()
        _lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def _scal_rows(rows, keys):
    """SWOF/SGOF rows as an (n, 4) array: saturation, kr, kro and the capillary value (NaN
    for None), from the reference's row objects (dicts) or from 3- or 4-sequences."""
    out = []
    for r in rows:
        if isinstance(r, dict):
            v = [r.get(k) for k in keys]
        else:
            v = list(r) + [None] * (4 - len(r))
        out.append([float("nan") if x is None else float(x) for x in v[:4]])
    return np.ascontiguousarray(np.array(out, dtype=np.float64).reshape(-1, 4))


def _arr(a, dtype=np.float64):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a


def _dp(a):
    return a.ctypes.data_as(_DP)


def default_solver_options() -> SolverOptions:
    o = SolverOptions()
    lib().rsf_default_solver_options(ctypes.byref(o))
    return o


class ReservoirSimulator:
    """HIP FIM simulator with the reference's method names (rs/frontend.rs)."""

    def __init__(self, nx, ny, nz, porosity):
        self._l = lib()
        h = _P()
        rc = self._l.rsf_create(int(nx), int(ny), int(nz), float(porosity), ctypes.byref(h))
        if rc != RSF_OK:
            raise ValueError("invalid grid dimensions or porosity")
        self._h = h
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.n = self.nx * self.ny * self.nz

    def close(self):
        if getattr(self, "_h", None):
            self._l.rsf_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc == RSF_OK:
            return
        msg = self.last_warning()
        if rc == RSF_ERR_INVALID:
            raise ValueError(f"{what}: invalid argument {msg}".strip())
        if rc == RSF_ERR_HIP:
            raise RuntimeError(f"{what}: HIP error: {msg}")
        raise RuntimeError(f"{what}: status {rc} {msg}")

    # ---- setters (reference names) ----
    def set_cell_dimensions(self, dx, dy, dz):
        self._chk(self._l.rsf_set_cell_dimensions(self._h, dx, dy, dz), "set_cell_dimensions")

    def set_cell_dimensions_per_layer(self, dx, dy, dz):
        a = _arr(dz)
        self._chk(self._l.rsf_set_cell_dimensions_per_layer(self._h, dx, dy, _dp(a), a.size), "dims")

    def set_permeability_field(self, kx, ky, kz):
        a, b, c = _arr(kx), _arr(ky), _arr(kz)
        self._chk(self._l.rsf_set_permeability_field(self._h, _dp(a), _dp(b), _dp(c), a.size), "perm")

    def set_permeability_per_layer(self, kx, ky, kz):
        """rs/frontend.rs:1243-1285: one permeability per layer, copied to every cell of it."""
        a, b, c = _arr(kx), _arr(ky), _arr(kz)
        self._chk(self._l.rsf_set_permeability_per_layer(self._h, _dp(a), _dp(b), _dp(c), a.size), "perm per layer")

    def set_perm_field_log(self, ln_k, kz_over_kh=0.1):
        """EXTENSION: kx = ky = exp(ln k), kz = kz_over_kh kx."""
        a = _arr(ln_k)
        self._chk(self._l.rsf_set_perm_field_log(self._h, _dp(a), float(kz_over_kh), a.size), "perm log")

    def set_porosity_field(self, phi):
        a = _arr(phi)
        self._chk(self._l.rsf_set_porosity_field(self._h, _dp(a), a.size), "porosity")

    def set_capillary_scale_field(self, scale):
        a = _arr(scale)
        self._chk(self._l.rsf_set_capillary_scale_field(self._h, _dp(a), a.size), "pc scale")

    def set_initial_pressure(self, p):
        self._chk(self._l.rsf_set_initial_pressure(self._h, p), "initial pressure")

    def set_initial_saturation(self, sw):
        self._chk(self._l.rsf_set_initial_saturation(self._h, sw), "initial saturation")

    def set_initial_gas_saturation(self, sg):
        self._chk(self._l.rsf_set_initial_gas_saturation(self._h, sg), "initial gas saturation")

    def set_initial_rs(self, rs):
        self._chk(self._l.rsf_set_initial_rs(self._h, rs), "initial rs")

    def set_initial_state_arrays(self, p, sw, sg, rs):
        a, b, c, d = _arr(p), _arr(sw), _arr(sg), _arr(rs)
        self._chk(self._l.rsf_set_initial_state_arrays(self._h, _dp(a), _dp(b), _dp(c), _dp(d), a.size), "state")

    def set_fluid_properties(self, mu_o, mu_w):
        self._chk(self._l.rsf_set_fluid_properties(self._h, mu_o, mu_w), "fluid")

    def set_fluid_compressibilities(self, c_o, c_w):
        self._chk(self._l.rsf_set_fluid_compressibilities(self._h, c_o, c_w), "compressibility")

    def set_fluid_densities(self, rho_o, rho_w):
        self._chk(self._l.rsf_set_fluid_densities(self._h, rho_o, rho_w), "densities")

    def set_rock_properties(self, c_r, depth_reference_m, b_o, b_w):
        self._chk(self._l.rsf_set_rock_properties(self._h, c_r, depth_reference_m, b_o, b_w), "rock")

    def set_gas_fluid_properties(self, mu_g, c_g, rho_g):
        self._chk(self._l.rsf_set_gas_fluid_properties(self._h, mu_g, c_g, rho_g), "gas")

    def set_rel_perm_props(self, s_wc, s_or, n_w, n_o, k_rw_max, k_ro_max):
        self._chk(self._l.rsf_set_rel_perm_props(self._h, s_wc, s_or, n_w, n_o, k_rw_max, k_ro_max), "relperm")

    def set_three_phase_rel_perm_props(self, *args):
        self._chk(self._l.rsf_set_three_phase_rel_perm_props(self._h, *[float(a) for a in args]), "relperm3")

    def set_three_phase_scal_tables(self, tables):
        """rs/frontend.rs:1145-1158 setThreePhaseScalTables: {"swof": [{sw, krw, krow, pcow}],
        "sgof": [{sg, krg, krog, pcog}]} (rows may also be 3- or 4-sequences; a missing or None
        capillary value is NaN). Raises ValueError with the reference's message."""
        sw = _scal_rows(tables["swof"], ("sw", "krw", "krow", "pcow"))
        sg = _scal_rows(tables["sgof"], ("sg", "krg", "krog", "pcog"))
        self._chk(self._l.rsf_set_three_phase_scal_tables(self._h, _dp(sw), sw.shape[0], _dp(sg), sg.shape[0]),
                  "scal tables")

    def set_capillary_params(self, p_entry, lam):
        self._chk(self._l.rsf_set_capillary_params(self._h, p_entry, lam), "pc")

    def set_gas_oil_capillary_params(self, p_entry, lam):
        self._chk(self._l.rsf_set_gas_oil_capillary_params(self._h, p_entry, lam), "pcog")

    def set_three_phase_mode_enabled(self, on):
        self._chk(self._l.rsf_set_three_phase_mode_enabled(self._h, int(bool(on))), "three phase")

    def set_gas_redissolution_enabled(self, on):
        self._chk(self._l.rsf_set_gas_redissolution_enabled(self._h, int(bool(on))), "set_gas_redissolution_enabled")

    def set_gravity_enabled(self, on):
        self._chk(self._l.rsf_set_gravity_enabled(self._h, int(bool(on))), "gravity")

    def set_fim_nested_well_solve(self, on):
        """frontend.rs:375: nested per-well Newton in place of the well relaxation."""
        self._chk(self._l.rsf_set_fim_nested_well_solve(self._h, int(bool(on))), "nested well solve")

    def set_fim_corey_table_points(self, pts):
        self._chk(self._l.rsf_set_fim_corey_table_points(self._h, int(pts)), "corey points")

    def set_injected_fluid(self, fluid):
        self._chk(self._l.rsf_set_injected_fluid(self._h, fluid.encode()), "injected fluid")

    def set_pvt_table(self, rows):
        r = _arr(np.asarray(rows, dtype=np.float64).reshape(-1, 6))
        self._chk(self._l.rsf_set_pvt_table(self._h, _dp(r), r.shape[0]), "pvt table")

    def set_well_control_modes(self, inj, prod):
        self._chk(self._l.rsf_set_well_control_modes(self._h, inj.encode(), prod.encode()), "control modes")

    def set_injector_enabled(self, on):
        self._chk(self._l.rsf_set_injector_enabled(self._h, int(bool(on))), "injector enabled")

    def set_target_well_rates(self, inj, prod):
        self._chk(self._l.rsf_set_target_well_rates(self._h, inj, prod), "target rates")

    def set_target_well_surface_rates(self, inj, prod):
        self._chk(self._l.rsf_set_target_well_surface_rates(self._h, inj, prod), "surface rates")

    def set_well_bhp_limits(self, bhp_min, bhp_max):
        self._chk(self._l.rsf_set_well_bhp_limits(self._h, bhp_min, bhp_max), "bhp limits")

    def add_well(self, i, j, k, bhp, rw, skin, injector):
        self._chk(self._l.rsf_add_well(self._h, i, j, k, bhp, rw, skin, int(bool(injector))), "add_well")

    def add_well_with_id(self, i, j, k, bhp, rw, skin, injector, well_id):
        self._chk(self._l.rsf_add_well_with_id(self._h, i, j, k, bhp, rw, skin, int(bool(injector)),
                                               well_id.encode()), "add_well_with_id")

    def set_well_schedule(self, well_id, mode, target, surface, limit, enabled):
        self._chk(self._l.rsf_set_well_schedule(self._h, well_id.encode(), mode.encode(), target, surface, limit,
                                                int(bool(enabled))), "schedule")

    def set_well_datum(self, well_id, datum, density):
        self._chk(self._l.rsf_set_well_datum(self._h, well_id.encode(), datum, density), "datum")

    def set_well_injected_fluid(self, well_id, fluid):
        self._chk(self._l.rsf_set_well_injected_fluid(self._h, well_id.encode(), fluid.encode()), "well fluid")

    def set_solver_options(self, **kw):
        o = default_solver_options()
        for k, v in kw.items():
            setattr(o, k, v)
        self._chk(self._l.rsf_set_solver_options(self._h, ctypes.byref(o)), "solver options")

    # ---- stepping ----
    def step(self, dt):
        """Advance one report step; returns the status code (0 = converged)."""
        rc = self._l.rsf_step(self._h, float(dt))
        if rc in (RSF_ERR_INVALID, RSF_ERR_HIP):
            self._chk(rc, "step")
        return rc

    def run_timestep(self, dt):
        rep = StepReport()
        rc = self._l.rsf_run_timestep(self._h, float(dt), ctypes.byref(rep))
        if rc in (RSF_ERR_INVALID, RSF_ERR_HIP):
            self._chk(rc, "run_timestep")
        return rc, rep

    def run_timestep_from(self, dt, cell_state=None, regimes=None, bhp=None, q=None):
        """run_fim_timestep with an initial iterate of the caller's (None: the stored state's own)."""
        keep = []

        def opt(v, dtype, ctype):
            if v is None:
                return None
            arr = np.ascontiguousarray(np.ravel(v), dtype=dtype)
            keep.append(arr)
            return arr.ctypes.data_as(ctype)

        rep = StepReport()
        rc = self._l.rsf_run_timestep_from(self._h, float(dt), opt(cell_state, np.float64, _DP),
                                           opt(regimes, np.int32, _IP), opt(bhp, np.float64, _DP),
                                           opt(q, np.float64, _DP), ctypes.byref(rep))
        if rc in (RSF_ERR_INVALID, RSF_ERR_HIP):
            self._chk(rc, "run_timestep_from")
        return rc, rep

    def run_fim_timestep(self, dt, previous=None, initial=None, options=None):
        """run_fim_timestep(sim, previous_state, initial_iterate, dt, options): one Newton attempt
        that commits nothing. previous / initial: (cell_state [N, 3], regimes) or None, initial
        may also carry (cell_state, regimes, bhp, q); options: a dict of solver options (the
        handle's when None). Returns (rc, report dict, (cell_state [N, 3], regimes, bhp, q))
        with the attempt's final iterate."""
        keep = []

        def opt(v, dtype, ctype):
            if v is None:
                return None
            arr = np.ascontiguousarray(np.ravel(v), dtype=dtype)
            keep.append(arr)
            return arr.ctypes.data_as(ctype)

        prev = previous if previous is not None else (None, None)
        init = tuple(initial) + (None,) * (4 - len(initial)) if initial is not None else (None,) * 4
        o = None
        if options is not None:
            o = default_solver_options()
            for k, v in options.items():
                setattr(o, k, v)
        nw, npf = self._l.rsf_num_wells(self._h), self._l.rsf_num_perforations(self._h)
        cs = np.empty(3 * self.n)
        rg = np.empty(self.n, dtype=np.int32)
        b = np.empty(max(nw, 1))
        q = np.empty(max(npf, 1))
        rep = FimStepReport()
        rc = self._l.rsf_run_fim_timestep(
            self._h, float(dt), opt(prev[0], np.float64, _DP), opt(prev[1], np.int32, _IP),
            opt(init[0], np.float64, _DP), opt(init[1], np.int32, _IP), opt(init[2], np.float64, _DP),
            opt(init[3], np.float64, _DP), ctypes.byref(o) if o is not None else None, ctypes.byref(rep), _dp(cs),
            rg.ctypes.data_as(_IP), _dp(b), _dp(q))
        if rc in (RSF_ERR_INVALID, RSF_ERR_HIP):
            self._chk(rc, "run_fim_timestep")
        report = {f: getattr(rep, f) for f, _ in FimStepReport._fields_}
        return rc, report, (cs.reshape(self.n, 3), rg, b[:nw], q[:npf])

    # ---- getters ----
    def _field(self, fn):
        out = np.empty(self.n)
        self._chk(fn(self._h, _dp(out), self.n), "getter")
        return out

    def get_pressures(self):
        return self._field(self._l.rsf_get_pressures)

    def get_sat_water(self):
        return self._field(self._l.rsf_get_sat_water)

    def get_sat_oil(self):
        return self._field(self._l.rsf_get_sat_oil)

    def get_sat_gas(self):
        return self._field(self._l.rsf_get_sat_gas)

    def get_rs(self):
        return self._field(self._l.rsf_get_rs)

    def get_field(self, name):
        return {"pressure": self.get_pressures, "sw": self.get_sat_water, "so": self.get_sat_oil,
                "sg": self.get_sat_gas, "rs": self.get_rs}[name]()

    def get_well_state(self):
        nw = self._l.rsf_num_wells(self._h)
        npf = self._l.rsf_num_perforations(self._h)
        bhp = np.empty(max(nw, 1))
        q = np.empty(max(npf, 1))
        self._chk(self._l.rsf_get_well_state(self._h, _dp(bhp), nw, _dp(q), npf), "well state")
        return bhp[:nw], q[:npf]

    def get_time(self):
        return self._l.rsf_get_time(self._h)

    def stats(self):
        s = Stats()
        self._chk(self._l.rsf_get_stats(self._h, ctypes.byref(s)), "stats")
        return {f: getattr(s, f) for f, _ in Stats._fields_}

    def accepted_dts(self):
        n = self._l.rsf_get_accepted_dts(self._h, None, 0)
        out = np.empty(max(n, 1))
        self._l.rsf_get_accepted_dts(self._h, _dp(out), n)
        return out[:n]

    def step_stats_history(self):
        """getFimStepStatsHistory (rs/frontend.rs:597-600): one FimStepStats per step() call."""
        out = []
        for i in range(self._l.rsf_step_stats_count(self._h)):
            s = StepStats()
            self._chk(self._l.rsf_get_step_stats(self._h, i, ctypes.byref(s)), "step_stats")
            rungs = (StepRung * max(s.n_rungs, 1))()
            k = self._l.rsf_get_step_rungs(self._h, i, rungs, s.n_rungs)
            if k < 0:
                self._chk(k, "step_rungs")
            out.append(decode_step_stats(s, rungs[:k]))
        return out

    def last_step_stats(self):
        """getLastFimStepStats (rs/frontend.rs:592-595): the last step's record, or None."""
        h = self.step_stats_history()
        return h[-1] if h else None

    # the reference's own method names (rs/frontend.rs:592-600)
    get_last_fim_step_stats = last_step_stats
    get_fim_step_stats_history = step_stats_history

    def last_warning(self):
        buf = ctypes.create_string_buffer(1024)
        self._l.rsf_last_warning(self._h, buf, 1024)
        return buf.value.decode(errors="replace")

    # ---- low-level entry points (parity harness) ----
    def debug_assemble(self, dt, cell_state=None, regimes=None, bhp=None, q=None, eliminate_wells=False):
        nw = self._l.rsf_num_wells(self._h)
        npf = self._l.rsf_num_perforations(self._h)
        R = np.empty(3 * self.n + nw + npf)
        A = np.empty(63 * self.n)
        keep = []

        def opt(a, dtype, ctype):
            if a is None:
                return None
            arr = np.ascontiguousarray(np.ravel(a), dtype=dtype)
            keep.append(arr)
            return arr.ctypes.data_as(ctype)

        self._chk(self._l.rsf_debug_assemble(self._h, dt, opt(cell_state, np.float64, _DP),
                                             opt(regimes, np.int32, _IP), opt(bhp, np.float64, _DP),
                                             opt(q, np.float64, _DP), int(bool(eliminate_wells)), _dp(R), _dp(A)),
                  "debug_assemble")
        return R, A.reshape(7, 9, self.n)

    def debug_newton_record(self, dt, cell_state=None, regimes=None, bhp=None, q=None):
        """rsf_debug_newton_record: the device Newton record at an iterate, in the layout of
        the oracle's residual_check (families, CNV, MB, violating pore-volume fraction and
        its 3 % rule, the OPM verdict with the strict and the final-iteration tiers, the
        wells' own test) plus the squared norm of the full residual and the row of each
        family's peak (-1 for a family with no rows)."""
        out = np.empty(22)
        keep = []

        def opt(a, dtype, ctype):
            if a is None:
                return None
            arr = np.ascontiguousarray(np.ravel(a), dtype=dtype)
            keep.append(arr)
            return arr.ctypes.data_as(ctype)

        self._chk(self._l.rsf_debug_newton_record(self._h, dt, opt(cell_state, np.float64, _DP),
                                                  opt(regimes, np.int32, _IP), opt(bhp, np.float64, _DP),
                                                  opt(q, np.float64, _DP), _dp(out)), "debug_newton_record")
        return dict(families=out[:5].copy(), cnv=out[5:8].copy(), mb=out[8:11].copy(),
                    violating_pv_fraction=float(out[11]), pv_rule_relaxes=bool(out[12]),
                    accept=bool(out[13]), accept_final=bool(out[14]), wells_converged=bool(out[15]),
                    residual_norm2=float(out[16]), peak_rows=out[17:22].astype(np.int64))

    def debug_solve_wells_locally(self, cell_state, regimes, bhp, q, max_iterations=-1):
        """Nested per-well Newton at an iterate (cells frozen): (bhp, q, converged,
        iterations, peak) after the solve."""
        nw = self._l.rsf_num_wells(self._h)
        npf = self._l.rsf_num_perforations(self._h)
        b = np.array(bhp, dtype=np.float64).reshape(nw)
        qq = np.array(q, dtype=np.float64).reshape(npf)
        cs = None if cell_state is None else np.ascontiguousarray(np.ravel(cell_state), dtype=np.float64)
        rg = None if regimes is None else np.ascontiguousarray(regimes, dtype=np.int32)
        conv = np.zeros(max(nw, 1), dtype=np.int32)
        its = np.zeros(max(nw, 1), dtype=np.int32)
        peak = np.zeros(max(nw, 1))
        self._chk(self._l.rsf_debug_solve_wells_locally(
            self._h, None if cs is None else _dp(cs), None if rg is None else rg.ctypes.data_as(_IP), _dp(b),
            _dp(qq), int(max_iterations), conv.ctypes.data_as(_IP), its.ctypes.data_as(_IP), _dp(peak)),
            "debug_solve_wells_locally")
        return b, qq, conv[:nw].astype(bool), its[:nw], peak[:nw]

    def debug_newton_direction(self, dt, cell_state=None, regimes=None, bhp=None, q=None):
        nw = self._l.rsf_num_wells(self._h)
        npf = self._l.rsf_num_perforations(self._h)
        dx = np.empty(3 * self.n + nw + npf)
        its = ctypes.c_int(0)
        keep = []

        def opt(a, dtype, ctype):
            if a is None:
                return None
            arr = np.ascontiguousarray(np.ravel(a), dtype=dtype)
            keep.append(arr)
            return arr.ctypes.data_as(ctype)

        self._chk(self._l.rsf_debug_newton_direction(self._h, dt, opt(cell_state, np.float64, _DP),
                                                    opt(regimes, np.int32, _IP), opt(bhp, np.float64, _DP),
                                                    opt(q, np.float64, _DP), _dp(dx), ctypes.byref(its)),
                  "debug_newton_direction")
        return dx, its.value

    def profile_kernels(self, reps=20):
        """Mean per-launch duration and algorithmic bytes of the hot kernels at the current system."""
        t = KernelTiming()
        self._chk(self._l.rsf_profile_kernels(self._h, int(reps), ctypes.byref(t)), "profile_kernels")
        return {f: getattr(t, f) for f, _ in KernelTiming._fields_}

    # ---- discrete adjoint ----
    def adjoint_record(self, checkpoint_interval=10):
        """Record the following accepted substeps for the adjoint (0 stops recording)."""
        self._chk(self._l.rsf_adjoint_record(self._h, int(checkpoint_interval)), "adjoint_record")

    def adjoint_objective(self):
        """(J, n_substeps): surface oil produced by the producers over the recorded substeps."""
        v = ctypes.c_double(0.0)
        k = ctypes.c_int(0)
        self._chk(self._l.rsf_adjoint_objective(self._h, ctypes.byref(v), ctypes.byref(k)), "adjoint_objective")
        return v.value, k.value

    def adjoint_gradient(self):
        """(dJ/d ln k, dJ/d phi, report dict) from the backward pass."""
        gl = np.empty(self.n)
        gp = np.empty(self.n)
        rep = AdjointReport()
        self._chk(self._l.rsf_adjoint_gradient(self._h, _dp(gl), _dp(gp), self.n, ctypes.byref(rep)),
                  "adjoint_gradient")
        return gl, gp, {f: getattr(rep, f) for f, _ in AdjointReport._fields_}

    def set_ilu_sweep(self, mode=1, lanes_pressure=256, lanes_block=64):
        """Natural-order ILU solves: one launch per sweep (mode 1, tiles of grid lines marching
        along x) or one launch per hyperplane (mode 0); the same numbers either way."""
        self._chk(self._l.rsf_set_ilu_sweep(self._h, int(mode), int(lanes_pressure), int(lanes_block)),
                  "set_ilu_sweep")

    def set_kernel_probe(self, enable=True, every=1):
        """Time the roofline kernel's launches of the following steps with HIP events:
        every `every`-th iteration's launch (1: all of them)."""
        self._chk(self._l.rsf_set_kernel_probe(self._h, int(every) if enable else 0), "set_kernel_probe")

    def kernel_probe(self):
        t = KernelProbe()
        self._chk(self._l.rsf_get_kernel_probe(self._h, ctypes.byref(t)), "kernel_probe")
        out = {f: getattr(t, f) for f, _ in KernelProbe._fields_}
        out["kernel"] = out["kernel"].decode()
        return out

    def rate_history(self):
        """getRateHistory: one TimePointRates per accepted substep, as {field: array}."""
        k = self._l.rsf_rate_history_len(self._h)
        buf = (TimePoint * max(k, 1))()
        self._chk(self._l.rsf_get_rate_history(self._h, buf, k), "rate_history")
        return {f: np.array([getattr(buf[i], f) for i in range(k)]) for f, _ in TimePoint._fields_}

    def rate_history_since(self, start_index):
        """getRateHistorySince: the entries from min(start_index, count) on, as {field: array}."""
        k = self._l.rsf_get_rate_history_since(self._h, int(start_index), None, 0)
        if k < 0:
            raise ValueError("rate_history_since: invalid start index")
        buf = (TimePoint * max(k, 1))()
        self._l.rsf_get_rate_history_since(self._h, int(start_index), buf, k)
        return {f: np.array([getattr(buf[i], f) for i in range(k)]) for f, _ in TimePoint._fields_}

    def latest_rate_point(self):
        """getLatestRatePoint: the last entry as {field: value}, or None before the first."""
        k = self._l.rsf_rate_history_len(self._h)
        if k <= 0:
            return None
        h = self.rate_history_since(k - 1)
        return {f: float(v[0]) for f, v in h.items()}

    def set_sweep_config(self, config):
        """setSweepConfig: {geometry, swept_threshold, initial_oil_saturation,
        residual_oil_saturation}, or None to turn the sweep metrics off."""
        if config is None:
            self._chk(self._l.rsf_set_sweep_config(self._h, None, 0.0, 0.0, 0.0), "sweep config")
            return
        self._chk(self._l.rsf_set_sweep_config(self._h, str(config["geometry"]).encode(),
                                               float(config["swept_threshold"]),
                                               float(config["initial_oil_saturation"]),
                                               float(config["residual_oil_saturation"])), "sweep config")

    def load_state(self, time_days, grid_state, completion_bhp, rate_history):
        """loadState: grid_state {pressure, sat_water, sat_oil[, sat_gas][, rs]}, one BHP per
        completion, and a rate history as returned by rate_history()."""
        def arr(key):
            v = grid_state.get(key)
            return None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        p, sw, so, sg, rs = (arr(k) for k in ("pressure", "sat_water", "sat_oil", "sat_gas", "rs"))
        for name, v in (("pressure", p), ("sat_water", sw), ("sat_oil", so), ("sat_gas", sg), ("rs", rs)):
            if v is not None and len(v) != self.n:
                raise ValueError(f"Mismatch grid size. Expected {self.n}, got {name} len: {len(v)}")
        b = np.ascontiguousarray(completion_bhp, dtype=np.float64)
        k = len(rate_history["time"]) if rate_history else 0
        buf = (TimePoint * max(k, 1))()
        for i in range(k):
            for f, _ in TimePoint._fields_:
                setattr(buf[i], f, float(rate_history[f][i]) if f in rate_history else float("nan"))
        self._chk(self._l.rsf_load_state(self._h, float(time_days), _dp(p), _dp(sw), _dp(so),
                                         None if sg is None else _dp(sg), None if rs is None else _dp(rs), self.n,
                                         _dp(b) if len(b) else None, len(b), buf, k), "load_state")

    def set_capture_dir(self, directory, mode=1):
        """Write the linear systems the GPU solves as fim-capture-v2 files into
        `directory`: mode 1 every solve, 2 only those that do not converge, 0 off."""
        d = None if directory is None else os.fsencode(directory)
        self._chk(self._l.rsf_set_capture_dir(self._h, d, int(mode)), "set_capture_dir")

    def solve_captured_system(self, jacobian, rhs):
        """Solve a captured system (scipy CSR in the reference's unknown order: three
        interleaved unknowns per cell, then the well tail) through the GPU linear stack
        with this handle's solver options. Returns (x, report dict)."""
        import scipy.sparse as sp

        J = sp.csr_matrix(jacobian)
        J.sort_indices()
        rows = J.shape[0]
        ptr = np.ascontiguousarray(J.indptr, dtype=np.int32)
        col = np.ascontiguousarray(J.indices, dtype=np.int32)
        val = np.ascontiguousarray(J.data, dtype=np.float64)
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        if b.shape != (rows,) or J.shape[1] != rows:
            raise ValueError("captured system must be square with one rhs value per row")
        x = np.empty(rows)
        rep = LinearReport()
        self._chk(self._l.rsf_solve_captured_system(self._h, rows, ptr.ctypes.data_as(_IP), col.ctypes.data_as(_IP),
                                                    _dp(val), _dp(b), _dp(x), ctypes.byref(rep)),
                  "solve_captured_system")
        return x, {f: getattr(rep, f) for f, _ in LinearReport._fields_}

    def debug_last_iterate(self):
        """Converged Newton iterate of the last accepted substep: (cell_state[N,3], regimes, bhp, q)."""
        nw, npf = self._l.rsf_num_wells(self._h), self._l.rsf_num_perforations(self._h)
        cs = np.empty(3 * self.n)
        rg = np.empty(self.n, dtype=np.int32)
        b = np.empty(max(nw, 1))
        q = np.empty(max(npf, 1))
        self._chk(self._l.rsf_debug_last_iterate(self._h, _dp(cs), rg.ctypes.data_as(_IP), _dp(b), nw, _dp(q), npf),
                  "last iterate")
        return cs.reshape(-1, 3), rg, b[:nw], q[:npf]

    def debug_initial_state(self):
        cs = np.empty(3 * self.n)
        rg = np.empty(self.n, dtype=np.int32)
        self._chk(self._l.rsf_debug_initial_state(self._h, _dp(cs), rg.ctypes.data_as(_IP)), "initial state")
        return cs.reshape(-1, 3), rg
