"""The step-statistics history (FimStepStats, rs/reporting.rs:141-243;
getFimStepStatsHistory, rs/frontend.rs:597-600) against the oracle: with Newton capped
below what the 1e-10 tolerance needs, the first trials of a step are rejected and
retried, and both sides record the same rungs, dominant residual families and rows, dt
bookkeeping and growth limiters.

The retry class depends on the linear backend (rs/fim/newton.rs:439-481). Both sides solve
these systems (278 rows with the well tail) directly, as the reference does at or below 512
rows (rs/fim/linear/mod.rs:313-335), so every rung carries the oracle's class, and a direct
solve counts one linear iteration on both sides."""
import sys

import numpy as np
import pytest

from ressim_amd import ReservoirSimulator, decks

pytestmark = pytest.mark.gpu

CAP = 200  # FGMRES iteration cap of the runs below


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("max_newton", [3, 4])
def test_step_stats_match_oracle(max_newton, oracle_mod):
    d = decks.small_black_oil(6, 5, 3, n_steps=2)
    o = d.build(oracle_mod.OracleSimulator)
    o.set_newton_options(max_iterations=max_newton, min_iteration_index=1, residual_tol=1e-10, linear_rel_tol=1e-10,
                         linear_max_iterations=200)
    g = d.build(ReservoirSimulator)
    g.set_solver_options(newton_residual_tol=1e-10, linear_rel_tol=1e-10, linear_max_iterations=CAP,
                         linear_restart=60, max_newton_iterations=max_newton, min_newton_iteration_index=1,
                         precision=0, ilu_ordering=0)
    assert g.last_step_stats() is None
    for dt in d.schedule:
        assert o.step(dt) == 0, o.last_warning()
        assert g.step(dt) == 0, g.last_warning()
    ho, hg = o.step_stats_history(), g.step_stats_history()
    assert len(ho) == len(hg) == len(d.schedule)
    for k, (a, b) in enumerate(zip(ho, hg)):
        log(f"[step_stats] max_newton={max_newton} step {k}: substeps {b['accepted_substeps']} retries "
            f"{b['linear_bad_retries']}/{b['nonlinear_bad_retries']}/{b['mixed_retries']} last class "
            f"{b['last_retry_class']} family {b['last_retry_dominant_family']} limiter {b['growth_limiter']}")
        for f in ("accepted_substeps", "mixed_retries", "linear_bad_retries", "nonlinear_bad_retries",
                  "growth_limiter", "last_retry_class", "last_retry_dominant_family", "last_retry_dominant_row"):
            assert b[f] == a[f], (k, f, a[f], b[f])
        for f in ("time_days", "target_dt_days", "advanced_dt_days", "min_accepted_dt_days", "max_accepted_dt_days",
                  "last_accepted_dt_days", "last_growth_factor"):
            assert b[f] == pytest.approx(a[f], rel=1e-9), (k, f)
        for kind in ("accepted_rungs", "retry_rungs"):
            assert len(b[kind]) == len(a[kind]), (k, kind)
            for ra, rb in zip(a[kind], b[kind]):
                for f in ("substep", "newton_iterations", "linear_solve_count", "linear_iteration_count",
                          "linear_iterations", "last_linear_converged", "dominant_family", "dominant_row",
                          "retry_class"):
                    assert rb.get(f) == ra.get(f), (k, kind, f, ra, rb)
                assert rb["dt_days"] == pytest.approx(ra["dt_days"], rel=1e-9)
        if b["retry_rungs"]:
            assert b["last_retry_class"] == b["retry_rungs"][-1]["retry_class"]
        assert b["solver_ms"] >= b["accepted_solver_ms"] + b["retry_solver_ms"] - 1e-9
        assert min(b["assembly_ms"], b["linear_solve_ms"], b["retry_solver_ms"]) >= 0.0
    assert sum(h["nonlinear_bad_retries"] for h in hg) > 0
    assert g.last_step_stats() == hg[-1] == g.get_last_fim_step_stats()
    assert g.get_fim_step_stats_history() == hg
    np.testing.assert_allclose(g.accepted_dts(), o.accepted_dts(), rtol=1e-9)


@pytest.mark.parametrize("target", [2.0, 1e-10])
def test_controller_exits_match_oracle(target, oracle_mod):
    """The retry-budget and collapse exits (rs/fim/timestep.rs:1672-1767) with a tolerance
    no iterate meets: the same warning, rungs and dts as the oracle."""
    d = decks.small_black_oil(6, 5, 3, n_steps=1)
    o = d.build(oracle_mod.OracleSimulator)
    o.set_newton_options(max_iterations=1, min_iteration_index=1, residual_tol=1e-30)
    g = d.build(ReservoirSimulator)
    g.set_solver_options(newton_residual_tol=1e-30, max_newton_iterations=1, min_newton_iteration_index=1)
    assert o.step(target) == 2 and g.step(target) == 2
    log(f"[step_stats] exit at target {target}: {g.last_warning()}")
    assert g.last_warning() == o.last_warning()
    a, b = o.step_stats_history()[-1], g.step_stats_history()[-1]
    assert len(b["retry_rungs"]) == len(a["retry_rungs"]) and b["accepted_substeps"] == 0
    for ra, rb in zip(a["retry_rungs"], b["retry_rungs"]):
        assert rb["dt_days"] == pytest.approx(ra["dt_days"], rel=1e-12)
        assert rb["newton_iterations"] == ra["newton_iterations"] == 1
