"""rsf_run_fim_timestep: run_fim_timestep(sim, previous_state, initial_iterate, dt, options)
(rs/fim/newton.rs:970-976) with every argument per call, one Newton attempt that commits
nothing, and the FimStepReport fields (:921-944)."""
import dataclasses
import sys

import numpy as np
import pytest

from ressim_amd import ReservoirSimulator, decks

pytestmark = pytest.mark.gpu

TIGHT = dict(newton_residual_tol=1e-10, linear_rel_tol=1e-10, linear_max_iterations=200, linear_restart=60,
             max_newton_iterations=30, min_newton_iteration_index=1)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bhp_waterflood():
    """Two-phase 7x6x3 deck, BHP wells, gravity off: nothing but the previous state's
    accumulation depends on the stored state, so an attempt from a given previous state can
    be compared with the committed run that reached it."""
    d = decks.small_waterflood(7, 6, 3, n_steps=2, rate=False)
    calls = [("set_gravity_enabled", (False,)) if c[0] == "set_gravity_enabled" else c for c in d.calls]
    return dataclasses.replace(d, calls=calls)


def test_stored_state_attempt_equals_run_timestep_and_commits_nothing():
    d = decks.small_black_oil(7, 7, 3, n_steps=1, five_wells=True)
    a = d.build(ReservoirSimulator)
    b = d.build(ReservoirSimulator)
    a.set_solver_options(**TIGHT)
    b.set_solver_options(**TIGHT)
    dt = d.schedule[0]
    p0 = a.get_pressures()
    rc, rep, (cs, rg, bhp, q) = a.run_fim_timestep(dt)
    assert rc == 0 and rep["converged"] == 1
    assert np.array_equal(a.get_pressures(), p0), "run_fim_timestep must not commit"
    assert len(a.accepted_dts()) == 0
    rb, repb = b.run_timestep(dt)
    assert rb == 0
    cb, rgb, bb, qb = b.debug_last_iterate()
    assert np.array_equal(cs.ravel(), np.ravel(cb)) and np.array_equal(rg, rgb)
    assert np.array_equal(bhp, bb) and np.array_equal(q, qb)
    assert rep["newton_iterations"] == repb.newton_iterations
    assert rep["linear_iteration_count"] == repb.linear_iterations
    # FimStepReport bookkeeping of a converged attempt
    assert rep["applied_update_count"] == rep["newton_iterations"] == rep["linear_solve_count"]
    assert rep["has_last_linear_report"] == 1 and rep["last_linear_converged"] == 1
    assert rep["final_residual_inf_norm"] <= 1e-10 and rep["final_material_balance_inf_norm"] <= 1e-7
    assert rep["failure_class"] == -1 and rep["retry_factor"] == 0.5
    assert np.isnan(rep["final_update_inf_norm"])
    assert rep["total_time_ms"] >= rep["assembly_ms"] + rep["linear_solve_time_ms"] > 0
    log(f"[fim] stored-state attempt: newton {rep['newton_iterations']} linear {rep['linear_iteration_count']}")


def test_previous_state_sets_the_accumulation():
    d = bhp_waterflood()
    a = d.build(ReservoirSimulator)
    a.set_solver_options(**TIGHT)
    dt1, dt2 = 1.0, 2.0
    assert a.run_timestep(dt1)[0] == 0
    cs1, rg1, b1, q1 = a.debug_last_iterate()
    assert a.run_timestep(dt2)[0] == 0
    cs2, rg2, b2, q2 = a.debug_last_iterate()
    b = d.build(ReservoirSimulator)  # stored state: the initial one
    b.set_solver_options(**TIGHT)
    p_stored = b.get_pressures()
    rc, rep, (cs, rg, bhp, q) = b.run_fim_timestep(dt2, previous=(cs1, rg1), initial=(cs1, rg1, b1, q1))
    assert rc == 0
    assert np.array_equal(b.get_pressures(), p_stored)
    cs2 = np.reshape(cs2, (-1, 3))
    dp = np.max(np.abs(cs[:, 0] - cs2[:, 0]) / np.abs(cs2[:, 0]))
    ds = np.max(np.abs(cs[:, 1] - cs2[:, 1]))
    dq = np.max(np.abs(q - q2) / np.maximum(np.abs(q2), 1e-12))
    log(f"[fim] from the previous state of substep 2: pressure {dp:.2e} sw {ds:.2e} rates {dq:.2e}")
    assert dp <= 1e-8 and ds <= 1e-7 and dq <= 1e-7
    # the same call with the stored state as previous state lands elsewhere (a different step)
    rc0, _, (cs0, _, _, _) = b.run_fim_timestep(dt2)
    assert rc0 == 0 and np.max(np.abs(cs0[:, 0] - cs2[:, 0]) / np.abs(cs2[:, 0])) > 1e-6


def test_options_per_call_leave_the_handle_options():
    d = decks.small_black_oil(12, 10, 6, n_steps=1)
    mixed = dict(TIGHT, precision=1, ilu_ordering=1)
    a = d.build(ReservoirSimulator)
    a.set_solver_options(**TIGHT)  # FP64, natural order
    b = d.build(ReservoirSimulator)
    b.set_solver_options(**mixed)
    c = d.build(ReservoirSimulator)
    c.set_solver_options(**TIGHT)
    dt = 0.5
    _, ra, (csa, _, _, _) = a.run_fim_timestep(dt, options=mixed)
    _, rb, (csb, _, _, _) = b.run_fim_timestep(dt)
    assert np.array_equal(csa, csb) and ra["linear_iteration_count"] == rb["linear_iteration_count"]
    _, ra2, (csa2, _, _, _) = a.run_fim_timestep(dt)  # back on the handle's FP64 options
    _, rc_, (csc, _, _, _) = c.run_fim_timestep(dt)
    assert np.array_equal(csa2, csc) and ra2["linear_iteration_count"] == rc_["linear_iteration_count"]
    log(f"[fim] per-call mixed: linear {ra['linear_iteration_count']}; handle FP64: {ra2['linear_iteration_count']}")


def test_failed_attempt_reports_its_class():
    d = decks.small_black_oil(7, 7, 3, n_steps=1, five_wells=True)
    a = d.build(ReservoirSimulator)
    a.set_solver_options(**TIGHT)
    rc, rep, _ = a.run_fim_timestep(5.0, options=dict(TIGHT, max_newton_iterations=1))
    assert rc == 2 and rep["converged"] == 0
    assert rep["failure_class"] in (0, 1, 2) and 0 <= rep["failure_dominant_family"] <= 4
    # the dominant row lies in the dominant family's rows (3 per cell, wells, completions)
    bhp, q = a.get_well_state()
    n, nw, row, fam = d.n_cells, len(bhp), rep["failure_dominant_row"], rep["failure_dominant_family"]
    assert 0 <= row < 3 * n + nw + len(q)
    assert (row % 3 if row < 3 * n else 3 if row < 3 * n + nw else 4) == fam, (row, fam)
    assert rep["retry_factor"] == (0.25 if rep["failure_class"] == 0 else 0.5)
    assert len(a.accepted_dts()) == 0


def test_invalid_arguments():
    d = decks.small_black_oil(6, 5, 3, n_steps=1)
    a = d.build(ReservoirSimulator)
    cs, rg = a.debug_initial_state()
    with pytest.raises(ValueError):
        a.run_fim_timestep(-1.0)
    with pytest.raises(ValueError):
        a.run_fim_timestep(1.0, previous=(cs, None))
    with pytest.raises(ValueError):
        a.run_fim_timestep(1.0, options=dict(linear_restart=1))
