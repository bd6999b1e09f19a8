"""Pins of the oracle's nested per-well Newton (rs/fim/wells_inner.rs).

The cases follow the reference's own tests of that file (wells_inner.rs:742-1062): the local
rows agree with the global assembly, the BHP chop's known answers, convergence from perturbed
states for BHP- and rate-controlled wells, the exhausted-budget report, and the read-only
convergence test before and after a solve. A central-difference check of the local Jacobian
on a multi-completion well adds what the reference leaves to construction.
"""
import numpy as np
import pytest

import oracle
from oracle import OracleSimulator

BHP_LOWER_LIMIT_BAR = 1.0 - 1.0e-5


def _pair(rate_controlled=False):
    """ReservoirSimulator::new(2, 1, 1, 0.2) with an injector in cell 0 and a producer in
    cell 1, as in wells_inner.rs:742-800."""
    s = OracleSimulator(2, 1, 1, 0.2)
    if rate_controlled:
        s.set_injected_fluid("water")
        s.add_well(0, 0, 0, 400.0, 0.1, 0.0, True)
        s.add_well(1, 0, 0, 100.0, 0.1, 0.0, False)
        s.set_well_control_modes("rate", "rate")
        s.set_target_well_surface_rates(30.0, 20.0)
        s.set_well_bhp_limits(50.0, 500.0)
    else:
        s.add_well(0, 0, 0, 500.0, 0.1, 0.0, True)
        s.add_well(1, 0, 0, 100.0, 0.1, 0.0, False)
    return s


def _iterate(s):
    cells, rg = s.initial_state()
    bhp, q = s.initial_well_state()
    return cells, rg, bhp, q


def _global_slice(s, cells, rg, bhp, q, well, perfs):
    """Rows and columns of the global assembly that belong to one well: (residual, dense J)."""
    R, (ptr, col, val) = s.assemble(1.0, cells, rg, bhp, q)
    n, nw = len(rg), len(bhp)
    idx = [3 * n + well] + [3 * n + nw + e for e in perfs]
    J = np.zeros((len(idx), len(idx)))
    for a, r in enumerate(idx):
        for k in range(ptr[r], ptr[r + 1]):
            if col[k] in idx:
                J[a, idx.index(col[k])] = val[k]
    return R[idx], J


@pytest.mark.parametrize("case", ["bhp_controlled", "rate_controlled", "away_from_convergence"])
def test_local_system_matches_global_assembly(case):
    """wells_inner.rs:659-805: local residual equals the global rows; Jacobian entries agree
    to 1e-12 (the global assembly drops |v| <= 1e-14); no perforation row couples to
    another perforation's q."""
    s = _pair(rate_controlled=case == "rate_controlled")
    cells, rg, bhp, q = _iterate(s)
    if case == "away_from_convergence":
        q = q + 500.0
    for w in range(2):
        r, J = s.well_local_system(cells, rg, bhp, q, w)
        assert r.shape == (2,) and J.shape == (2, 2)  # wells_inner.rs:807-828
        gr, gJ = _global_slice(s, cells, rg, bhp, q, w, [w])
        np.testing.assert_array_equal(r, gr)
        np.testing.assert_allclose(J, gJ, rtol=0, atol=1e-12)


def test_chop_bhp_update_known_answers():
    """wells_inner.rs:831-844."""
    assert abs(oracle.chop_bhp_update(200.0, 10.0, 1.0) - 210.0) < 1e-12
    assert abs(oracle.chop_bhp_update(200.0, 500.0, 1.0) - 400.0) < 1e-12
    assert abs(oracle.chop_bhp_update(200.0, -500.0, 1.0) - BHP_LOWER_LIMIT_BAR) < 1e-12
    assert abs(oracle.chop_bhp_update(200.0, 500.0, 0.1) - 220.0) < 1e-12
    assert abs(oracle.chop_bhp_update(0.5, -100.0, 1.0) - BHP_LOWER_LIMIT_BAR) < 1e-12


def test_bhp_controlled_well_converges_from_perturbed_state():
    """wells_inner.rs:862-909: q + 800 converges in at most 3 iterations, BHP stays on
    target, and the local residual is independently small afterwards."""
    s = _pair()
    cells, rg, bhp, q = _iterate(s)
    b, qq, conv, its, _ = s.solve_wells_locally(cells, rg, bhp, q + 800.0)
    assert conv.all() and (its <= 3).all()
    np.testing.assert_allclose(b, [500.0, 100.0], rtol=0, atol=1e-9)
    for w in range(2):
        r, _ = s.well_local_system(cells, rg, b, qq, w)
        assert np.all(np.abs(r) < 1e-6)


def test_rate_controlled_well_converges_to_slack_feasible_state():
    """wells_inner.rs:911-952: q * 0.3 converges to a (bhp, q) with either slack feasible."""
    s = _pair(rate_controlled=True)
    cells, rg, bhp, q = _iterate(s)
    b, qq, conv, its, _ = s.solve_wells_locally(cells, rg, bhp, q * 0.3)
    assert conv.all()
    eps = 1e-6  # rs/fim/wells.rs:1582
    for w, (inj, limit) in enumerate([(True, 500.0), (False, 50.0)]):
        phi, _ = s.well_local_system(cells, rg, b, qq, w)
        phi = phi[0]
        assert abs(phi) <= 1e-4
        a = ((limit - b[w]) if inj else (b[w] - limit)) / max(abs(limit), 1.0)
        # the scaled rate slack fixed by phi = sqrt(a^2 + r^2 + 2 eps^2) - a - r
        r = (2 * eps * eps - phi * phi - 2 * phi * a) / (2 * (phi + a))
        assert a > -1e-3 or r > -1e-3
    assert qq[0] <= 0.0 and qq[1] >= 0.0


def test_exhausted_budget_reports_not_converged():
    """wells_inner.rs:955-976: a zero budget reports not converged, 0 iterations, finite peak."""
    s = _pair()
    cells, rg, bhp, q = _iterate(s)
    b, qq, conv, its, peak = s.solve_wells_locally(cells, rg, bhp, q + 800.0, max_iterations=0)
    assert not conv[1] and its[1] == 0 and np.isfinite(peak[1])
    np.testing.assert_array_equal(qq, q + 800.0)


def test_convergence_check_before_and_after_solve():
    """wells_inner.rs:978-1062: every well covered; the read-only check (a zero-iteration
    solve) reads not converged before and converged after, one well at a time."""
    s = _pair()
    cells, rg, bhp, q = _iterate(s)
    q0 = q + 800.0
    _, _, before, _, _ = s.solve_wells_locally(cells, rg, bhp, q0, max_iterations=0)
    assert not before.any()
    b, qq, conv, _, _ = s.solve_wells_locally(cells, rg, bhp, q0)
    assert conv.all() and len(conv) == 2
    _, _, after, _, _ = s.solve_wells_locally(cells, rg, b, qq, max_iterations=0)
    assert after.all()
    # only well 0 converged: the aggregate stays false
    mixed = q0.copy()
    mixed[0] = qq[0]
    _, _, part, _, _ = s.solve_wells_locally(cells, rg, b, mixed, max_iterations=0)
    assert part[0] and not part[1]


def _multi_completion(surface):
    """3x3x3 box: a rate-controlled producer completed in three layers and a water injector."""
    s = OracleSimulator(3, 3, 3, 0.2)
    s.set_cell_dimensions(50.0, 50.0, 5.0)
    rng = np.random.default_rng(7)
    k = rng.uniform(50.0, 400.0, 27)
    s.set_permeability_field(k, k, 0.1 * k)
    s.set_initial_pressure(250.0)
    s.set_initial_saturation(0.25)
    s.set_injected_fluid("water")
    for z in range(3):
        s.add_well_with_id(2, 2, z, 150.0, 0.1, 0.5, False, "P1")
    s.add_well_with_id(0, 0, 1, 350.0, 0.1, 0.0, True, "I1")
    s.set_well_control_modes("rate", "rate")
    if surface:
        s.set_target_well_surface_rates(80.0, 60.0)
    else:
        s.set_target_well_rates(80.0, 60.0)
    s.set_well_bhp_limits(100.0, 400.0)
    return s


@pytest.mark.parametrize("surface", [False, True])
def test_local_jacobian_matches_central_differences(surface):
    """The local Jacobian of a three-completion rate-controlled well against central
    differences of the local residual in (bhp, q), at a state off the kinks (live
    connections, q away from 0)."""
    s = _multi_completion(surface)
    cells, rg, bhp, q = _iterate(s)
    q = q * 1.3 + np.where(q >= 0, 5.0, -5.0)
    for w in range(len(bhp)):
        r, J = s.well_local_system(cells, rg, bhp, q, w)
        dim = len(r)
        # unknown map: 0 -> bhp[w]; 1 + t -> the t-th completion of well w (P1: 0..2, I1: 3)
        order = [0, 1, 2] if w == 0 else [3]
        assert dim == 1 + len(order)
        fd = np.zeros_like(J)
        for j in range(dim):
            h = 1e-5 * max(abs(bhp[w] if j == 0 else q[order[j - 1]]), 1.0)
            bp, bm, qp, qm = bhp.copy(), bhp.copy(), q.copy(), q.copy()
            if j == 0:
                bp[w] += h
                bm[w] -= h
            else:
                qp[order[j - 1]] += h
                qm[order[j - 1]] -= h
            rp, _ = s.well_local_system(cells, rg, bp, qp, w)
            rm, _ = s.well_local_system(cells, rg, bm, qm, w)
            fd[:, j] = (rp - rm) / (2 * h)
        np.testing.assert_allclose(J, fd, rtol=1e-6, atol=1e-9)


def test_nested_solve_converges_multi_completion_wells():
    """A three-completion rate-controlled producer converges its local rows to the inner
    tolerance with the BHP chop active."""
    s = _multi_completion(True)
    cells, rg, bhp, q = _iterate(s)
    b, qq, conv, its, peak = s.solve_wells_locally(cells, rg, bhp * 1.2, q * 0.5)
    assert conv.all() and (peak <= 1e-4).all() and (its >= 1).all()
