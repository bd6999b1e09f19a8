"""Pins of the oracle's substep controller, Newton relaxation and well relaxation.

The known answers are the reference's own unit tests: rs/fim/timestep.rs:2100-2165 (OPM PID
and iteration-count growth) and rs/fim/newton/tests.rs:1534-1652 (oscillation detection and
the relaxation factor). The well relaxation (rs/fim/state.rs:494-551), which the reference
tests only through whole runs, is pinned by its fixed points, its caps and its sign clamp,
and the rate-controlled well's consistent BHP by the target it must reproduce.
"""
import math

import numpy as np
import pytest

import oracle
from oracle import OracleSimulator

TINY = [1e-6, 1e-6, 1e-6]
FLOAT_MAX = np.finfo(np.float64).max


def test_pid_dt_known_answers():
    """timestep.rs:2100-2121."""
    assert abs(oracle.pid_dt(1.0, [0.1, 0.1, 0.2]) - 0.5) < 1e-12
    assert oracle.pid_dt(1.0, [0.0, 0.05, 0.05]) == FLOAT_MAX
    expected = (0.05 / 0.05) ** 0.075 * (0.1 / 0.05) ** 0.175 * (0.05 * 0.05 / (0.05 * 0.05)) ** 0.01
    dt = oracle.pid_dt(1.0, [0.05, 0.05, 0.05])
    assert abs(dt - expected) < 1e-9 and dt > 1.0


def test_iteration_count_dt_known_answers():
    """timestep.rs:2124-2133: 3.4 at 2 iterations, 1 at the target 8, 0.5 at 16."""
    assert abs(oracle.iteration_count_dt(1.0, 2) - 3.4) < 1e-12
    assert abs(oracle.iteration_count_dt(1.0, 8) - 1.0) < 1e-12
    assert abs(oracle.iteration_count_dt(1.0, 16) - 0.5) < 1e-12


def test_growth_decision_known_answers():
    """timestep.rs:2135-2160: the 3x ceiling wins on a clean history, the iteration count
    binds at 16 iterations, and a retry caps growth at 2x."""
    assert oracle.growth_decision(1.0, 0, TINY, 0) == (pytest.approx(3.0, abs=1e-9), "opm-max-growth")
    assert oracle.growth_decision(1.0, 16, TINY, 0) == (pytest.approx(0.5, abs=1e-9), "opm-iter")
    assert oracle.growth_decision(1.0, 0, TINY, 1) == (pytest.approx(2.0, abs=1e-9), "opm-restart-growth")
    # the PID half binds when the last relative change exceeds the 0.1 tolerance
    f, lim = oracle.growth_decision(2.0, 8, [0.1, 0.1, 0.4], 0)
    assert lim == "opm-pid" and f == pytest.approx(0.25, abs=1e-12)


def _norms(**kw):
    d = dict(water=1.0, oil_component=1.0, gas_component=1.0, well_constraint=1.0, perforation_flow=1.0)
    d.update(kw)
    return d


def test_detect_oscillation_known_answers():
    """newton/tests.rs:1534-1639."""
    assert oracle.detect_oscillation(_norms(), _norms(water=2.0), _norms(water=1.01)) == 1
    assert oracle.detect_oscillation(_norms(perforation_flow=2.137e-5), _norms(perforation_flow=3.419e-5),
                                     _norms(perforation_flow=2.137e-5)) == 1
    missing = _norms(well_constraint=math.inf, perforation_flow=math.inf)
    assert oracle.detect_oscillation(missing, missing, missing) == 0
    assert oracle.detect_oscillation(_norms(), _norms(water=2.0), _norms(water=4.0)) == 0
    assert oracle.detect_oscillation(_norms(), _norms(), _norms()) == 0


def test_next_relaxation_factor_known_answers():
    """newton/tests.rs:1641-1652: floors at 0.5, holds without oscillation, one step is 0.1."""
    relax = 1.0
    for _ in range(20):
        relax = oracle.next_relaxation_factor(relax, 1)
    assert abs(relax - 0.5) < 1e-12
    assert abs(oracle.next_relaxation_factor(0.7, 0) - 0.7) < 1e-12
    assert abs(oracle.next_relaxation_factor(1.0, 1) - 0.9) < 1e-12


def test_relative_change_closed_form():
    """timestep.rs:436-466: sum of squared changes of p, Sw, Sg (Sg = 0 undersaturated) and
    the implied So over the sum of squares of the new values."""
    a = np.array([[100.0, 0.2, 0.1], [200.0, 0.3, 50.0]])
    b = np.array([[110.0, 0.25, 0.05], [190.0, 0.35, 60.0]])
    ra = rb = np.array([0, 1], dtype=np.int32)  # saturated (hc = Sg), undersaturated (hc = Rs)
    num = (10.0 ** 2 + 0.05 ** 2 + 0.05 ** 2 + 0.0 ** 2) + (10.0 ** 2 + 0.05 ** 2 + 0.0 + 0.05 ** 2)
    den = (110.0 ** 2 + 0.25 ** 2 + 0.05 ** 2 + 0.7 ** 2) + (190.0 ** 2 + 0.35 ** 2 + 0.0 + 0.65 ** 2)
    assert oracle.relative_change(a, ra, b, rb) == pytest.approx(num / den, rel=1e-12)
    assert oracle.relative_change(a, ra, a, ra) == 0.0


# --------------------------------------------------------------- well relaxation
def _pair(rate_controlled=False):
    s = OracleSimulator(2, 1, 1, 0.2)
    if rate_controlled:
        s.set_injected_fluid("water")
        s.add_well(0, 0, 0, 400.0, 0.1, 0.0, True)
        s.add_well(1, 0, 0, 100.0, 0.1, 0.0, False)
        s.set_well_control_modes("rate", "rate")
        s.set_target_well_rates(30.0, 20.0)
        s.set_well_bhp_limits(50.0, 500.0)
    else:
        s.add_well(0, 0, 0, 500.0, 0.1, 0.0, True)
        s.add_well(1, 0, 0, 100.0, 0.1, 0.0, False)
    cells, rg = s.initial_state()
    return s, cells, rg


def _connection(s, cells, rg, bhp, q, w):
    """Connection rate of well w's single completion at bhp, from the local row q - conn."""
    r, _ = s.well_local_system(cells, rg, bhp, q, w)
    return q[w] - r[1]


def test_relaxation_fixed_point_and_bhp_caps():
    """A BHP-controlled well at its target with q at its connection rate stays put; a BHP
    100 bar off moves 90 % of the way (to 10 bar off); 1,000 bar off is capped at 25 bar."""
    s, cells, rg = _pair()
    targets = np.array([500.0, 100.0])
    q0 = np.array([_connection(s, cells, rg, targets, np.zeros(2), w) for w in range(2)])
    b, q = s.relax_wells(cells, rg, targets, q0)
    np.testing.assert_array_equal(b, targets)
    np.testing.assert_allclose(q, q0, rtol=1e-14)
    b, _ = s.relax_wells(cells, rg, targets + 100.0, q0)
    np.testing.assert_allclose(b, targets + 10.0, rtol=1e-14)
    b, _ = s.relax_wells(cells, rg, targets - 1000.0, q0)
    np.testing.assert_allclose(b, targets - 25.0, rtol=1e-14)


def test_relaxation_rate_blend_cap_and_sign_clamp():
    """q moves 75 % toward the connection rate within max(10 % of it, 250 m3/d), and keeps
    the completion's sign (injectors q <= 0, producers q >= 0)."""
    s, cells, rg = _pair()
    targets = np.array([500.0, 100.0])
    cq = np.array([_connection(s, cells, rg, targets, np.zeros(2), w) for w in range(2)])
    assert cq[0] < 0.0 < cq[1]
    tr = np.maximum(0.1 * np.abs(cq), 250.0)
    _, q = s.relax_wells(cells, rg, targets, cq + 40.0)  # within the cap: 25 % of the gap left
    np.testing.assert_allclose(q, cq + 10.0, rtol=1e-12)
    _, q = s.relax_wells(cells, rg, targets, cq + 100.0 * tr)  # far off: capped at tr
    expect = cq + tr
    expect[0] = min(expect[0], 0.0)
    np.testing.assert_allclose(q, expect, rtol=1e-12, atol=1e-12)
    _, q = s.relax_wells(cells, rg, targets, cq - 100.0 * tr)
    expect = cq - tr
    expect[1] = max(expect[1], 0.0)
    np.testing.assert_allclose(q, expect, rtol=1e-12, atol=1e-12)


def test_relaxation_rate_controlled_bhp_reproduces_the_target():
    """For a rate-controlled well the consistent BHP is the bisection root of its total
    rate on the target (rs/fim/wells.rs:431-504): starting there, the BHP stays within the
    bisection's resolution and the completion rate meets the target, or the BHP sits at its
    limit when the target cannot be met."""
    s, cells, rg = _pair(rate_controlled=True)
    b0, q0 = s.initial_well_state()
    b, q = s.relax_wells(cells, rg, b0, q0)
    np.testing.assert_allclose(b, b0, rtol=1e-10)
    for w, target, limit in ((0, 30.0, 500.0), (1, 20.0, 50.0)):
        conn = abs(_connection(s, cells, rg, b, q, w))
        assert conn == pytest.approx(target, rel=1e-8) or b[w] == pytest.approx(limit)
