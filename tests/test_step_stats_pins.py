"""The retry classification and the dominant residual family behind the step-statistics
history (FimStepStats, rs/reporting.rs:141-243), in both the library (host functions of
the C ABI, no GPU needed) and the oracle.

Known answers: the reference's unit tests rs/fim/newton/tests.rs:858-1300 (clean CPR
failure, clean direct backend and converged fallback: nonlinear-bad; weak CPR with
reductions 0.6 / 0.8: mixed; unconverged fallback: linear-bad; family peaks 1.0 / 0.5 /
0.25: water) and the rule's boundaries (rs/fim/newton.rs:50-51: average <= 0.25 and last
<= 0.5 inclusive; no coarse application is weak; no linear solve is mixed). The dominant
row: the reference's residual_family_diagnostics test (tests.rs:795-830) and numpy's
first-occurrence argmax on residuals with ties."""
import math

import numpy as np
import pytest

import oracle
from ressim_amd import simulator

IMPLS = {"library": (simulator.debug_classify_retry, simulator.debug_dominant_family),
         "oracle": (oracle.classify_retry, oracle.dominant_family)}

CASES = [
    # (have, converged, direct, apps, avg, last) -> class
    ((True, True, False, 4, 1e-12, 1e-12), "nonlinear-bad"),   # tests.rs:858-915
    ((True, True, True, 0, 1.0, 1.0), "nonlinear-bad"),        # tests.rs:917-965 direct backend
    ((True, True, False, 4, 0.6, 0.8), "mixed"),                # tests.rs:967-1024
    ((True, False, True, 0, 1.0, 1.0), "linear-bad"),          # tests.rs:1251-1300
    ((True, False, False, 4, 1e-12, 1e-12), "linear-bad"),     # unconverged CPR, strong or not
    ((False, False, False, 0, 1.0, 1.0), "mixed"),             # no linear report (newton.rs:479-481)
    ((True, True, False, 0, 1e-12, 1e-12), "mixed"),           # no coarse application
    ((True, True, False, 3, 0.25, 0.5), "nonlinear-bad"),      # both bounds inclusive
    ((True, True, False, 3, 0.25 + 1e-12, 0.5), "mixed"),
    ((True, True, False, 3, 0.25, 0.5 + 1e-12), "mixed"),
    ((True, True, False, 3, math.nan, 0.1), "mixed"),          # NaN never passes a bound
]


@pytest.mark.parametrize("impl", list(IMPLS))
@pytest.mark.parametrize("args,expected", CASES)
def test_retry_class_known_answers(impl, args, expected):
    classify, _ = IMPLS[impl]
    assert classify(*args) == expected


@pytest.mark.parametrize("impl", list(IMPLS))
def test_dominant_family_known_answers(impl):
    _, fam = IMPLS[impl]
    assert fam([1.0, 0.5, 0.25, 0.0, 0.0], False) == "water"  # tests.rs:858-915
    assert fam([0.1, 0.5, 0.25, 9.0, 0.0], False) == "oil"    # well families absent without wells
    assert fam([0.1, 0.5, 0.25, 9.0, 0.0], True) == "well"
    assert fam([0.1, 0.5, 0.25, 0.3, 0.7], True) == "perf"
    assert fam([0.5, 0.5, 0.5, 0.5, 0.5], True) == "water"   # ties keep the earlier family
    assert fam([0.2, 0.1, 0.3, 0.3, 0.1], True) == "gas"
    assert fam([0.2, math.nan, 0.1, 0.0, 0.0], True) == "water"  # NaN never wins
    assert fam([math.nan, 0.1, 0.0, 0.0, 0.0], False) == "water"  # nor is it replaced (x > NaN is false)


def family_of_row(row, n_cells, n_wells):
    """The residual family of a row in the reference's unknown numbering (3 per cell, then
    wells, then completions)."""
    if row < 3 * n_cells:
        return ("water", "oil", "gas")[row % 3]
    return "well" if row < 3 * n_cells + n_wells else "perf"


def test_family_peaks_known_answer():
    """rs/fim/newton/tests.rs:795-830 residual_family_diagnostics_reports_global_peak_family:
    two cells, two wells and one completion; the peaks are water 0.5 at cell 0, oil 1.2 at
    cell 0, gas 0.8 at cell 0, well 8.0 at well 1 and perforation 0.5 at completion 0, and
    the global peak is the second well's row (3 n + 1 = 7)."""
    residual = [5.0, 12.0, 8.0, 4.0, 9.0, 6.0, 3.0, 40.0, 1.0]
    scale = [10.0] * 6 + [10.0, 5.0] + [2.0]
    vals, rows, g = oracle.family_peaks(2, 2, 1, residual, scale)
    np.testing.assert_allclose(vals, [0.5, 1.2, 0.8, 8.0, 0.5], rtol=0, atol=1e-12)
    assert rows.tolist() == [0, 1, 2, 7, 8] and g == "well"


@pytest.mark.parametrize("seed", range(6))
def test_family_peaks_first_row_of_the_maximum(seed):
    """Each family's peak row is the first row holding its maximum (a later row replaces
    the peak only when strictly larger, rs/fim/newton/convergence.rs:325-345), checked
    against numpy's argmax (first occurrence) on residuals drawn from a few levels so that
    ties are common; all-zero families give their first row, and without wells the well
    families have no row and an infinite peak (never the global one)."""
    rng = np.random.default_rng(seed)
    n, W = int(rng.integers(1, 40)), int(rng.integers(0, 4))
    P = int(rng.integers(W, 3 * W + 1)) if W else 0
    m = 3 * n + W + P
    levels = np.array([0.0, 1.0, 2.0, 4.0])
    residual = rng.choice(levels, size=m) * rng.choice([-1.0, 1.0], size=m)
    scale = rng.choice([1.0, 2.0, 4.0], size=m)
    if seed == 0:
        residual[:] = 0.0
    vals, rows, g = oracle.family_peaks(n, W, P, residual, scale)
    a = np.abs(residual) / scale
    want_rows, want_vals = [], []
    for c in range(3):
        k = int(np.argmax(a[c:3 * n:3]))
        want_rows.append(3 * k + c)
        want_vals.append(a[3 * k + c])
    for lo, cnt in ((3 * n, W), (3 * n + W, P)):
        if cnt:
            k = int(np.argmax(a[lo:lo + cnt]))
            want_rows.append(lo + k)
            want_vals.append(a[lo + k])
        else:
            want_rows.append(-1)
            want_vals.append(np.inf)
    assert rows.tolist() == want_rows
    np.testing.assert_array_equal(vals, want_vals)
    fams = 5 if W else 3
    best = max(range(fams), key=lambda f: (want_vals[f], -f))
    assert g == ("water", "oil", "gas", "well", "perf")[best]
    assert family_of_row(want_rows[best], n, W) == g


def test_oracle_records_retries_and_rungs():
    """The oracle's record of a step whose first trials fail (Newton capped at 4 iterations
    at a 1e-10 tolerance): every rejected attempt is a rung with its class, the counts add
    up, the accepted dts are the accepted rungs' and the dt bookkeeping is consistent."""
    from ressim_amd import decks
    d = decks.small_black_oil(6, 5, 3, n_steps=2)
    o = d.build(oracle.OracleSimulator)
    o.set_newton_options(max_iterations=4, min_iteration_index=1, residual_tol=1e-10, linear_rel_tol=1e-10,
                         linear_max_iterations=200)
    for dt in d.schedule:
        assert o.step(dt) == 0
    hist = o.step_stats_history()
    assert len(hist) == len(d.schedule)
    acc_all = []
    for h, target in zip(hist, d.schedule):
        retries = h["linear_bad_retries"] + h["nonlinear_bad_retries"] + h["mixed_retries"]
        assert len(h["retry_rungs"]) == retries and len(h["accepted_rungs"]) == h["accepted_substeps"]
        dts = [r["dt_days"] for r in h["accepted_rungs"]]
        acc_all += dts
        assert h["advanced_dt_days"] == pytest.approx(target, rel=1e-12) == pytest.approx(sum(dts), rel=1e-12)
        assert h["min_accepted_dt_days"] == min(dts) and h["max_accepted_dt_days"] == max(dts)
        assert h["last_accepted_dt_days"] == dts[-1]
        if retries:
            assert h["last_retry_class"] == h["retry_rungs"][-1]["retry_class"]
            assert h["last_retry_dominant_family"] == h["retry_rungs"][-1]["dominant_family"]
            assert h["last_retry_dominant_row"] == h["retry_rungs"][-1]["dominant_row"]
            n_wells = len(o.get_well_state()[0])
            for r in h["retry_rungs"]:
                assert r["newton_iterations"] == 4  # rejected at the post-loop check
                assert family_of_row(r["dominant_row"], d.n_cells, n_wells) == r["dominant_family"]
        else:
            assert h["last_retry_dominant_row"] is None
    assert hist[0]["nonlinear_bad_retries"] > 0
    np.testing.assert_array_equal(np.array(acc_all), o.accepted_dts())


def _never_converging(d, factory_opts):
    o = d.build(oracle.OracleSimulator)
    o.set_newton_options(**factory_opts)
    return o


@pytest.mark.parametrize("target,rungs,text", [
    (2.0, 16, "FIM Newton step exceeded retry budget at t=0.000000 days after 16 retries"),
    (1e-10, 5, "FIM Newton step collapsed timestep at t=0.000000 days after 1 iterations"),
])
def test_oracle_controller_exits(target, rungs, text):
    """rs/fim/timestep.rs:1589-1767: with a tolerance no iterate meets, every attempt is
    rejected; the trial shrinks by 0.33 until it falls to 1e-12 (a 1e-10 day target: five
    attempts) or the 16th retry of the substep ends the step (:1000), with the reference's
    warning, and the record holds every rejected attempt."""
    from ressim_amd import decks
    d = decks.small_black_oil(6, 5, 3, n_steps=1)
    o = _never_converging(d, dict(max_iterations=1, min_iteration_index=1, residual_tol=1e-30))
    assert o.step(target) == 2
    assert o.last_warning() == text
    h = o.step_stats_history()[-1]
    assert len(h["retry_rungs"]) == rungs and h["accepted_substeps"] == 0 and h["advanced_dt_days"] == 0.0
    assert h["nonlinear_bad_retries"] == rungs  # each attempt's direct solve converged
    assert h["min_accepted_dt_days"] is None and h["growth_limiter"] is None
    dts = [r["dt_days"] for r in h["retry_rungs"]]
    np.testing.assert_allclose(np.array(dts[1:]) / np.array(dts[:-1]), 0.33, rtol=1e-12)
