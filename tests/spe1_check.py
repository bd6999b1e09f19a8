"""The reference's SPE1 acceptance criteria (rs/tests/spe1_acceptance.rs:98-312), applied to
a rate history: checkpoint errors against the published series and the material-balance
drifts. Shared by the oracle's and the GPU's SPE1 tests."""
import json
import os

import numpy as np

SERIES = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                     "spe1_reference_series.json")))


def interpolate(series, t):
    s = np.asarray(series)
    if t <= s[0, 0]:
        return float(s[0, 1])
    for (t0, v0), (t1, v1) in zip(s[:-1], s[1:]):
        if t <= t1:
            return float(v0 + (v1 - v0) * (t - t0) / (t1 - t0))
    return float(s[-1, 1])


def checkpoint_report(history, t, oil_in_place, gas_in_place):
    """Errors at checkpoint t from the last entry of `history` (the rate history at t)."""
    k = len(history["time"]) - 1
    assert abs(history["time"][k] - t) < 1e-6
    tol = SERIES["tolerance"]
    ref_p, ref_q, ref_gor = (interpolate(SERIES[s], t) for s in ("pressure_bar", "oil_rate_sc_day", "gor"))
    out = dict(t=t,
               pressure=abs(history["avg_reservoir_pressure"][k] - ref_p) / ref_p,
               oil_rate=abs(history["total_production_oil"][k] - ref_q) / ref_q,
               gor=abs(history["producing_gor"][k] - ref_gor) / ref_gor,
               plateau=abs(history["total_production_oil"][k] - SERIES["target_oil_rate_sc_day"])
               / SERIES["target_oil_rate_sc_day"],
               oil_mb=abs(history["material_balance_error_oil_m3"][k]) / oil_in_place,
               gas_mb=abs(history["material_balance_error_gas_m3"][k])
               / (gas_in_place + SERIES["target_gas_injection_sc_day"] * t))
    out["ok"] = (out["pressure"] <= tol["pressure"] and out["oil_rate"] <= tol["oil_rate"] and out["gor"] <= tol["gor"]
                 and out["oil_mb"] <= tol["oil_material_balance"] and out["gas_mb"] <= tol["gas_material_balance"]
                 and (t > 730.0 or out["plateau"] <= tol["plateau"]))
    return out


def run_and_check(sim, deck, oil_in_place, gas_in_place, log=print):
    """Steps `sim` through the deck's schedule; at every yearly checkpoint applies the
    reference's criteria. Returns the list of checkpoint reports."""
    reports, t = [], 0.0
    checkpoints = list(deck.info["checkpoints"])
    for dt in deck.schedule:
        rc = sim.step(dt)
        assert rc == 0, sim.last_warning()
        assert sim.last_warning() == ""  # the reference asserts a clean run (spe1_acceptance.rs:177-183)
        t += dt
        if checkpoints and abs(t - checkpoints[0]) < 1e-6:
            r = checkpoint_report(sim.rate_history(), checkpoints.pop(0), oil_in_place, gas_in_place)
            log(f"[spe1] t={r['t']:7.1f} pressure_err={100 * r['pressure']:6.3f}% oil_rate_err="
                f"{100 * r['oil_rate']:6.3f}% gor_err={100 * r['gor']:6.3f}% oil_mb={100 * r['oil_mb']:.3f}% "
                f"gas_mb={100 * r['gas_mb']:.3f}%")
            reports.append(r)
    return reports
