"""The reference's Buckley-Leverett breakthrough check (rs/tests/buckley.rs:68-143,165-205):
the pore volumes injected when the producer's water cut first reaches 1 %, against the Welge
tangent of the Corey fractional-flow curve."""
import numpy as np


def fractional_flow(sw, c):
    d = 1.0 - c["s_wc"] - c["s_or"]
    se_w = np.clip((sw - c["s_wc"]) / d, 0.0, 1.0)
    se_o = np.clip((1.0 - sw - c["s_or"]) / d, 0.0, 1.0)
    lw, lo = se_w ** c["n_w"] / c["mu_w"], se_o ** c["n_o"] / c["mu_o"]
    lt = lw + lo
    return 0.0 if lt <= np.finfo(float).eps else float(np.clip(lw / lt, 0.0, 1.0))


def welge_breakthrough_pv(c):
    """1 / f'(S_shock) with S_shock the tangent point from (S_wc, 0), scanned in steps of 5e-4."""
    s0, s_max, best, shock = c["s_wc"], 1.0 - c["s_or"], 0.0, c["s_wc"]
    s = s0 + 5e-4
    while s <= s_max:
        slope = fractional_flow(s, c) / (s - s0)
        if slope > best and np.isfinite(slope):
            best, shock = slope, s
        s += 5e-4
    e = 1e-4
    d = (fractional_flow(np.clip(shock + e, s0, s_max), c) - fractional_flow(np.clip(shock - e, s0, s_max), c)) / (2 * e)
    return np.inf if d <= np.finfo(float).eps else 1.0 / d


def breakthrough_pv(sim, deck, total_pv):
    """Steps until the water cut reaches the case's 1 %; returns the PV injected then."""
    cum, prev_t = 0.0, 0.0
    for dt in deck.schedule:
        assert sim.step(dt) == 0, sim.last_warning()
        h = sim.rate_history()
        t = h["time"][-1]
        cum += max(h["total_injection"][-1], 0.0) * (t - prev_t)
        prev_t = t
        liq, oil = h["total_production_liquid"][-1], h["total_production_oil"][-1]
        if liq > 1e-9 and np.clip(max(liq - oil, 0.0) / liq, 0.0, 1.0) >= 0.01:
            return cum / total_pv
    raise AssertionError("no breakthrough")
