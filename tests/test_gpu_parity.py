"""GPU parity: the HIP path (through the C ABI) against the oracle on identical inputs.

Bars: assembly residual and Jacobian blocks to 1e-10 relative (FP64, different
summation order); Newton directions to 1e-8 relative against a direct solve of
the oracle's Jacobian; converged substeps with pressure within 1e-8 relative,
saturations within 1e-7 absolute and well rates within 1e-7 relative.
"""
import dataclasses
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ressim_amd import ReservoirSimulator, decks

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def csr_dense_blocks(ptr, col, val, n_cells):
    J = sp.csr_matrix((val, col, ptr), shape=(len(ptr) - 1, len(ptr) - 1))
    return J


def slot_neighbor(nx, ny, nz, i, s):
    z, r = divmod(i, nx * ny)
    y, x = divmod(r, nx)
    if s == 0:
        return i
    if s == 1:
        return i - 1 if x > 0 else -1
    if s == 2:
        return i + 1 if x + 1 < nx else -1
    if s == 3:
        return i - nx if y > 0 else -1
    if s == 4:
        return i + nx if y + 1 < ny else -1
    if s == 5:
        return i - nx * ny if z > 0 else -1
    return i + nx * ny if z + 1 < nz else -1


def blocks_to_csr(A, nx, ny, nz):
    n = nx * ny * nz
    rows, cols, vals = [], [], []
    for i in range(n):
        for s in range(7):
            j = slot_neighbor(nx, ny, nz, i, s)
            if j < 0:
                continue
            for c in range(3):
                for v in range(3):
                    rows.append(3 * i + c)
                    cols.append(3 * j + v)
                    vals.append(A[s, 3 * c + v, i])
    return sp.csr_matrix((vals, (rows, cols)), shape=(3 * n, 3 * n))


# (precision, ILU ordering): the FP64 reference-ordering path, the mixed-precision
# red-black path the benchmark runs and its FP64 leg (with nx even both store the
# preconditioner colour-compacted; the 7x7x3 and 1D decks cover the natural layout)
MODES = {"fp64_natural": (0, 0), "mixed_redblack": (1, 1), "fp64_redblack": (0, 1)}


def tight(sim, mode="fp64_natural", **kw):
    prec, order = MODES[mode]
    opts = dict(newton_residual_tol=1e-10, linear_rel_tol=1e-10, linear_max_iterations=200, linear_restart=60,
                max_newton_iterations=30, min_newton_iteration_index=1, precision=prec, ilu_ordering=order)
    opts.update(kw)
    sim.set_solver_options(**opts)


def tight_oracle(o):
    o.set_newton_options(max_iterations=30, min_iteration_index=1, residual_tol=1e-10, linear_rel_tol=1e-10,
                         linear_max_iterations=200)


DECKS = {
    "waterflood_1d": lambda: decks.waterflood_1d(5),
    "small_waterflood": lambda: decks.small_waterflood(6, 5, 3, n_steps=3),
    "small_black_oil": lambda: decks.small_black_oil(6, 5, 3, n_steps=3),
    "small_black_oil_5w": lambda: decks.small_black_oil(7, 7, 3, n_steps=2, five_wells=True),
    "small_black_oil_drsdt0": lambda: decks.without_redissolution(decks.small_black_oil(6, 5, 3, n_steps=3)),
}


@pytest.mark.parametrize("name", list(DECKS))
def test_assembly_matches_oracle(name, oracle_mod):
    d = DECKS[name]()
    log(f"[assembly] deck={d.name} n={d.n_cells}")
    o = d.build(oracle_mod.OracleSimulator)
    g = d.build(ReservoirSimulator)
    cs, rg = o.initial_state()
    cs_g, rg_g = g.debug_initial_state()
    np.testing.assert_array_equal(rg, rg_g)
    np.testing.assert_allclose(cs_g, cs, rtol=1e-13, atol=0)
    # perturbed iterate so that every term is exercised
    rng = np.random.default_rng(1)
    it = cs.copy()
    it[:, 0] += rng.uniform(-2, 2, size=len(it))
    it[:, 1] += rng.uniform(-0.02, 0.02, size=len(it))
    sat = rg == 0
    it[sat, 2] = np.maximum(it[sat, 2] + rng.uniform(0, 0.01, size=sat.sum()), 0)
    bhp, q = o.get_well_state()
    dt = d.schedule[0]
    R_o, (ptr, col, val) = o.assemble(dt, it, rg, bhp, q)
    R_g, A_g = g.debug_assemble(dt, it, rg, bhp, q, eliminate_wells=False)
    scale = np.maximum(np.abs(R_o), 1e-8 * np.abs(R_o).max() + 1e-300)
    err = np.abs(R_g - R_o) / scale
    log(f"[assembly] residual max rel err {err.max():.3e}")
    assert err.max() < 1e-9
    n = d.n_cells
    J = sp.csr_matrix((val, col, ptr), shape=(len(R_o), len(R_o)))
    Jrr = J[: 3 * n, : 3 * n].toarray()
    Jg = blocks_to_csr(A_g, d.nx, d.ny, d.nz).toarray()
    diff = np.abs(Jg - Jrr)
    tol = 1e-10 * np.abs(Jrr).max(axis=1, keepdims=True) + 1e-14
    log(f"[assembly] jacobian max abs err {diff.max():.3e}")
    assert (diff <= tol).all()


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", list(DECKS))
def test_newton_direction_matches_direct_solve(name, mode, oracle_mod):
    d = DECKS[name]()
    log(f"[direction] deck={d.name} mode={mode}")
    o = d.build(oracle_mod.OracleSimulator)
    g = d.build(ReservoirSimulator)
    tight(g, mode)
    cs, rg = o.initial_state()
    bhp, q = o.get_well_state()
    dt = d.schedule[0]
    R_o, (ptr, col, val) = o.assemble(dt, cs, rg, bhp, q)
    J = sp.csc_matrix(sp.csr_matrix((val, col, ptr), shape=(len(R_o), len(R_o))))
    dx_ref = spla.spsolve(J, -R_o)
    dx_g, its = g.debug_newton_direction(dt, cs, rg, bhp, q)
    rel = np.linalg.norm(dx_g - dx_ref) / np.linalg.norm(dx_ref)
    res = np.linalg.norm(J @ dx_g + R_o) / np.linalg.norm(R_o)
    log(f"[direction] iterations={its} rel err {rel:.3e} residual {res:.3e}")
    assert res < 1e-8
    assert rel < 1e-6


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", list(DECKS))
def test_substeps_match_oracle(name, mode, oracle_mod):
    d = DECKS[name]()
    log(f"[substeps] deck={d.name} mode={mode}")
    # the oracle's substep controller fixes the schedule; both sides then replay it
    sched = d.build(oracle_mod.OracleSimulator)
    tight_oracle(sched)
    for dt in d.schedule:
        assert sched.step(dt) == 0, sched.last_warning()
    dts = sched.accepted_dts()
    log(f"[substeps] schedule {np.round(dts, 4).tolist()}")
    o = d.build(oracle_mod.OracleSimulator)
    g = d.build(ReservoirSimulator)
    tight_oracle(o)
    tight(g, mode)
    for k, dt in enumerate(dts):
        rc_o, ni, li = o.run_substep(dt)
        rc_g, rep = g.run_timestep(dt)
        log(f"[substeps] step {k} oracle rc={rc_o} newton={ni}; gpu rc={rc_g} newton={rep.newton_iterations} "
            f"lin={rep.linear_iterations}")
        assert rc_o == 0 and rc_g == 0
        p_o, p_g = o.get_field("pressure"), g.get_pressures()
        assert np.max(np.abs(p_g - p_o) / np.abs(p_o)) < 1e-8
        for f in ("sw", "sg", "so", "rs"):
            a, b = o.get_field(f), g.get_field(f)
            tol = 1e-7 if f != "rs" else 1e-7 * max(1.0, np.abs(a).max())
            assert np.max(np.abs(a - b)) < tol, f
        bo, qo = o.get_well_state()
        bg, qg = g.get_well_state()
        assert np.max(np.abs(bg - bo) / np.abs(bo)) < 1e-8
        # per completion, relative to its own rate (a floor of 1e-9 of the largest rate for
        # completions that carry almost nothing)
        assert np.all(np.abs(qg - qo) <= 1e-7 * np.abs(qo) + 1e-9 * max(1.0, np.abs(qo).max()))


@pytest.mark.parametrize("mode", ["fp64_natural", "mixed_redblack"])
@pytest.mark.parametrize("name", list(DECKS))
def test_rate_history_matches_oracle(name, mode, oracle_mod):
    """The step report (rs/reporting.rs:553-734) recorded after every accepted substep on
    the oracle's schedule: rates and averages to the state tolerance, the cumulative
    balance errors (differences of near-equal volumes) to 1e-7 of the volumes in place."""
    d = DECKS[name]()
    sched = d.build(oracle_mod.OracleSimulator)
    tight_oracle(sched)
    for dt in d.schedule:
        assert sched.step(dt) == 0, sched.last_warning()
    dts = sched.accepted_dts()
    o = d.build(oracle_mod.OracleSimulator)
    g = d.build(ReservoirSimulator)
    tight_oracle(o)
    tight(g, mode)
    for dt in dts:
        assert o.run_substep(dt)[0] == 0
        assert g.run_timestep(dt)[0] == 0
    ho, hg = o.rate_history(), g.rate_history()
    assert len(ho["time"]) == len(hg["time"]) == len(dts)
    inv = o.inventory_totals()
    log(f"[report] deck={d.name} mode={mode} entries={len(dts)} oil={hg['total_production_oil'][-1]:.6e} "
        f"mb_oil={hg['material_balance_error_oil_m3'][-1]:.3e} (oracle {ho['material_balance_error_oil_m3'][-1]:.3e})")
    for f in ho:
        a, b = ho[f], hg[f]
        if f.startswith("material_balance_error"):
            k = {"material_balance_error_m3": 0, "material_balance_error_oil_m3": 1,
                 "material_balance_error_gas_m3": 2}[f]
            np.testing.assert_allclose(b, a, rtol=0, atol=1e-7 * max(abs(inv[k]), 1.0), err_msg=f)
        elif f.endswith("fraction") or f == "time":
            np.testing.assert_array_equal(b, a, err_msg=f)
        else:
            np.testing.assert_allclose(b, a, rtol=1e-7, atol=1e-9 * max(np.abs(a).max(), 1.0), err_msg=f)


@pytest.mark.parametrize("name", ["waterflood_1d", "small_waterflood", "small_black_oil"])
def test_controller_schedule_matches_oracle(name, oracle_mod):
    """Each side runs its own substep controller (rs/fim/timestep.rs:988-1890: first trial,
    OPM PID and iteration-count growth, 2x after a retry, retry x0.33) over the deck's
    report steps in FP64 natural order: the accepted dt sequences agree to rounding and the
    states to the substep bars."""
    d = DECKS[name]()
    o = d.build(oracle_mod.OracleSimulator)
    g = d.build(ReservoirSimulator)
    tight_oracle(o)
    tight(g, "fp64_natural")
    for dt in d.schedule:
        assert o.step(dt) == 0, o.last_warning()
        assert g.step(dt) == 0, g.last_warning()
    do, dg = o.accepted_dts(), g.accepted_dts()
    log(f"[controller] deck={d.name} oracle {np.round(do, 5).tolist()} gpu {np.round(dg, 5).tolist()}")
    assert len(do) == len(dg)
    np.testing.assert_allclose(dg, do, rtol=1e-9, atol=0)
    p_o, p_g = o.get_field("pressure"), g.get_pressures()
    assert np.max(np.abs(p_g - p_o) / np.abs(p_o)) < 1e-8


def _no_wells_segregation():
    """Gravity segregation with no wells: the 6x5x3 three-phase recipe with its wells
    removed and free gas placed in the bottom layer, so that buoyancy alone drives flow."""
    d = decks.small_black_oil(6, 5, 3, n_steps=3)
    calls = []
    for name, args in d.calls:
        if name in ("add_well_with_id", "set_well_schedule"):
            continue
        if name == "set_initial_state_arrays":
            p, sw, sg, rs = (np.array(a, dtype=np.float64) for a in args)
            n_layer = 6 * 5
            sg = sg.copy()
            sg[-n_layer:] = 0.3
            args = (p, sw, sg, rs)
        calls.append((name, args))
    return dataclasses.replace(d, name="segregation_6x5x3_no_wells", calls=calls)


EDGE_DECKS = {
    "areal_12x10x1": lambda: decks.small_waterflood(12, 10, 1, n_steps=2),
    "segregation_no_wells": _no_wells_segregation,
}


@pytest.mark.parametrize("mode", ["fp64_natural", "mixed_redblack"])
@pytest.mark.parametrize("name", list(EDGE_DECKS))
def test_edge_decks_match_oracle(name, mode, oracle_mod):
    """A one-layer areal grid and a grid with no wells at all (empty well tail), on the
    oracle controller's schedule, to the substep bars."""
    d = EDGE_DECKS[name]()
    sched = d.build(oracle_mod.OracleSimulator)
    tight_oracle(sched)
    for dt in d.schedule:
        assert sched.step(dt) == 0, sched.last_warning()
    dts = sched.accepted_dts()
    o = d.build(oracle_mod.OracleSimulator)
    g = d.build(ReservoirSimulator)
    tight_oracle(o)
    tight(g, mode)
    for dt in dts:
        assert o.run_substep(dt)[0] == 0
        rc, rep = g.run_timestep(dt)
        assert rc == 0
    log(f"[edge] deck={d.name} mode={mode} substeps={len(dts)}")
    p_o, p_g = o.get_field("pressure"), g.get_pressures()
    assert np.max(np.abs(p_g - p_o) / np.abs(p_o)) < 1e-8
    for f in ("sw", "sg", "so"):
        assert np.max(np.abs(o.get_field(f) - g.get_field(f))) < 1e-7, f
    if name == "segregation_no_wells":
        sg = g.get_field("sg").reshape(3, 5, 6)
        assert sg[1].mean() > 0.0  # free gas has risen out of the bottom layer
