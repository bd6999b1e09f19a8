"""Pins of the oracle (CPU restatement) against the reference's own known answers,
closed forms and invariants. CPU only.

Each test names the reference test or formula it reproduces (paths relative to
src/lib/ressim/src/).
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from oracle import OracleSimulator
from ressim_amd import decks


# --------------------------------------------------------------- AD rules
def test_ad_closed_forms():
    """fim/ad.rs:429-499: x*y/(x+y) at (2,3); exp(sqrt(4)); 4^1.5; 1/2; 10-1.5x; max branches."""
    v = oracle.ad_selftest()
    assert v[0] == pytest.approx(1.2, abs=1e-15)
    assert v[1] == pytest.approx(9 / 25, abs=1e-15)
    assert v[2] == pytest.approx(4 / 25, abs=1e-15)
    assert v[3] == pytest.approx(math.exp(2.0), rel=1e-15)
    assert v[4] == pytest.approx(math.exp(2.0) * 0.25, rel=1e-15)
    assert (v[5], v[6]) == (8.0, 3.0)
    assert (v[7], v[8]) == (0.5, -0.25)
    assert v[9] == pytest.approx(10 - 1.5 * 3) and v[10] == pytest.approx(-1.5)
    assert (v[11], v[12]) == (2.0, 0.0)  # constant branch wins -> no derivative
    assert (v[13], v[14]) == (1.0, 1.0)  # value above floor keeps its derivative


# --------------------------------------------------------------- PVT
PVTO = np.array([[50.0, 20.0, 1.10, 1.00, 0.02, 0.015],
                 [350.0, 20.0, 1.0675, 1.030, 0.02, 0.015],
                 [150.0, 80.0, 1.25, 0.70, 0.008, 0.018],
                 [350.0, 80.0, 1.2252, 0.714, 0.008, 0.018]])


def _pvto_sim():
    s = OracleSimulator(1, 1, 1, 0.2)
    s.set_fluid_compressibilities(1e-4, 3e-6)
    s.set_three_phase_mode_enabled(True)
    s.set_pvt_table(PVTO)
    return s


def _seg(r0, r1, t):
    ib = 1 / r0[2] + t * (1 / r1[2] - 1 / r0[2])
    ibm = 1 / (r0[2] * r0[3]) + t * (1 / (r1[2] * r1[3]) - 1 / (r0[2] * r0[3]))
    return 1 / ib, ib / ibm


def test_pvto_guided_interpolation_known_answer():
    """pvt.rs:1040-1100: Bo, mu_o at (200 bar, Rs 50) interpolate 1/Bo and 1/(Bo mu) along
    the branches at the guided pressures 150 and 250 bar, then in Rs."""
    s = _pvto_sim()
    lo_bo, lo_mu = _seg(PVTO[0], PVTO[1], 1 / 3)
    hi_bo, hi_mu = _seg(PVTO[2], PVTO[3], 0.5)
    ib = 1 / lo_bo + 0.5 * (1 / hi_bo - 1 / lo_bo)
    ibm = 1 / (lo_bo * lo_mu) + 0.5 * (1 / (hi_bo * hi_mu) - 1 / (lo_bo * lo_mu))
    out = s.probe(1, 200.0, 50.0)
    assert out[0] == pytest.approx(1 / ib, abs=1e-14)
    assert out[3] == pytest.approx(ib / ibm, abs=1e-14)


def test_pvt_derivatives_match_central_differences():
    """AD derivatives of the table interpolation against central differences
    (fim/numjac.rs:13-82 tolerances: rel 1e-6, abs 1e-9)."""
    s = _pvto_sim()
    for p, rs in [(120.0, 30.0), (200.0, 50.0), (300.0, 70.0), (400.0, 60.0)]:
        out = s.probe(1, p, rs)
        h = 1e-6 * (1 + abs(p))
        fp = (s.probe(1, p + h, rs)[0] - s.probe(1, p - h, rs)[0]) / (2 * h)
        hr = 1e-6 * (1 + abs(rs))
        fr = (s.probe(1, p, rs + hr)[0] - s.probe(1, p, rs - hr)[0]) / (2 * hr)
        assert abs(out[1] - fp) <= 1e-9 + 1e-6 * abs(fp)
        assert abs(out[2] - fr) <= 1e-9 + 1e-6 * abs(fr)


def test_saturated_extrapolation_above_table():
    """pvt.rs:281-297: above the last saturated row (the first row of the highest-Rs
    branch, here p=150) Bo = Bo_last exp(-c_o (p - p_last)) and Bg follows Boyle's law."""
    s = _pvto_sim()
    out = s.probe(0, 400.0)
    last = PVTO[2]
    assert out[2] == pytest.approx(last[2] * math.exp(-1e-4 * (400 - 150)), rel=1e-14)
    assert out[4] == pytest.approx(last[4] * 150 / 400, rel=1e-14)
    assert out[0] == last[1]


# --------------------------------------------------------------- relperm / Pc
def test_corey_table_knots_and_slopes():
    """relperm.rs:602-635: 21 equal knots over [Swc, 1-Sor]; values at knots equal Corey,
    the derivative is the segment slope, constant extension outside."""
    s = OracleSimulator(1, 1, 1, 0.2)
    s.set_rel_perm_props(0.1, 0.1, 2.0, 2.0, 1.0, 1.0)
    step = 0.8 / 20
    for k in range(21):
        sw = 0.1 + k * step
        se = (sw - 0.1) / 0.8
        out = s.probe(2, sw)
        assert out[0] == pytest.approx(se ** 2, abs=1e-14)
        assert out[2] == pytest.approx((1 - se) ** 2, abs=1e-14)
    sw = 0.1 + 3.5 * step
    out = s.probe(2, sw)
    x0, x1 = 0.1 + 3 * step, 0.1 + 4 * step
    slope = (((x1 - 0.1) / 0.8) ** 2 - ((x0 - 0.1) / 0.8) ** 2) / step
    assert out[1] == pytest.approx(slope, rel=1e-12)
    assert s.probe(2, 0.05)[1] == 0.0 and s.probe(2, 0.95)[1] == 0.0


def test_stone2_endpoint_derivative_is_frozen():
    """relperm.rs:446-471: Stone II at Sw=Swc, Sg=0 sits at kro_max with zero derivative."""
    s = OracleSimulator(1, 1, 1, 0.2)
    s.set_three_phase_rel_perm_props(0.15, 0.15, 0.05, 0.05, 0.1, 2, 2, 3, 0.6, 1.0, 0.8)
    out = s.probe(3, 0.15, 0.0)
    assert out[0] == 1.0 and out[1] == 0.0 and out[2] == 0.0
    out = s.probe(3, 0.4, 0.2)
    h = 1e-7
    fd = (s.probe(3, 0.4 + h, 0.2)[0] - s.probe(3, 0.4 - h, 0.2)[0]) / (2 * h)
    assert out[1] == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_brooks_corey_pc_cap_and_value():
    """capillary.rs:73-90: Pc = Pe Se^(-1/lambda), capped at 20 Pe, zero at Se=1."""
    s = OracleSimulator(1, 1, 1, 0.2)
    s.set_rel_perm_props(0.1, 0.1, 2, 2, 1, 1)
    s.set_capillary_params(5.0, 2.0)
    assert s.probe(4, 0.1)[0] == 100.0
    assert s.probe(4, 0.9)[0] == 0.0
    se = (0.5 - 0.1) / 0.8
    assert s.probe(4, 0.5)[0] == pytest.approx(5.0 * se ** -0.5, rel=1e-15)


# --------------------------------------------------------------- geometry / wells
def test_peaceman_well_index_closed_form():
    """fim/wells.rs:800-826 and fim/tests/wells.rs:13-50: WI = c 2 pi sqrt(kx ky) dz / (ln(r_eq/rw)+s)."""
    s = OracleSimulator(1, 1, 1, 0.2)
    s.set_cell_dimensions(20.0, 30.0, 4.0)
    s.set_permeability_field([200.0], [50.0], [10.0])
    s.add_well(0, 0, 0, 100.0, 0.1, 1.5, False)
    kx, ky, dx, dy = 200.0, 50.0, 20.0, 30.0
    r_eq = 0.28 * math.sqrt(math.sqrt(kx / ky) * dx ** 2 + math.sqrt(ky / kx) * dy ** 2) / (
        (kx / ky) ** 0.25 + (ky / kx) ** 0.25)
    wi = 8.5269888e-3 * 2 * math.pi * math.sqrt(kx * ky) * 4.0 / (math.log(r_eq / 0.1) + 1.5)
    assert s.probe(5, 0.0)[0] == pytest.approx(wi, rel=1e-14)


def test_tpfa_transmissibility_formulas():
    """grid.rs:278-311: harmonic mean in x/y, half-cell series in z (fim/assembly_tests.rs:684-771)."""
    s = OracleSimulator(2, 2, 2, 0.2)
    s.set_cell_dimensions_per_layer(10.0, 20.0, [2.0, 6.0])
    kx = np.array([100, 300, 50, 80, 10, 20, 30, 40.0])
    ky = kx * 2
    kz = kx / 10
    s.set_permeability_field(kx, ky, kz)
    assert s.probe(6, 0, 1, init=0.0)[0] == pytest.approx(2 * 100 * 300 / 400 * (20 * 2) / 10, rel=1e-15)
    assert s.probe(6, 0, 2, init=1.0)[0] == pytest.approx(2 * 200 * 100 / 300 * (10 * 2) / 20, rel=1e-15)
    assert s.probe(6, 0, 4, init=2.0)[0] == pytest.approx(2 * 10 * 1 * 200 / (1 * 2 + 10 * 6), rel=1e-15)


# --------------------------------------------------------------- Newton policy
def test_cnv_mb_hand_values():
    """fim/newton/tests.rs:294-324."""
    d = oracle.cnv_mb([2.0, -0.6, 0.0, -1.0, 0.9, 0.0], [100.0, 300.0], [[1.0, 1.2, 0.01], [1.0, 1.4, 0.03]])
    assert d["cnv"][0] == pytest.approx(0.02, abs=1e-12)
    assert d["cnv"][1] == pytest.approx(1.3 * 0.006, abs=1e-12)
    assert d["cnv"][2] == 0.0
    assert d["mb"][0] == pytest.approx(1.0 / 400, abs=1e-12)
    assert d["mb"][1] == pytest.approx(1.3 * 0.3 / 400, abs=1e-12)
    assert d["violating_pv_fraction"] == pytest.approx(0.25, abs=1e-12)
    assert not d["pv_rule_relaxes"] and not d["would_accept"]


def test_cnv_pv_rule_and_final_iteration_relaxation():
    """fim/newton/tests.rs:326-384."""
    d = oracle.cnv_mb([0.5, 0, 0, -0.5, 0, 0], [10.0, 990.0], np.ones((2, 3)))
    assert d["mb"][0] == 0.0 and d["pv_rule_relaxes"] and d["would_accept"]
    r = [25.0, 0, 0, -25.0, 0, 0]
    assert not oracle.cnv_mb(r, [500.0, 500.0], np.ones((2, 3)))["would_accept"]
    assert oracle.cnv_mb(r, [500.0, 500.0], np.ones((2, 3)), relax_final=True)["would_accept"]


def _two_cells(regime):
    hc = 0.1 if regime == 0 else 50.0
    return np.array([[200.0, 0.3, hc], [200.0, 0.3, hc]]), np.array([regime, regime])


def test_chop_counts_implied_oil_saturation():
    """fim/newton/tests.rs:159-180: dSw=dSg=0.15 -> implied dSo=-0.3 sets satAlpha=0.2/0.3."""
    cells, rg = _two_cells(0)
    up = np.zeros(6)
    up[1], up[2], up[4], up[5] = 0.15, 0.15, 0.01, -0.02
    c = oracle.chop(cells, rg, [], up)
    a = 0.2 / 0.3
    assert c[1] == pytest.approx(0.15 * a, abs=1e-12) and c[2] == pytest.approx(0.15 * a, abs=1e-12)
    assert c[4] == 0.01 and c[5] == -0.02


def test_chop_pressure_relative_and_rs_guard():
    """fim/newton/tests.rs:182-223."""
    cells, rg = _two_cells(0)
    up = np.zeros(6)
    up[0], up[1], up[3] = -100.0, 0.5, 30.0
    c = oracle.chop(cells, rg, [], up)
    assert c[0] == pytest.approx(-60.0) and c[1] == pytest.approx(0.2) and c[3] == 30.0
    cells, rg = _two_cells(1)
    up = np.zeros(6)
    up[2], up[1], up[5] = -80.0, 0.25, -10.0
    c = oracle.chop(cells, rg, [], up)
    assert c[2] == pytest.approx(-50.0) and c[1] == pytest.approx(0.2) and c[5] == -10.0


def test_chop_relaxation_first_and_bhp_limits():
    """fim/newton/tests.rs:225-292: relaxation before chop; dBHP <= |BHP|; floor at 1 bar."""
    cells, rg = _two_cells(0)
    up = np.zeros(6)
    up[1] = 0.3
    assert oracle.chop(cells, rg, [], up, relax=0.5)[1] == pytest.approx(0.15)
    one = np.array([[300.0, 0.3, 0.0]])
    assert oracle.chop(one, [0], [500.0], [0, 0, 0, 600.0])[3] == pytest.approx(500.0)
    assert oracle.chop(one, [0], [500.0], [0, 0, 0, 200.0])[3] == 200.0
    assert 1.5 + oracle.chop(one, [0], [1.5], [0, 0, 0, -3.0])[3] == pytest.approx(1.0)


# --------------------------------------------------------------- assembly invariants
def _fd_jacobian_check(deck, seed, rounding=False):
    """Columns of the AD Jacobian against central differences (rs/fim/numjac.rs:13-82:
    h = 1e-6 (1 + |x|), rel 1e-6 / abs 1e-9). With `rounding`, the tolerance also admits the
    rounding error of the difference itself, 4 eps max|R| / h per row: for couplings many
    orders below their row's magnitude (SPE1's water relperm is 1e-5 at most) it, not the
    Jacobian, decides the last digits."""
    o = deck.build(OracleSimulator)
    cs, rg = o.initial_state()
    rng = np.random.default_rng(seed)
    it = cs.copy()
    it[:, 0] += rng.uniform(-3, 3, size=len(it))
    it[:, 1] += rng.uniform(-0.03, 0.03, size=len(it))
    sat = rg == 0
    it[sat, 2] = np.maximum(it[sat, 2] + rng.uniform(0.001, 0.02, size=sat.sum()), 0.0)
    bhp, q = o.get_well_state()
    dt = deck.schedule[0]
    R, (ptr, col, val) = o.assemble(dt, it, rg, bhp, q)
    J = sp.csr_matrix((val, col, ptr), shape=(len(R), len(R))).toarray()
    n = deck.n_cells
    x0 = np.concatenate([it.ravel(), bhp, q])
    worst = 0.0
    for k in range(len(x0)):
        h = 1e-6 * (1 + abs(x0[k]))
        xp, xm = x0.copy(), x0.copy()
        xp[k] += h
        xm[k] -= h

        def res(x):
            return o.assemble(dt, x[:3 * n].reshape(-1, 3), rg, x[3 * n:3 * n + len(bhp)], x[3 * n + len(bhp):])[0]

        rp, rm = res(xp), res(xm)
        fd = (rp - rm) / (2 * h)
        tol = 1e-9 + 1e-6 * np.maximum(np.abs(fd), np.abs(J[:, k]))
        if rounding:
            tol = tol + 4 * np.finfo(float).eps * np.maximum(np.abs(rp), np.abs(rm)) / h
        bad = np.abs(J[:, k] - fd) > tol
        # at a clamp or upwind kink inside [x-h, x+h] the two one-sided slopes differ;
        # the reference keeps the live branch there, so AD must equal one of them
        fwd, bwd = (rp - R) / h, (R - rm) / h
        kink = np.abs(fwd - bwd) > 1e-3 * np.maximum(np.abs(fwd), np.abs(bwd)) + 1e-9
        one_sided = (np.abs(J[:, k] - fwd) <= 1e-4 * np.abs(fwd) + 1e-7) | (
            np.abs(J[:, k] - bwd) <= 1e-4 * np.abs(bwd) + 1e-7)
        bad &= ~(kink & one_sided)
        worst = max(worst, bad.sum())
    return worst


def test_ad_jacobian_matches_central_differences_waterflood():
    """fim/numjac.rs:13-82 harness (h = 1e-6(1+|x|), rel 1e-6 / abs 1e-9) on a two-phase deck."""
    assert _fd_jacobian_check(decks.small_waterflood(4, 3, 2, n_steps=1), 5) == 0


def test_ad_jacobian_matches_central_differences_black_oil():
    """Same harness on a three-phase deck with gravity, Leverett-J Pc and rate-controlled wells."""
    assert _fd_jacobian_check(decks.small_black_oil(3, 3, 2, n_steps=1), 6) == 0


def test_ad_jacobian_matches_central_differences_without_redissolution():
    """The harness with gas redissolution disabled (rs/fim/assembly_ad.rs:1857
    three_phase_drsdt0_disabled_with_wells): the capped Rs and the capped overflow
    flash keep the Jacobian exact."""
    assert _fd_jacobian_check(decks.without_redissolution(decks.small_black_oil(3, 3, 2, n_steps=1)), 6) == 0


def test_rs_never_rises_without_redissolution():
    """DRSDT 0: gas injected into undersaturated oil stays free; no cell's Rs rises
    above its value at the start of a substep (rs/fim/state.rs:1041-1070)."""
    d = decks.without_redissolution(decks.small_black_oil(5, 4, 3, n_steps=3, dt=2.0))
    o = d.build(OracleSimulator)
    o.set_newton_options(residual_tol=1e-10, min_iteration_index=1, linear_rel_tol=1e-10, linear_max_iterations=200)
    rs_prev = o.get_field("rs")
    sg_max = 0.0
    for dt in d.schedule:
        assert o.step(dt) == 0, o.last_warning()
        rs = o.get_field("rs")
        assert np.all(rs <= rs_prev + 1e-9)
        rs_prev = rs
        sg_max = max(sg_max, o.get_field("sg").max())
    assert sg_max > 1e-3  # free gas did appear


def test_face_fluxes_conserve_every_component():
    """fim/assembly_tests.rs:426-590: each face adds +F to one cell and -F to the other, so
    the flux part of the residual sums to zero per component (no wells)."""
    d = decks.small_black_oil(4, 3, 2, n_steps=1)
    d.calls = [c for c in d.calls if not c[0].startswith(("add_well", "set_well_s", "set_well_i"))]
    o = d.build(OracleSimulator)
    cs, rg = o.initial_state()
    rng = np.random.default_rng(2)
    it = cs.copy()
    it[:, 0] += rng.uniform(-5, 5, size=len(it))
    it[:, 1] += rng.uniform(0, 0.05, size=len(it))
    R1 = o.assemble(5.0, it, rg, [], [])[0]
    R0 = o.assemble(0.0, it, rg, [], [])[0]  # dt = 0 leaves the accumulation only
    flux = (R1 - R0).reshape(-1, 3)
    assert np.abs(flux.sum(axis=0)).max() <= 1e-9 * np.abs(flux).max()
    assert np.abs(flux).max() > 0


def test_uniform_state_has_zero_flux():
    """fim/assembly_tests.rs:426: a uniform state without gravity or wells has no flux."""
    d = decks.small_waterflood(4, 3, 2, n_steps=1)
    d.calls = [c for c in d.calls if not c[0].startswith(("add_well", "set_gravity", "set_porosity",
                                                          "set_permeability"))]
    o = d.build(OracleSimulator)
    cs, rg = o.initial_state()
    R = o.assemble(1.0, cs, rg, [], [])[0]
    assert np.abs(R).max() == 0.0


# --------------------------------------------------------------- linear stack
@pytest.mark.parametrize("deck", [lambda: decks.small_black_oil(6, 5, 3, n_steps=1, five_wells=True),
                                  lambda: decks.waterflood_1d(1)])
def test_schur_cpr_fgmres_matches_direct_solve(deck):
    """fim/linear/well_schur.rs:534-567 (Schur = direct full solve) and
    fim/linear/gmres_block_jacobi.rs:2736 (FGMRES = direct under a fixed preconditioner)."""
    d = deck()
    o = d.build(OracleSimulator)
    cs, rg = o.initial_state()
    bhp, q = o.get_well_state()
    R, (ptr, col, val) = o.assemble(d.schedule[0], cs, rg, bhp, q)
    J = sp.csc_matrix(sp.csr_matrix((val, col, ptr), shape=(len(R), len(R))))
    ref = spla.spsolve(J, -R)
    rc, dx, its = o.newton_direction(d.schedule[0], cs, rg, bhp, q, rel_tol=1e-11)
    assert rc == 0
    assert np.linalg.norm(J @ dx + R) <= 1e-9 * np.linalg.norm(R)
    assert np.linalg.norm(dx - ref) <= 1e-6 * np.linalg.norm(ref)


# --------------------------------------------------------------- full runs
def test_waterflood_mass_balance_and_front():
    """Injected water is stored in the reservoir (closed boundaries, producer barely
    producing water), and the front moves monotonically from the injector."""
    d = decks.waterflood_1d(30)
    o = d.build(OracleSimulator)
    o.set_newton_options(residual_tol=1e-9, min_iteration_index=1)
    for dt in d.schedule:
        assert o.step(dt) == 0
    sw, p = o.get_field("sw"), o.get_field("pressure")
    pv = 1.0 * 10 * 10 * 0.2
    inv_bw = lambda pr: 1 + 3e-6 * (pr - 300) * (1 + 3e-6 * (pr - 300) / 2)  # pvt.rs:603-606, Bw_ref 1
    stored_sc = (pv * sw * inv_bw(p)).sum() - (pv * 0.1 * inv_bw(300.0)) * len(sw)
    # 0.5 rm3/d injected at the injector cell's pressure for 30 d; the producer sits at Swc
    assert stored_sc == pytest.approx(0.5 * 30 * inv_bw(p[0]), rel=2e-4)
    front = sw[:40]
    assert np.all(np.diff(front) <= 1e-7)  # ahead of the front Sw only moves by water expansion


def test_producer_oil_objective_equals_oil_in_place_drop():
    """The adjoint objective (surface oil of the producers, summed per accepted
    substep) equals the drop of oil in place: the gas injector carries no oil and the
    boundaries are closed, so the oil component balance closes it exactly (to the
    Newton tolerance)."""
    d = decks.black_oil_3d(8, 8, 3, 30.0, 30.0, 5.0, sigma=1.0, corr=2.0, seed=11, n_steps=4, dt=2.0)
    o = d.build(OracleSimulator)
    o.set_newton_options(residual_tol=1e-11, min_iteration_index=1, linear_rel_tol=1e-11, linear_max_iterations=300)
    inv0 = o.inventory_totals()
    for dt in d.schedule:
        assert o.step(dt) == 0, o.last_warning()
    drop = inv0[1] - o.inventory_totals()[1]
    assert drop > 0
    assert o.producer_oil() == pytest.approx(drop, rel=1e-9)


# ---------------------------------------------------------------- step reporting
def _report_deck(**kw):
    d = decks.small_black_oil(6, 5, 3, n_steps=3, **kw)
    return d


def _pore_volumes(d):
    phi = dict(d.calls)["set_porosity_field"][0]
    dx, dy, dz = dict(d.calls)["set_cell_dimensions"]
    return phi * dx * dy * dz


def test_report_averages_and_gor_follow_the_stored_fields():
    """rs/reporting.rs:304-327 PV-weighted pressure and :646-666 plain saturation means,
    recomputed from the fields the oracle stores; GOR = gas / oil (:668-672)."""
    d = _report_deck()
    o = d.build(OracleSimulator)
    for dt in d.schedule:
        assert o.step(dt) == 0, o.last_warning()
    h = o.rate_history()
    assert len(h["time"]) == len(o.accepted_dts())
    np.testing.assert_allclose(h["time"], np.cumsum(o.accepted_dts()), rtol=1e-14)
    pv = _pore_volumes(d)
    p, sw, sg = o.get_field("pressure"), o.get_field("sw"), o.get_field("sg")
    assert abs(h["avg_reservoir_pressure"][-1] - (pv @ p) / pv.sum()) < 1e-10 * p.max()
    assert abs(h["avg_water_saturation"][-1] - sw.mean()) < 1e-14
    assert abs(h["avg_gas_saturation"][-1] - sg.mean()) < 1e-14
    np.testing.assert_allclose(h["producing_gor"], h["total_production_gas"] / h["total_production_oil"], rtol=1e-14)


def test_report_rate_controlled_injector_meets_its_target_and_balances_close():
    """The deck's gas injector is rate-controlled at 10,800 sc m3/day and never reaches its
    BHP limit, so the reported injection equals the target (within the Newton
    tolerance), with no BHP-limited injector. Water is not injected and oil is not
    created: their cumulative balance errors (reported minus inventory change, on the
    reference's geometric pore volume, rock compressibility 0) stay at round-off of the
    converged state relative to what moved."""
    d = _report_deck()
    o = d.build(OracleSimulator)
    o.set_newton_options(max_iterations=30, min_iteration_index=1, residual_tol=1e-10, linear_rel_tol=1e-10,
                         linear_max_iterations=200)
    for dt in d.schedule:
        assert o.step(dt) == 0, o.last_warning()
    h = o.rate_history()
    np.testing.assert_allclose(h["total_injection"], 10800.0, rtol=1e-6)
    assert np.all(h["injector_bhp_limited_fraction"] == 0.0)
    assert np.all(h["producer_bhp_limited_fraction"] == 0.0)  # BHP-controlled: not rate-controlled
    produced_oil = np.sum(h["total_production_oil"] * np.diff(np.concatenate([[0.0], h["time"]])))
    assert h["material_balance_error_oil_m3"][-1] < 1e-6 * produced_oil
    water_in_place = float(_pore_volumes(d) @ o.get_field("sw"))
    assert h["material_balance_error_m3"][-1] < 1e-9 * water_in_place
    # gas: injected minus produced minus the change in place (free and dissolved) closes
    # to round-off of what was injected
    injected_gas = np.sum(h["total_injection"] * np.diff(np.concatenate([[0.0], h["time"]])))
    assert abs(h["material_balance_error_gas_m3"][-1]) < 1e-9 * injected_gas


def test_report_bhp_limited_producer():
    """A rate-controlled producer asked for far more oil than it can deliver sits at its BHP
    limit: producer_bhp_limited_fraction is 1 and the reported oil rate is below the
    target (the case of rs/fim/tests/wells.rs:195-250)."""
    d = _report_deck()
    d.calls.append(("set_well_schedule", ("PROD", "rate", np.nan, 1e5, 150.0, True)))
    o = d.build(OracleSimulator)
    assert o.step(d.schedule[0]) == 0, o.last_warning()
    h = o.rate_history()
    assert h["producer_bhp_limited_fraction"][-1] == 1.0
    assert 0.0 < h["total_production_oil"][-1] < 1e5
    bhp, q = o.get_well_state()
    assert abs(bhp[1] - 150.0) < 1e-6


# --------------------------------------------------------------- sweep metrics
def _swept_run(config, n_steps=4):
    d = decks.small_waterflood(8, 6, 3, n_steps=n_steps, dt=2.0)
    o = d.build(OracleSimulator)
    o.set_sweep_config(config)
    for dt in d.schedule:
        assert o.step(dt) == 0, o.last_warning()
    return o


@pytest.mark.parametrize("geometry", ["areal", "vertical", "both", "unknown"])
def test_sweep_metrics_count_swept_cells_and_columns(geometry):
    """rs/reporting.rs:52-123: a cell is swept when sw > threshold (strict); E_vol is the
    swept fraction of cells, E_A the fraction of (i, j) columns holding a swept cell;
    "areal" reports E_V = 1, "vertical" E_A = 1 and E_V = E_vol, "both" (and any other
    name) only the recovered fraction of the initial mobile oil."""
    cfg = dict(geometry=geometry, swept_threshold=0.21, initial_oil_saturation=0.8, residual_oil_saturation=0.1)
    o = _swept_run(cfg)
    h = o.rate_history()
    sw = o.get_field("sw").reshape(3, 6, 8)
    so = o.get_field("so")
    swept = sw > 0.21
    e_vol = swept.mean()
    e_a = swept.any(axis=0).mean()
    assert 0.0 < e_vol < 1.0 and e_a > e_vol  # the front has entered but not filled the box
    assert h["sweep_e_vol"][-1] == pytest.approx(e_vol, abs=1e-15)
    if geometry == "areal":
        assert h["sweep_e_a"][-1] == pytest.approx(e_a, abs=1e-15) and h["sweep_e_v"][-1] == 1.0
        assert np.isnan(h["sweep_mobile_oil_recovered"][-1])
    elif geometry == "vertical":
        assert h["sweep_e_a"][-1] == 1.0 and h["sweep_e_v"][-1] == h["sweep_e_vol"][-1]
        assert np.isnan(h["sweep_mobile_oil_recovered"][-1])
    else:
        assert np.isnan(h["sweep_e_a"][-1]) and np.isnan(h["sweep_e_v"][-1])
        rec = 1.0 - np.maximum(so - 0.1, 0.0).sum() / (so.size * 0.7)
        assert h["sweep_mobile_oil_recovered"][-1] == pytest.approx(min(max(rec, 0.0), 1.0), rel=1e-13)
    # one entry per accepted substep, each with its own sweep state; E_vol never falls
    assert len(h["sweep_e_vol"]) == len(o.accepted_dts())
    assert np.all(np.diff(h["sweep_e_vol"]) >= 0.0)


def test_sweep_metrics_threshold_is_strict_and_off_without_config():
    """A threshold at the largest water saturation sweeps nothing; with no configuration
    every sweep field is None (NaN here); an initial oil saturation at or below Sor makes
    the recovered fraction 0 (rs/reporting.rs:105-117)."""
    o = _swept_run(None, n_steps=2)
    h = o.rate_history()
    for f in ("sweep_e_a", "sweep_e_v", "sweep_e_vol", "sweep_mobile_oil_recovered"):
        assert np.all(np.isnan(h[f]))
    top = o.get_field("sw").max()
    d = decks.small_waterflood(8, 6, 3, n_steps=2, dt=2.0)
    o2 = d.build(OracleSimulator)
    for dt in d.schedule:
        assert o2.step(dt) == 0
    np.testing.assert_array_equal(o2.get_field("sw"), o.get_field("sw"))
    o2.set_sweep_config(dict(geometry="areal", swept_threshold=top, initial_oil_saturation=0.1,
                             residual_oil_saturation=0.2))
    assert o2.step(1e-3) == 0
    h2 = o2.rate_history()
    assert h2["sweep_e_vol"][-1] == 0.0 or o2.get_field("sw").max() > top
    o2.set_sweep_config(dict(geometry="both", swept_threshold=top, initial_oil_saturation=0.1,
                             residual_oil_saturation=0.2))
    assert o2.step(1e-3) == 0
    assert o2.rate_history()["sweep_mobile_oil_recovered"][-1] == 0.0


@pytest.mark.parametrize("case", ["A", "B"])
def test_buckley_leverett_breakthrough(case):
    """rs/tests/buckley.rs:229-283 on the FIM path: breakthrough PV within 25 % (case A) or
    30 % (case B) of the Welge tangent."""
    from buckley_check import breakthrough_pv, welge_breakthrough_pv

    d = decks.buckley_leverett(case)
    o = d.build(OracleSimulator)
    total_pv = 0.2 * 10.0 * 10.0 * 1.0 * d.n_cells  # default 10 m x 10 m x 1 m cells, phi 0.2
    pv = breakthrough_pv(o, d, total_pv)
    ref = welge_breakthrough_pv(d.info)
    print(f"[buckley] case {case}: breakthrough {pv:.4f} PV, Welge {ref:.4f} PV")
    assert abs(pv - ref) / ref <= d.info["rel_tol"], (pv, ref)
