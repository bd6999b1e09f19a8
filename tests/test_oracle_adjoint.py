"""The oracle's FP64 discrete adjoint (oracle/adjoint.py) against central finite
differences of the oracle's own forward map on the same substep schedule.

Random directions in (ln k, phi) test the whole gradient at once; coordinate
directions at single cells (an injector cell, a producer cell, an interior cell) catch
an index that lands on the wrong cell or the wrong parameter. The error of a central
difference at h is the truncation c h^2 plus rounding; the Richardson combination
(4 D(h/2) - D(h)) / 3 removes the h^2 term, so it is held to a tighter bound.
"""
import numpy as np
import pytest

from oracle import OracleSimulator
from oracle.adjoint import OracleAdjoint, nested_dissection
from ressim_amd import decks

OPTS = dict(max_iterations=30, min_iteration_index=1, residual_tol=1e-11, linear_rel_tol=1e-11,
            linear_max_iterations=300)


@pytest.fixture(scope="module")
def case(oracle_mod):
    d = decks.black_oil_3d(8, 8, 3, 30.0, 30.0, 5.0, sigma=1.0, corr=2.0, seed=11, n_steps=4, dt=2.0,
                           name="adjoint_8x8x3")
    o = d.build(OracleSimulator)
    o.set_newton_options(**OPTS)
    for dt in d.schedule:
        assert o.step(dt) == 0, o.last_warning()
    dts = o.accepted_dts()
    a = OracleAdjoint(d, OPTS)
    J = a.forward(dts)
    g_lnk, g_phi = a.gradient()
    fields = dict(d.calls)
    k = [np.asarray(x) for x in fields["set_permeability_field"]]
    phi0 = np.asarray(fields["set_porosity_field"][0])

    def objective(shift, phi):
        s = d.build(OracleSimulator)
        m = np.exp(shift)
        s.set_permeability_field(*(x * m for x in k))
        s.set_porosity_field(phi)
        s.set_newton_options(**OPTS)
        for dt in dts:
            rc, _, _ = s.run_substep(dt)
            assert rc == 0, s.last_warning()
        return s.producer_oil()

    return dict(deck=d, dts=dts, J=J, g_lnk=g_lnk, g_phi=g_phi, phi0=phi0, objective=objective)


def _fd(case, dk, dp, h):
    f = case["objective"]
    return (f(h * dk, case["phi0"] + h * dp) - f(-h * dk, case["phi0"] - h * dp)) / (2 * h)


def test_objective_replays_the_controlled_run(case):
    assert case["objective"](np.zeros(case["deck"].n_cells), case["phi0"]) == case["J"]
    assert len(case["dts"]) > len(case["deck"].schedule)  # the controller cut at least one step


def test_nested_dissection_is_a_permutation():
    p = nested_dissection(7, 5, 3)
    assert sorted(p.tolist()) == list(range(105))


def test_oracle_adjoint_random_directions(case):
    rng = np.random.default_rng(3)
    n = case["deck"].n_cells
    for _ in range(4):
        dk, dp = rng.standard_normal(n), 0.05 * rng.standard_normal(n)
        ad = case["g_lnk"] @ dk + case["g_phi"] @ dp
        d1, d2 = _fd(case, dk, dp, 1e-4), _fd(case, dk, dp, 5e-5)
        rich = (4 * d2 - d1) / 3
        print(f"fd {d1:.10e} richardson {rich:.10e} adjoint {ad:.10e}")
        assert abs(d1 - ad) <= 1e-6 * abs(ad)
        assert abs(rich - ad) <= 1e-7 * abs(ad)


@pytest.mark.parametrize("which,ijk", [("lnk", (0, 0, 1)), ("lnk", (7, 7, 2)), ("lnk", (4, 3, 1)),
                                       ("phi", (2, 5, 0)), ("phi", (7, 7, 0))])
def test_oracle_adjoint_single_cells(case, which, ijk):
    d = case["deck"]
    i = ijk[0] + d.nx * (ijk[1] + d.ny * ijk[2])
    dk, dp = np.zeros(d.n_cells), np.zeros(d.n_cells)
    (dk if which == "lnk" else dp)[i] = 1.0
    ad = (case["g_lnk"] if which == "lnk" else case["g_phi"])[i]
    fd = _fd(case, dk, dp, 1e-4 if which == "lnk" else 1e-5)
    print(f"{which}{ijk} fd {fd:.10e} adjoint {ad:.10e}")
    assert abs(fd - ad) <= 1e-6 * abs(ad) + 1e-9 * np.abs(case["g_lnk"]).max()
