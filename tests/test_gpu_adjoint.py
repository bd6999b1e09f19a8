"""Discrete adjoint of the FIM step (extension): gradients of the producers' surface
oil with respect to per-cell log-permeability and porosity.

The adjoint gradient is checked along seeded random directions against central
finite differences (h = 1e-4) of the forward map on the same substep schedule,
computed (a) with the GPU forward and (b) with the oracle's forward on the CPU.
(b) shares no code with the GPU. The gradients themselves are compared entry by entry
with the oracle's FP64 adjoint (transposed direct solves of its AD Jacobian, pinned on
the CPU against the oracle's own finite differences in test_oracle_adjoint). The
objective is pinned on the CPU against the oil-in-place balance (test_oracle_pins).
"""
import sys

import numpy as np
import pytest

from ressim_amd import decks
from ressim_amd.adjoint import FimAdjoint

pytestmark = pytest.mark.gpu

H = 1e-4
TIGHT = dict(newton_residual_tol=1e-11, newton_mb_tol=1e-13, min_newton_iteration_index=1,
             linear_rel_tol=1e-11, linear_max_iterations=300, linear_restart=60, max_newton_iterations=30)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_deck():
    return decks.black_oil_3d(8, 8, 3, 30.0, 30.0, 5.0, sigma=1.0, corr=2.0, seed=11, n_steps=4, dt=2.0,
                              name="adjoint_8x8x3")


def directions(n, k, seed):
    rng = np.random.default_rng(seed)
    return [(rng.standard_normal(n), 0.05 * rng.standard_normal(n)) for _ in range(k)]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_adjoint_matches_fd_of_gpu_forward(precision):
    d = small_deck()
    fn = FimAdjoint(d, checkpoint_interval=2, solver_options=dict(TIGHT, precision=precision))
    J0, dts = fn.forward()
    g_lnk, g_phi = fn.backward()
    log(f"[adjoint] J={J0:.10e} substeps={len(dts)} report={fn.report}")
    phi0 = fn._phi
    fd_fn = FimAdjoint(d, checkpoint_interval=1000, solver_options=dict(TIGHT, precision=precision), replay=dts)
    worst = 0.0
    for dk, dp in directions(d.n_cells, 4, 3):
        jp, _ = fd_fn.forward(H * dk, phi0 + H * dp)
        jm, _ = fd_fn.forward(-H * dk, phi0 - H * dp)
        fd = (jp - jm) / (2 * H)
        ad = g_lnk @ dk + g_phi @ dp
        rel = abs(fd - ad) / abs(fd)
        worst = max(worst, rel)
        log(f"[adjoint] fd {fd:.10e} adjoint {ad:.10e} rel {rel:.2e}")
    fn.close()
    fd_fn.close()
    assert worst < 1e-6


def test_adjoint_matches_fd_of_oracle_forward(oracle_mod):
    d = small_deck()
    fn = FimAdjoint(d, checkpoint_interval=2, solver_options=dict(TIGHT, precision=0))
    J0, dts = fn.forward()
    g_lnk, g_phi = fn.backward()
    fields = dict(d.calls)
    kx, ky, kz = (np.asarray(a) for a in fields["set_permeability_field"])
    phi0 = np.asarray(fields["set_porosity_field"][0])

    def oracle_J(shift, phi):
        o = d.build(oracle_mod.OracleSimulator)
        m = np.exp(shift)
        o.set_permeability_field(kx * m, ky * m, kz * m)
        o.set_porosity_field(phi)
        o.set_newton_options(max_iterations=30, min_iteration_index=1, residual_tol=1e-11, linear_rel_tol=1e-11,
                             linear_max_iterations=300)
        for dt in dts:
            rc, _, _ = o.run_substep(dt)
            assert rc == 0, o.last_warning()
        return o.producer_oil()

    j_cpu = oracle_J(np.zeros(d.n_cells), phi0)
    log(f"[adjoint] objective gpu {J0:.12e} oracle {j_cpu:.12e}")
    assert abs(J0 - j_cpu) <= 1e-9 * abs(j_cpu)
    worst = 0.0
    for dk, dp in directions(d.n_cells, 3, 5):
        fd = (oracle_J(H * dk, phi0 + H * dp) - oracle_J(-H * dk, phi0 - H * dp)) / (2 * H)
        ad = g_lnk @ dk + g_phi @ dp
        rel = abs(fd - ad) / abs(fd)
        worst = max(worst, rel)
        log(f"[adjoint] oracle fd {fd:.10e} gpu adjoint {ad:.10e} rel {rel:.2e}")
    fn.close()
    assert worst < 1e-5


def _oracle_gradient(oracle_mod, d, dts, opts):
    from oracle.adjoint import OracleAdjoint

    a = OracleAdjoint(d, opts)
    J = a.forward(dts)
    g_lnk, g_phi = a.gradient()
    return J, g_lnk, g_phi


def _compare_to_oracle(name, J0, g, Jo, go, tol):
    log(f"[adjoint] {name} objective gpu {J0:.12e} oracle {Jo:.12e}")
    assert abs(J0 - Jo) <= 1e-10 * abs(Jo)
    for label, a, b in zip(("lnk", "phi"), g, go):
        err = np.abs(a - b).max() / np.abs(b).max()
        log(f"[adjoint] {name} d/d{label}: max|gpu - oracle| / max|oracle| = {err:.2e} (max {np.abs(b).max():.4e})")
        assert err <= tol


ORACLE_TIGHT = dict(max_iterations=30, min_iteration_index=1, residual_tol=1e-11, linear_rel_tol=1e-11,
                    linear_max_iterations=300)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_adjoint_matches_oracle_adjoint_8x8x3(oracle_mod, precision):
    """GPU gradients against the oracle's FP64 adjoint (transposed direct solves of its AD
    Jacobian) on the GPU's own substep schedule."""
    d = small_deck()
    fn = FimAdjoint(d, checkpoint_interval=2, solver_options=dict(TIGHT, precision=precision))
    J0, dts = fn.forward()
    g = fn.backward()
    fn.close()
    Jo, *go = _oracle_gradient(oracle_mod, d, dts, ORACLE_TIGHT)
    _compare_to_oracle("8x8x3", J0, g, Jo, go, 1e-8)


def test_adjoint_matches_oracle_adjoint_48x48x8(oracle_mod):
    """The 48x48x8 adjoint deck over its first 4 report steps of 5 d."""
    d = decks.black_oil_adjoint(n_steps=4)
    opts = dict(TIGHT, precision=1, ilu_ordering=1)
    fn = FimAdjoint(d, checkpoint_interval=10, solver_options=opts)
    J0, dts = fn.forward()
    g = fn.backward()
    fn.close()
    log(f"[adjoint] 48x48x8 substeps {len(dts)}")
    Jo, *go = _oracle_gradient(oracle_mod, d, dts, ORACLE_TIGHT)
    _compare_to_oracle("48x48x8", J0, g, Jo, go, 1e-8)


@pytest.fixture(scope="module")
def fifty_steps():
    """The full 48x48x8 adjoint deck (50 report steps of 5 d): gradient at interval 10 and a
    simulator that replays the same substeps for finite differences."""
    d = decks.black_oil_adjoint(n_steps=50)
    # The scaled residual of a short substep (the 0.05 d remainders at report ends) has a
    # rounding floor near 5e-10 on this deck: at 1e-10 a replay at perturbed parameters can
    # stall above the tolerance, so the forward converges to 1e-9 (Newton's last step
    # still takes the iterate far below it)
    opts = dict(TIGHT, precision=1, ilu_ordering=1, newton_residual_tol=1e-9, newton_mb_tol=1e-11,
                linear_rel_tol=1e-10)
    fn = FimAdjoint(d, checkpoint_interval=10, solver_options=opts)
    J0, dts = fn.forward()
    g_lnk, g_phi = fn.backward()
    log(f"[adjoint] 48x48x8 x 50: J={J0:.10e} substeps={len(dts)} report={fn.report}")
    fn.close()
    fd_fn = FimAdjoint(d, checkpoint_interval=1000, solver_options=opts, replay=dts)
    yield dict(d=d, g_lnk=g_lnk, g_phi=g_phi, phi0=fn._phi, fd=fd_fn, dirs=directions(d.n_cells, 20, 7))
    fd_fn.close()


@pytest.mark.parametrize("block", range(4))
def test_adjoint_fifty_steps_twenty_directions(fifty_steps, block):
    """20 seeded directions, five per block, central differences at h = 1e-4. Where one is
    off by more than 1e-5 the error is truncation: halving h cuts it about fourfold, and the
    Richardson value (4 D(h/2) - D(h)) / 3 meets the adjoint to 1e-7. A pair whose replay
    does not converge at h is taken at h/2."""
    c = fifty_steps
    fd_fn, phi0 = c["fd"], c["phi0"]

    def D(dk, dp, h):
        try:
            jp, _ = fd_fn.forward(h * dk, phi0 + h * dp)
            jm, _ = fd_fn.forward(-h * dk, phi0 - h * dp)
        except RuntimeError as e:
            log(f"[adjoint] replay at h = {h:g} did not converge: {e}")
            return None
        return (jp - jm) / (2 * h)

    h = 1e-4
    for t, (dk, dp) in enumerate(c["dirs"][5 * block:5 * block + 5]):
        ad = c["g_lnk"] @ dk + c["g_phi"] @ dp
        d1 = D(dk, dp, h)
        if d1 is None:
            # Direction 19 at -h: a 5 d substep stalls at a scaled residual of 3.4e-8, and the
            # oracle's Newton stalls on the same substep at the same point, so the stall is
            # the reference iteration's. The pair at h/2 converges and stands in for it.
            d2 = D(dk, dp, h / 2)
            assert d2 is not None
            rel = abs(d2 - ad) / abs(ad)
            log(f"[adjoint] dir {5 * block + t}: fd(h/2) {d2:.10e} adjoint {ad:.10e} rel {rel:.2e}")
            assert rel <= 1e-5
            continue
        rel = abs(d1 - ad) / abs(ad)
        msg = f"[adjoint] dir {5 * block + t}: fd(h) {d1:.10e} adjoint {ad:.10e} rel {rel:.2e}"
        if rel > 1e-5:
            d2 = D(dk, dp, h / 2)
            rich = (4 * d2 - d1) / 3
            ratio = (d1 - ad) / (d2 - ad)
            rrel = abs(rich - ad) / abs(ad)
            log(msg + f" | fd(h/2) {d2:.10e} ratio {ratio:.2f} richardson rel {rrel:.2e}")
            assert 3.0 < ratio < 5.0
            assert rrel <= 1e-7
        else:
            log(msg)


def test_kept_iterates_give_the_replayed_gradient():
    """Checkpoint interval 1 keeps every converged iterate and skips the forward replay of
    the backward pass; the replay is deterministic, so the gradient is the same bits as
    with an interval of 2."""
    d = small_deck()
    a = FimAdjoint(d, checkpoint_interval=2, solver_options=dict(TIGHT, precision=1))
    b = FimAdjoint(d, checkpoint_interval=1, solver_options=dict(TIGHT, precision=1))
    ja, dta = a.forward()
    jb, dtb = b.forward()
    assert ja == jb and list(dta) == list(dtb)
    ga = a.backward()
    gb = b.backward()
    log(f"[adjoint] replay ms: interval 2 {a.report.get('replay_ms')}, interval 1 {b.report.get('replay_ms')}")
    for x, y in zip(ga, gb):
        np.testing.assert_array_equal(x, y)
    a.close()
    b.close()
