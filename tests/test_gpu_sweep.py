"""The reference ordering's ILU solves in one launch per sweep (rsf_set_ilu_sweep mode 1)
against the same solves as one launch per hyperplane level (mode 0).

Both evaluate FimBlockIlu0Factors::apply (rs/fim/linear/gmres_block_jacobi.rs:190-237)
and apply_pressure_ilu (:634-658) with the same operations in the same order, so whole
Newton substeps must come out bit for bit equal, iteration counts included, for every tile
shape (tiles of 1 to 256 grid lines, tile counts that do and do not divide the grid). At
128x128x64, where the per-level path is too slow to run beside it, the one-launch path's
converged substep must solve the oracle's equations at the benchmark tolerances.
"""
import sys

import numpy as np
import pytest

from ressim_amd import ReservoirSimulator, decks

pytestmark = pytest.mark.gpu

TOL, MB_TOL = 1e-6, 1e-7


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def substep(deck, precision, sweep, lanes_p=256, lanes_b=64, dt=0.5):
    g = deck.build(ReservoirSimulator)
    g.set_solver_options(newton_residual_tol=TOL, newton_mb_tol=MB_TOL, min_newton_iteration_index=0,
                         precision=precision, ilu_ordering=0, linear_restart=30)
    g.set_ilu_sweep(sweep, lanes_p, lanes_b)
    rc, rep = g.run_timestep(dt)
    out = dict(rc=rc, stats=g.stats(), warning=g.last_warning(), it=g.debug_last_iterate() if rc == 0 else None,
               fields={f: g.get_field(f) for f in ("pressure", "sw", "sg", "rs")})
    g.close()
    return out


# decks above the 512-row direct route, so the CPR preconditioner runs
SMALL = {
    "black_oil_12x10x6": lambda: decks.small_black_oil(12, 10, 6, n_steps=1),
    "black_oil_9x7x5_5w": lambda: decks.small_black_oil(9, 7, 5, n_steps=1, five_wells=True),
    "waterflood_11x9x4": lambda: decks.small_waterflood(11, 9, 4, n_steps=1),
}
TILES = [(256, 64), (16, 16), (64, 256), (1, 2), (8, 32)]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("name", list(SMALL))
def test_one_launch_sweeps_equal_level_launches(name, precision):
    d = SMALL[name]()
    ref = substep(d, precision, 0)
    assert ref["rc"] == 0, ref["warning"]
    for lp, lb in TILES:
        got = substep(d, precision, 1, lp, lb)
        assert got["rc"] == 0, got["warning"]
        for k in ("newton_iterations", "linear_iterations", "pressure_iterations"):
            assert got["stats"][k] == ref["stats"][k], (lp, lb, k)
        for f in ("pressure", "sw", "sg", "rs"):
            assert np.array_equal(got["fields"][f], ref["fields"][f]), (lp, lb, f)
        log(f"[sweep] {name} precision={precision} tiles {lp}/{lb}: bit-equal, "
            f"pressure iterations {got['stats']['pressure_iterations']}")


def test_65k_one_launch_sweeps_equal_level_launches():
    d = decks.black_oil_65k(n_steps=1)
    ref = substep(d, 1, 0)
    got = substep(d, 1, 1)
    assert ref["rc"] == 0 and got["rc"] == 0
    assert got["stats"]["pressure_iterations"] == ref["stats"]["pressure_iterations"]
    assert np.array_equal(got["fields"]["pressure"], ref["fields"]["pressure"])
    log(f"[sweep] 64x64x16 mixed: bit-equal, linear ms {got['stats']['linear_ms']:.1f} "
        f"(per level {ref['stats']['linear_ms']:.1f})")


@pytest.mark.parametrize("precision", [1, 0], ids=["mixed", "fp64"])
def test_1m_reference_ordering_solves_oracle_equations(precision, oracle_mod):
    d = decks.black_oil_1m(n_steps=1)
    got = substep(d, precision, 1, dt=0.25)
    assert got["rc"] == 0, got["warning"]
    o = d.build(oracle_mod.OracleSimulator)
    chk = o.residual_check(0.25, *got["it"])
    fam, mb = chk["families"], chk["mb"]
    log(f"[sweep] 128x128x64 natural order precision={precision}: oracle families "
        f"{np.array2string(fam, precision=3)} mb {np.array2string(mb, precision=3)}")
    assert fam[:3].max() <= TOL * (1 + 1e-6)
    assert (mb <= MB_TOL * (1 + 1e-6)).all()
