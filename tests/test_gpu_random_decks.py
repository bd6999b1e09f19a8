"""Seeded random decks against the oracle: grid sizes 3..8 per axis, random seeds of the
permeability and initial-state recipes, three-phase (with or without gravity, one or five
wells) and two-phase waterfloods (rate-controlled, capillary pressure), in the three
solver modes (FP64 natural ordering, mixed red-black, FP64 red-black). Each deck runs the
oracle's own accepted substeps; pressure (relative) and saturations (absolute) agree to
1e-7 after every substep. Odd nx sends a deck to the general pressure path, even nx to
the fused red-black one."""
import sys

import numpy as np
import pytest

from ressim_amd import ReservoirSimulator, decks

pytestmark = pytest.mark.gpu

MODES = [(0, 0), (1, 1), (0, 1)]  # (precision, ilu_ordering)


def _cases(n=12, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for case in range(n):
        nx, ny, nz = (int(v) for v in rng.integers(3, 9, size=3))
        s = int(rng.integers(1, 10000))
        five = bool(rng.integers(0, 2))
        grav = bool(rng.integers(0, 2))
        kind = int(rng.integers(0, 2))
        out.append((case, nx, ny, nz, s, five, grav, kind, MODES[case % 3]))
    return out


@pytest.mark.parametrize("case,nx,ny,nz,seed,five,grav,kind,mode", _cases())
def test_random_deck_matches_oracle(case, nx, ny, nz, seed, five, grav, kind, mode, oracle_mod):
    if kind == 0:
        d = decks.small_black_oil(nx, ny, nz, n_steps=2, gravity=grav, five_wells=five, seed=seed)
    else:
        d = decks.small_waterflood(nx, ny, nz, n_steps=2, seed=seed)
    print(f"[random] case {case} {d.name} seed={seed} mode={mode}", flush=True)
    sys.stdout.flush()
    opts = dict(max_iterations=30, min_iteration_index=1, residual_tol=1e-10, linear_rel_tol=1e-10,
                linear_max_iterations=200)
    sched = d.build(oracle_mod.OracleSimulator)
    sched.set_newton_options(**opts)
    for dt in d.schedule:
        assert sched.step(dt) == 0, sched.last_warning()
    o = d.build(oracle_mod.OracleSimulator)
    o.set_newton_options(**opts)
    g = d.build(ReservoirSimulator)
    g.set_solver_options(newton_residual_tol=1e-10, linear_rel_tol=1e-10, linear_max_iterations=200,
                         linear_restart=60, max_newton_iterations=30, min_newton_iteration_index=1,
                         precision=mode[0], ilu_ordering=mode[1])
    for dt in sched.accepted_dts():
        assert o.run_substep(dt)[0] == 0
        rc, _ = g.run_timestep(dt)
        assert rc == 0, g.last_warning()
        po = o.get_field("pressure")
        assert np.max(np.abs(g.get_pressures() - po) / np.abs(po)) < 1e-7
        for f in ("sw", "sg"):
            assert np.max(np.abs(o.get_field(f) - g.get_field(f))) < 1e-7, f
