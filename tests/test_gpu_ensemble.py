"""Batched ensemble: members stepped concurrently (one stream each) give bit-identical
fields to the same members stepped one after another."""
import numpy as np
import pytest

from ressim_amd import decks
from ressim_amd.ensemble import build_members, run_ensemble

pytestmark = pytest.mark.gpu

OPTS = dict(newton_residual_tol=1e-6, newton_mb_tol=1e-7, min_newton_iteration_index=0, precision=1, ilu_ordering=1)


def test_concurrent_members_match_sequential():
    mem = [decks.two_phase_member(m, n_steps=3) for m in range(4)]
    seq = build_members(mem, OPTS)
    r1 = run_ensemble(mem, seq, n_threads=1)
    par = build_members(mem, OPTS)
    r2 = run_ensemble(mem, par, n_threads=4)
    print(f"[ensemble] sequential {r1['cell_steps_per_s']:.3e} concurrent {r2['cell_steps_per_s']:.3e} cell-steps/s",
          flush=True)
    assert all(s == 0 for s in r1["status"]) and all(s == 0 for s in r2["status"])
    for a, b in zip(seq, par):
        for f in ("pressure", "sw"):
            np.testing.assert_array_equal(a.get_field(f), b.get_field(f))
        assert a.accepted_dts() == pytest.approx(b.accepted_dts(), rel=0, abs=0)
    for s in seq + par:
        s.close()
