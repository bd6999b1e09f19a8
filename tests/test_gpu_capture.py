"""Captured linear systems through the GPU linear stack (the reference's offline
solver-lab methodology, rs/fim/linear/solver_lab.rs, applied to CPU-vs-GPU parity).

- A full FIM system from the oracle (cells plus the tail of 5 wells: BHPs and
  perforation rates), written as fim-capture-v2 and replayed on the GPU: the answer
  must match a direct sparse solve of the same captured system.
- The library's own captures of the systems it solves (after the wells' Schur
  elimination): they parse as fim-capture-v2 with the layout "N 3 0 3N", every value
  is printed as Rust's `{:e}` prints it, and replaying them reproduces the direct solve.
"""
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from ressim_amd import ReservoirSimulator, decks
from ressim_amd import capture as cap

pytestmark = pytest.mark.gpu

TIGHT = dict(linear_rel_tol=1e-11, linear_max_iterations=200, linear_restart=60)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", [(0, 0), (1, 1), (0, 1)], ids=["fp64_natural", "mixed_redblack", "fp64_redblack"])
def test_oracle_capture_with_well_tail_replays_on_gpu(tmp_path, oracle_mod, mode):
    d = decks.small_black_oil(8, 7, 3, n_steps=1, five_wells=True)
    o = d.build(oracle_mod.OracleSimulator)
    cs, rg = o.initial_state()
    bhp, q = o.get_well_state()
    R, (ptr, col, val) = o.assemble(d.schedule[0], cs, rg, bhp, q)
    J = sp.csr_matrix((val, col, ptr), shape=(len(R), len(R)))
    n = d.n_cells
    cap.write_capture(tmp_path, 0, cap.CapturedSystem(cap.CaptureMetadata(0, "sequence", "none", 0),
                                                      cap.BlockLayout(n, 3, len(bhp), 3 * n + len(bhp)), J, -R))
    s = cap.load_captures(tmp_path)[0]
    ref = spla.spsolve(sp.csc_matrix(s.jacobian), s.rhs)
    g = d.build(ReservoirSimulator)
    g.set_solver_options(precision=mode[0], ilu_ordering=mode[1], **TIGHT)
    x, rep = g.solve_captured_system(s.jacobian, s.rhs)
    rel = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    res = np.linalg.norm(s.jacobian @ x - s.rhs) / np.linalg.norm(s.rhs)
    log(f"[capture] oracle system rows={len(R)} tail={len(R) - 3 * n} mode={mode} its={rep['iterations']} "
        f"rel err {rel:.3e} residual {res:.3e}")
    assert rep["converged"] == 1
    assert rel < 1e-8 and res < 1e-9
    # the handle is still a simulator afterwards: its wells are rebuilt at the next step
    assert g.step(d.schedule[0]) == 0, g.last_warning()
    g.close()


def test_gpu_captures_parse_and_replay(tmp_path):
    d = decks.small_black_oil(6, 5, 3, n_steps=1)
    g = d.build(ReservoirSimulator)
    g.set_solver_options(precision=1, ilu_ordering=1)
    g.set_capture_dir(str(tmp_path), 1)
    assert g.step(d.schedule[0]) == 0, g.last_warning()
    g.close()
    names = sorted(os.listdir(tmp_path))
    log(f"[capture] {len(names)} systems written")
    assert names and names[0] == "fim_capture_00000.txt"
    systems = cap.load_captures(tmp_path)
    n = d.n_cells
    for s in systems:
        assert s.layout == cap.BlockLayout(n, 3, 0, 3 * n)
        assert s.jacobian.shape == (3 * n, 3 * n) and s.metadata.failure_reason == "sequence"
    # every number in the files is printed as Rust's {:e} prints it
    text = (tmp_path / names[0]).read_text().splitlines()
    start = text.index("rhs") + 1
    trip = text.index("triplets")
    for line in text[start:trip]:
        assert cap.format_f64(float(line)) == line
    for line in text[trip + 1:]:
        v = line.split()[2]
        assert cap.format_f64(float(v)) == v
    # replaying the first and the last system reproduces the direct solve
    for s in (systems[0], systems[-1]):
        ref = spla.spsolve(sp.csc_matrix(s.jacobian), s.rhs)
        h = d.build(ReservoirSimulator)
        h.set_solver_options(precision=0, ilu_ordering=0, **TIGHT)
        x, rep = h.solve_captured_system(s.jacobian, s.rhs)
        h.close()
        rel = np.linalg.norm(x - ref) / np.linalg.norm(ref)
        log(f"[capture] replay newton_iteration={s.metadata.newton_iteration} its={rep['iterations']} rel {rel:.3e}")
        assert rep["converged"] == 1 and rel < 1e-8


def test_replay_carries_couplings_off_the_stencil(oracle_mod):
    """A coupling between cells that are not stencil neighbours (cells 0 and 10 of a 6x5
    layer) travels in the extras, like a well clique; and a system with fewer rows than
    three per cell is refused."""
    d = decks.small_black_oil(6, 5, 3, n_steps=1)
    n = d.n_cells
    o = d.build(oracle_mod.OracleSimulator)
    cs, rg = o.initial_state()
    bhp, q = o.get_well_state()
    R, (ptr, col, val) = o.assemble(d.schedule[0], cs, rg, bhp, q)
    J = sp.csr_matrix((val, col, ptr), shape=(len(R), len(R))).tolil()
    scale = abs(J[0, 0])
    for a in range(3):
        J[a, 30 + a] += 0.1 * scale
        J[30 + a, a] += 0.1 * scale
    J = sp.csr_matrix(J)
    ref = spla.spsolve(sp.csc_matrix(J), -R)
    g = d.build(ReservoirSimulator)
    g.set_solver_options(precision=0, ilu_ordering=1, **TIGHT)
    x, rep = g.solve_captured_system(J, -R)
    rel = np.linalg.norm(x - ref) / np.linalg.norm(ref)
    log(f"[capture] off-stencil coupling its={rep['iterations']} rel {rel:.3e}")
    assert rep["converged"] == 1 and rel < 1e-8
    with pytest.raises(ValueError, match="fewer rows"):
        g.solve_captured_system(sp.identity(3 * n - 3, format="csr"), np.ones(3 * n - 3))
    g.close()
