"""The fim-capture-v2/v3 reader and writer (rs/fim/linear/capture.rs:149-640),
checked as the reference's own tests check them (capture.rs:647-843) on the same sample
system, plus the exact text the reference's writer produces for it and a real FIM
system from the oracle read back bit for bit."""
import numpy as np
import pytest
import scipy.sparse as sp

from ressim_amd import capture as cap
from ressim_amd import decks


def sample_system():
    # capture.rs:649-658
    J = sp.csr_matrix(([4.0, -1.5e-8, 2.0, 1e12, 8.0], ([0, 0, 1, 2, 2], [0, 2, 1, 0, 2])), shape=(3, 3))
    return J, np.array([1.0, -2.5, 3e-4])


def sample_flow(J):
    # capture.rs:661-676
    return cap.FlowLifecycle(cap.FLOW_SOURCE_TAG, cap.FLOW_SOURCE_COMMIT, cap.DUNE_ISTL_VERSION,
                             cap.TRUE_IMPES_PRESSURE_SCALE_BAR, np.array([np.diag([1.0, 2.0, 3.0])]),
                             np.array([[1.0, 0.0, 0.0]]), 3, J.copy(), sp.csr_matrix((3, 0)), sp.csr_matrix((0, 3)),
                             sp.csr_matrix((0, 0)))


LAYOUT = cap.BlockLayout(1, 3, 0, 3)

# what the reference's writer produces for capture.rs:679-716 (Rust `{:e}` values)
SAMPLE_TEXT = """fim-capture-v2
newton_iteration 7
failure_reason dead-state
dominant_family water
dominant_item_index 405
layout 1 3 0 3
equation_scaling 1
water 1 1e1
oil_component 1 1e1
gas_component 1 1e1
well_constraint 0
perforation_flow 0
rows 3
cols 3
nnz 5
rhs
1e0
-2.5e0
3e-4
triplets
0 0 4e0
0 2 -1.5e-8
1 1 2e0
2 0 1e12
2 2 8e0
"""


def sample_capture():
    J, rhs = sample_system()
    md = cap.CaptureMetadata(7, "dead-state", "water", 405)
    sc = cap.EquationScaling(np.array([10.0]), np.array([10.0]), np.array([10.0]), np.array([]), np.array([]))
    return cap.CapturedSystem(md, LAYOUT, J, rhs, sc)


def test_format_matches_rust_exponent_notation():
    cases = [(1.0, "1e0"), (4.0, "4e0"), (-2.5, "-2.5e0"), (1.5e-8, "1.5e-8"), (1e12, "1e12"), (3e-4, "3e-4"),
             (0.0, "0e0"), (-0.0, "-0e0"), (1234.5, "1.2345e3"), (0.1 + 0.2, "3.0000000000000004e-1"),
             (5e-324, "5e-324"), (1.7976931348623157e308, "1.7976931348623157e308"), (float("inf"), "inf"),
             (float("nan"), "NaN")]
    for v, s in cases:
        assert cap.format_f64(v) == s, (v, cap.format_f64(v))
    rng = np.random.default_rng(3)
    for v in np.concatenate([rng.standard_normal(500) * 10.0 ** rng.integers(-300, 300, 500),
                             rng.standard_normal(500)]):
        assert float(cap.format_f64(v)) == v  # shortest digits still round-trip


def test_writer_produces_the_reference_text(tmp_path):
    path = cap.write_capture(tmp_path, 3, sample_capture())
    assert path.endswith("fim_capture_00003.txt")
    assert open(path).read() == SAMPLE_TEXT


def test_capture_round_trips_metadata_layout_rhs_and_triplets(tmp_path):
    # capture.rs:678-716
    c = sample_capture()
    cap.write_capture(tmp_path, 3, c)
    loaded = cap.load_captures(tmp_path)
    assert len(loaded) == 1
    s = loaded[0]
    assert s.metadata == c.metadata and s.layout == c.layout
    np.testing.assert_array_equal(s.rhs, c.rhs)
    np.testing.assert_array_equal(s.jacobian.toarray(), c.jacobian.toarray())
    assert s.equation_scaling == c.equation_scaling and s.flow_lifecycle is None


def test_capture_round_trips_none_equation_scaling(tmp_path):
    # capture.rs:718-741
    J, rhs = sample_system()
    cap.write_capture(tmp_path, 0, cap.CapturedSystem(cap.CaptureMetadata(0, "x", "oil", 1), None, J, rhs))
    s = cap.load_captures(tmp_path)[0]
    assert s.equation_scaling is None and s.flow_lifecycle is None and s.layout is None


def test_flow_lifecycle_capture_round_trips_and_validates(tmp_path):
    # capture.rs:743-775
    J, rhs = sample_system()
    flow = sample_flow(J)
    cap.write_capture(tmp_path, 0, cap.CapturedSystem(cap.CaptureMetadata(0, "y2d6a-proof", "water", 0), LAYOUT, J,
                                                      rhs, None, flow))
    text = (tmp_path / "fim_capture_00000.txt").read_text()
    assert text.startswith("fim-capture-v3\n")
    s = cap.load_captures(tmp_path)[0]
    np.testing.assert_array_equal(s.flow_lifecycle.storage_blocks, flow.storage_blocks)
    np.testing.assert_array_equal(s.flow_lifecycle.true_impes_weights, flow.true_impes_weights)
    np.testing.assert_array_equal(s.flow_lifecycle.j_rr.toarray(), J.toarray())


def test_flow_lifecycle_capture_rejects_mismatched_true_impes_weight(tmp_path):
    # capture.rs:777-801
    J, rhs = sample_system()
    flow = sample_flow(J)
    flow.true_impes_weights[0][0] = 0.5
    cap.write_capture(tmp_path, 0, cap.CapturedSystem(cap.CaptureMetadata(0, "y2d6a-reject", "water", 0), LAYOUT, J,
                                                      rhs, None, flow))
    with pytest.raises(cap.CaptureError):
        cap.load_captures(tmp_path)


def test_true_impes_weight_solves_the_scaled_transposed_storage_block():
    # S^T w = e0 with the pressure column scaled by 50 bar, max-abs normalised
    rng = np.random.default_rng(5)
    S = rng.standard_normal((3, 3)) + 3 * np.eye(3)
    w = cap.true_impes_weight(S)
    Ss = S.copy()
    Ss[:, 0] *= 50.0
    r = Ss.T @ w
    assert abs(np.max(np.abs(w)) - 1.0) < 1e-15
    np.testing.assert_allclose(r / r[0], [1.0, 0.0, 0.0], atol=1e-13)
    assert cap.true_impes_weight(np.diag([1.0, 2.0, 3.0])).tolist() == [1.0, 0.0, 0.0]


def test_parse_rejects_wrong_header_and_truncated_triplets():
    # capture.rs:803-842
    with pytest.raises(cap.CaptureError):
        cap.parse_capture("not-a-capture")
    J, rhs = sample_system()
    text = cap.capture_text(cap.CapturedSystem(cap.CaptureMetadata(0, "max-iters", "oil", 1), None, J, rhs))
    truncated = "\n".join(text.splitlines()[:-1])
    with pytest.raises(cap.CaptureError):
        cap.parse_capture(truncated)


def test_oracle_system_reads_back_bit_for_bit(tmp_path, oracle_mod):
    """A real FIM system (the oracle's Jacobian with its wells' tail, 5 wells) written and
    read back: every value and the sparsity pattern identical."""
    d = decks.small_black_oil(7, 7, 3, n_steps=1, five_wells=True)
    o = d.build(oracle_mod.OracleSimulator)
    cs, rg = o.initial_state()
    bhp, q = o.get_well_state()
    R, (ptr, col, val) = o.assemble(d.schedule[0], cs, rg, bhp, q)
    J = sp.csr_matrix((val, col, ptr), shape=(len(R), len(R)))
    n = d.n_cells
    c = cap.CapturedSystem(cap.CaptureMetadata(0, "sequence", "none", 0),
                           cap.BlockLayout(n, 3, len(bhp), 3 * n + len(bhp)), J, -R)
    cap.write_capture(tmp_path, 0, c)
    s = cap.load_captures(tmp_path)[0]
    assert s.layout == c.layout
    assert s.rhs.tobytes() == c.rhs.tobytes()
    a, b = sp.csr_matrix(s.jacobian), sp.csr_matrix(J)
    a.sort_indices(), b.sort_indices()
    assert np.array_equal(a.indptr, b.indptr) and np.array_equal(a.indices, b.indices)
    assert a.data.tobytes() == b.data.tobytes()
