"""Pins of the oracle's CPR internals (oracle/linear.hpp), one component at a time.

Each component is checked against the reference's own unit test of it
(rs/fim/linear/gmres_block_jacobi.rs:2860-3360), restated on the oracle's 3x3-block layout,
and against a textbook construction written here in numpy (ILU(0) on the pattern, block
ILU(0), quasi-IMPES weights as row 0 of the inverse diagonal block, the coarse operator
A_p[i][j] = sum_l w_i[l] A[3i+l][3j], the two-stage apply). A transposed block, a dropped
fill term or a wrong weight normalisation fails at least one of them.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import oracle


# ------------------------------------------------------------------ textbook references
def ilu0_textbook(A):
    """ILU(0) restricted to the nonzero pattern of A (IKJ order): unit-lower L, upper U."""
    A = np.array(A, dtype=float)
    n = len(A)
    pat = A != 0.0
    np.fill_diagonal(pat, True)
    LU = A.copy()
    for i in range(1, n):
        for k in range(i):
            if not pat[i, k]:
                continue
            LU[i, k] /= LU[k, k]
            for j in range(k + 1, n):
                if pat[i, j]:
                    LU[i, j] -= LU[i, k] * LU[k, j]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    return L, U


def block_ilu0_textbook(A, nb):
    """Block ILU(0) on the 3x3 block pattern of A (IKJ order); returns the apply."""
    A = np.array(A, dtype=float)
    blk = {}
    for i in range(nb):
        for j in range(nb):
            b = A[3 * i:3 * i + 3, 3 * j:3 * j + 3]
            if i == j or np.any(b != 0.0):
                blk[i, j] = b.copy()
    for i in range(nb):
        for k in range(i):
            if (i, k) not in blk:
                continue
            blk[i, k] = blk[i, k] @ np.linalg.inv(blk[k, k])
            for j in range(k + 1, nb):
                if (i, j) in blk and (k, j) in blk:
                    blk[i, j] = blk[i, j] - blk[i, k] @ blk[k, j]

    def apply(r):
        y = np.zeros((nb, 3))
        for i in range(nb):
            v = r[3 * i:3 * i + 3].copy()
            for k in range(i):
                if (i, k) in blk:
                    v -= blk[i, k] @ y[k]
            y[i] = v
        x = np.zeros((nb, 3))
        for i in reversed(range(nb)):
            v = y[i].copy()
            for j in range(i + 1, nb):
                if (i, j) in blk:
                    v -= blk[i, j] @ x[j]
            x[i] = np.linalg.solve(blk[i, i], v)
        return x.ravel()

    return apply


def quasi_impes(A, nb):
    """Row 0 of each diagonal block's inverse, scaled by its largest magnitude, and the
    coarse operator with prolongation e0."""
    w = np.zeros((nb, 3))
    for b in range(nb):
        r = np.linalg.inv(A[3 * b:3 * b + 3, 3 * b:3 * b + 3])[0]
        w[b] = r / np.abs(r).max()
    P = np.zeros((nb, nb))
    for i in range(nb):
        for j in range(nb):
            P[i, j] = w[i] @ A[3 * i:3 * i + 3, 3 * j]
    return w, P


def stencil_system(nx, ny, nz, seed, coupling=0.3):
    """A cell system of 3x3 blocks on a 7-point stencil: diagonal blocks dominant, random
    nonsymmetric neighbour blocks, as a Jacobian's sparsity has it."""
    rng = np.random.default_rng(seed)
    nb = nx * ny * nz
    A = np.zeros((3 * nb, 3 * nb))
    for c in range(nb):
        z, r = divmod(c, nx * ny)
        y, x = divmod(r, nx)
        nbrs = [(x - 1, y, z), (x + 1, y, z), (x, y - 1, z), (x, y + 1, z), (x, y, z - 1), (x, y, z + 1)]
        for (i, j, k) in nbrs:
            if 0 <= i < nx and 0 <= j < ny and 0 <= k < nz:
                d = k * nx * ny + j * nx + i
                A[3 * c:3 * c + 3, 3 * d:3 * d + 3] = -coupling * rng.uniform(0.2, 1.0, (3, 3))
        D = rng.uniform(-0.5, 0.5, (3, 3))
        D[np.diag_indices(3)] = rng.uniform(4.0, 8.0, 3)
        D[0, 1:] = rng.uniform(0.5, 2.0, 2)  # pressure couples to the transport unknowns
        A[3 * c:3 * c + 3, 3 * c:3 * c + 3] = D
    return A


# ------------------------------------------------------------------ the reference's tests
def test_dense_coarse_inverse_when_small():
    """:2860 pressure_correction_uses_exact_dense_inverse_when_small: 2 cells, the coarse
    system goes to the exact dense inverse and solves rp = (1, -2) to 1e-10."""
    A = np.zeros((6, 6))
    for (r, c, v) in [(0, 0, 4.0), (0, 3, -1.0), (1, 1, 2.0), (2, 2, 1.0), (3, 0, -1.5), (3, 3, 5.0), (4, 4, 3.0),
                      (5, 5, 7.0)]:
        A[r, c] = v
    out = oracle.cpr_probe(A, [1.0, -2.0], mode=0)
    assert out["dense"]
    res = np.array([1.0, -2.0]) - out["pressure_matrix"] @ out["x"]
    assert np.linalg.norm(res) < 1e-10


def test_transfer_weights_of_the_reference_block():
    """:2930 pressure_transfer_weights_follow_local_schur_elimination on
    [[4,1,0],[2,3,0],[0,0,1]]: restriction (1, -1/3, 0) and prolongation e0. For this block
    the quasi-IMPES weights (row 0 of the inverse, (0.3, -0.1, 0), over its largest
    magnitude) equal that Schur restriction."""
    blk = np.array([[4.0, 1.0, 0.0], [2.0, 3.0, 0.0], [0.0, 0.0, 1.0]])
    out = oracle.cpr_probe(sp.block_diag([blk]).toarray(), [1.0], mode=0)
    np.testing.assert_allclose(out["weights"][0], [1.0, -1.0 / 3.0, 0.0], atol=1e-12)
    # prolongation e0: the coarse operator takes the pressure column only
    np.testing.assert_allclose(out["pressure_matrix"], [[4.0 - 2.0 / 3.0]], rtol=1e-14)


def test_scalar_ilu0_lower_times_upper_recovers_the_matrix():
    """:3032 full_ilu0_lower_times_upper_recovers_original_matrix (4x4, tridiagonal)."""
    A = np.zeros((4, 4))
    for (r, c, v) in [(0, 0, 4.0), (0, 1, -1.0), (1, 0, -1.5), (1, 1, 5.0), (1, 2, -0.5), (2, 1, -0.25), (2, 2, 6.0),
                      (2, 3, -0.75), (3, 2, -0.8), (3, 3, 3.5)]:
        A[r, c] = v
    lu, _ = oracle.scalar_ilu0(A, np.ones(4))
    on = A != 0.0
    assert np.max(np.abs(lu[on] - A[on])) < 1e-10


def test_scalar_ilu0_solves_a_diagonal_system_exactly():
    """:3058 full_ilu0_apply_solves_diagonal_system_exactly."""
    A = np.diag([2.0, 3.0, 4.0, 5.0])
    x = np.array([1.0, -2.0, 3.0, -4.0])
    _, got = oracle.scalar_ilu0(A, A @ x)
    assert np.linalg.norm(got - x) < 1e-12


def _embed2(blocks2):
    """2x2 cell blocks of the reference's tests in 3x3 blocks with a unit third unknown."""
    nb = 1 + max(max(i, j) for (i, j) in blocks2)
    A = np.zeros((3 * nb, 3 * nb))
    for (i, j), b in blocks2.items():
        A[3 * i:3 * i + 2, 3 * j:3 * j + 2] = b
    for i in range(nb):
        A[3 * i + 2, 3 * i + 2] = 1.0
    return A


def test_block_ilu0_exact_for_uncoupled_blocks():
    """:3075 block_ilu0_solves_exactly_when_cell_blocks_are_uncoupled."""
    A = _embed2({(0, 0): [[4.0, 1.0], [2.0, 3.0]], (1, 1): [[5.0, -1.0], [1.0, 6.0]]})
    x = np.array([1.0, -2.0, 0.5, 3.0, -4.0, -0.25])
    got = oracle.block_ilu0_apply(A, A @ x)
    assert np.linalg.norm(got - x) < 1e-10


def test_block_ilu0_reduces_the_residual_with_coupling():
    """:3106 block_ilu0_reduces_residual_with_cell_coupling_and_scalar_tail, without the
    scalar tail row: in the FIM path the well tail is Schur-eliminated before CPR, so the
    block factor never sees one."""
    A = _embed2({(0, 0): [[6.0, 1.0], [1.0, 5.0]], (0, 1): [[0.5, 0.0], [0.0, 0.0]],
                 (1, 0): [[0.5, 0.0], [0.0, 0.0]], (1, 1): [[7.0, 1.0], [1.0, 6.0]]})
    x = np.array([1.0, -2.0, 0.7, 3.0, -1.5, 0.2])
    rhs = A @ x
    z = oracle.block_ilu0_apply(A, rhs)
    assert np.all(np.isfinite(z))
    assert np.linalg.norm(rhs - A @ z) < np.linalg.norm(rhs)


def test_pressure_bicgstab_on_a_small_nonsymmetric_system():
    """:3333 pressure_bicgstab_reduces_residual_on_small_nonsymmetric_system: 20 iterations,
    tolerance 1e-10, residual under 1e-8."""
    A = np.zeros((6, 6))
    for (r, c, v) in [(0, 0, 4.0), (0, 3, 1.0), (1, 1, 2.0), (2, 2, 1.5), (3, 0, -1.5), (3, 3, 5.0), (4, 4, 3.0),
                      (5, 5, 7.0)]:
        A[r, c] = v
    out = oracle.cpr_probe(A, [1.0, -2.0], mode=1, max_iterations=20, tol=1e-10)
    res = np.array([1.0, -2.0]) - out["pressure_matrix"] @ out["x"]
    assert np.linalg.norm(res) < 1e-8
    assert 1 <= out["bicgstab_iterations"] <= 20


# ------------------------------------------------------------------ textbook constructions
def test_scalar_ilu0_matches_textbook_on_a_7_point_matrix():
    """The pressure ILU(0) of a 5x4x3 7-point coarse matrix: the same L U and the same apply
    as ILU(0) by elimination on the pattern."""
    A = stencil_system(5, 4, 3, seed=5)
    _, P = quasi_impes(A, 60)
    L, U = ilu0_textbook(P)
    lu, x = oracle.scalar_ilu0(P, np.arange(60.0) - 30.0)
    np.testing.assert_allclose(lu, L @ U, rtol=0, atol=1e-12 * np.abs(P).max())
    ref = sla.solve_triangular(U, sla.solve_triangular(L, np.arange(60.0) - 30.0, lower=True, unit_diagonal=True))
    np.testing.assert_allclose(x, ref, rtol=1e-12, atol=1e-14)


def test_block_ilu0_matches_textbook_on_a_7_point_system():
    A = stencil_system(4, 3, 2, seed=9)
    r = np.random.default_rng(2).standard_normal(72)
    ref = block_ilu0_textbook(A, 24)(r)
    got = oracle.block_ilu0_apply(A, r)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-13)


def test_quasi_impes_weights_and_coarse_operator():
    A = stencil_system(4, 3, 2, seed=13)
    w, P = quasi_impes(A, 24)
    out = oracle.cpr_probe(A, np.ones(24), mode=0)
    np.testing.assert_allclose(out["weights"], w, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(out["pressure_matrix"], P, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("dims", [(4, 3, 2), (10, 10, 3)])
def test_cpr_two_stage_apply(dims):
    """:471-501: restrict with the weights, coarse solve, prolong into the pressure unknowns,
    correct the residual, block-ILU(0) smooth. 24 cells take the dense coarse inverse (exact);
    300 cells take BiCGStab to 1e-6 of |rp|, so the apply agrees to that order."""
    nx, ny, nz = dims
    nb = nx * ny * nz
    A = stencil_system(nx, ny, nz, seed=17)
    w, P = quasi_impes(A, nb)
    v = np.random.default_rng(4).standard_normal(3 * nb)
    rp = (w * v.reshape(nb, 3)).sum(axis=1)
    xp = np.linalg.solve(P, rp)
    res = np.zeros(3 * nb)
    res[0::3] = xp
    z_ref = res + block_ilu0_textbook(A, nb)(v - A @ res)
    out = oracle.cpr_probe(A, v, mode=2)
    assert out["dense"] == (nb <= 299)
    tol = 1e-11 if nb <= 299 else 1e-4
    np.testing.assert_allclose(out["x"], z_ref, rtol=0, atol=tol * np.abs(z_ref).max())
    if nb > 299:
        c = oracle.cpr_probe(A, rp, mode=0)
        assert np.linalg.norm(rp - P @ c["x"]) <= 1e-6 * np.linalg.norm(rp) * (1 + 1e-9)
        assert 1 <= c["bicgstab_iterations"] <= 50
