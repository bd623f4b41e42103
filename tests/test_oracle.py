"""Pins the CPU oracle (oracle/hipla_ref.c) before anything is compared against it.

Each check is chosen so a plausible restatement mistake fails it: a dropped update term, a
wrong pivot rule, a transposed operand, a wrong permutation direction, a sign flip.  Sources:
* the reference's own fixtures (source/tests/test_data/matrix, checked as in
  source/tests/tests/test_matrix.c:192-198,250-256 at EPSILON = 1e-6, python.h:98);
* GSL 2.6's known answers (linalg/test.c:359-383; Hilbert / Vandermonde constructors :136-205,
  right-hand side b_i = i + 1) and its invariants (linalg/test_lu.c:31-288: P L U = A, A inv(A) = I);
* exact rational arithmetic (fractions) on small integer systems;
* closed forms: the Neumann series of N = I - Q, the adjoint identity d(w.x)/dA = -lambda x^T.
"""
from fractions import Fraction

import numpy as np
import pytest

from tests.conftest import load_fixture


# ------------------------------------------------------------------------------- fixtures
def test_fixture_small_matrix(oracle_lib):
    n, A = load_fixture("small_matrix", "A.txt")
    _, b = load_fixture("small_matrix", "b.txt")
    _, x = load_fixture("small_matrix", "x.txt")
    code, xs = oracle_lib.solve_matrix(A.reshape(n, n), b)
    assert code == 0
    assert np.array_equal(xs, x)            # exactly [2, -1]


def test_fixture_matrix_ion(oracle_lib):
    n, A = load_fixture("matrix_ion", "A.txt")
    _, b = load_fixture("matrix_ion", "b.txt")
    _, x = load_fixture("matrix_ion", "x.txt")
    code, xs = oracle_lib.solve_matrix(A.reshape(n, n), b)
    assert code == 0
    assert np.abs(xs - x).max() <= 1e-6
    # planted mistakes: a transposed operand and a dropped conservation row both fail the check
    _, xt = oracle_lib.solve_matrix(A.reshape(n, n).T.copy(), b)
    assert not np.abs(xt - x).max() <= 1e-6
    A2 = A.reshape(n, n).copy()
    A2[2, :] = 0.0
    A2[2, 2] = 1.0
    _, xd = oracle_lib.solve_matrix(A2, b)
    assert not np.abs(xd - x).max() <= 1e-6


def test_fixture_inverse_small(oracle_lib):
    n, M = load_fixture("inverse_small", "matrix.txt")
    _, I = load_fixture("inverse_small", "inverse.txt")
    code, inv = oracle_lib.invert_matrix(M.reshape(n, n))
    assert code == 0
    assert np.array_equal(inv.ravel(), I)


def test_fixture_inverse_macro(oracle_lib):
    n, M = load_fixture("inverse_macro", "matrix.txt")
    _, I = load_fixture("inverse_macro", "inverse.txt")
    code, inv = oracle_lib.invert_matrix(M.reshape(n, n))
    assert code == 0
    # the reference test compares only the first n entries (test_matrix.c:198): row 0
    assert np.abs(inv.ravel()[:n] - I[:n]).max() <= 1e-6
    # full matrix: the fixture's matrix.txt is itself rounded to 7 digits and kappa ~ 9e3, so the
    # whole inverse agrees to ~1e-4 relative, not to the printed digits
    assert np.abs(inv.ravel() - I).max() / np.abs(I).max() < 2e-4
    # row 0 of this N^-1 is e_0, so the row-0 check pins little: a planted row/column mistake
    # (transposed inverse) passes it but fails the full-matrix comparison
    assert np.abs(inv.T.ravel()[:n] - I[:n]).max() <= 1e-6
    assert np.abs(inv.T.ravel() - I).max() / np.abs(I).max() > 1e-2


# ------------------------------------------------------------------------- GSL known answers
def _hilbert(n):
    i = np.arange(n)
    return 1.0 / (i[:, None] + i[None, :] + 1.0)


def _vander(n):
    i = np.arange(n, dtype=float)
    return (i[:, None] + 1.0) ** (n - np.arange(n)[None, :] - 1.0)


HILB = {2: [-8.0, 18.0], 3: [27.0, -192.0, 210.0], 4: [-64.0, 900.0, -2520.0, 1820.0]}


@pytest.mark.parametrize("n", [2, 3, 4])
def test_gsl_hilbert_known_answers(oracle_lib, n):
    b = np.arange(1, n + 1, dtype=float)
    st, lu_x = oracle_lib.solve_matrix(_hilbert(n), b)
    assert st == 0
    ref = np.array(HILB[n])
    # GSL test_LU_solve_eps tolerance for Hilbert: 4096 n eps relative (linalg/test_lu.c:218-220)
    assert np.all(np.abs(lu_x - ref) <= 4096 * n * np.finfo(float).eps * np.abs(ref))


@pytest.mark.parametrize("n", [2, 3, 4, 12])
def test_gsl_vandermonde_known_answers(oracle_lib, n):
    b = np.arange(1, n + 1, dtype=float)
    ref = np.zeros(n)
    ref[n - 2] = 1.0                  # vander*_solution, linalg/test.c:372-377
    st, x = oracle_lib.solve_matrix(_vander(n), b)
    assert st == 0
    assert np.abs(x - ref).max() <= 1e-6


# ------------------------------------------------------------------------- exact arithmetic
def _exact_solve(A, b):
    n = len(b)
    M = [[Fraction(int(A[i, j])) for j in range(n)] + [Fraction(int(b[i]))] for i in range(n)]
    for k in range(n):
        p = next(i for i in range(k, n) if M[i][k] != 0)
        M[k], M[p] = M[p], M[k]
        for i in range(k + 1, n):
            f = M[i][k] / M[k][k]
            for j in range(k, n + 1):
                M[i][j] -= f * M[k][j]
    x = [Fraction(0)] * n
    for i in range(n - 1, -1, -1):
        x[i] = (M[i][n] - sum(M[i][j] * x[j] for j in range(i + 1, n))) / M[i][i]
    return np.array([float(v) for v in x])


def test_exact_rational_small(oracle_lib):
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 4, 5, 6, 8):
        for _ in range(10):
            A = rng.integers(-9, 10, (n, n)).astype(float)
            if abs(np.linalg.det(A)) < 0.5:
                continue
            b = rng.integers(-9, 10, n).astype(float)
            xe = _exact_solve(A, b)
            code, x = oracle_lib.solve_matrix(A, b)
            assert code in (0, 2, 3)
            scale = max(1.0, np.abs(xe).max())
            assert np.abs(x - xe).max() <= 1e-10 * scale * np.linalg.cond(A)


# ------------------------------------------------------------------------------- invariants
def test_plu_equals_a_and_recursion_matches_right_looking(oracle_lib):
    rng = np.random.default_rng(7)
    for n in (1, 2, 7, 16, 24, 25, 40, 64, 100, 131, 200):
        for kind in range(3):
            A = rng.standard_normal((n, n))
            if kind == 1:
                A = np.where(rng.random((n, n)) < 0.75, 0.0, A)
                A[np.arange(n), np.arange(n)] += 1.0
            if kind == 2:
                A = np.round(A * 2) / 2
                A[np.arange(n), np.arange(n)] += 3.0
            LU, perm, ipiv, sgn = oracle_lib.lu_decomp(A)
            L = np.tril(LU, -1) + np.eye(n)
            U = np.triu(LU)
            PA = A[perm]                              # gsl_permutation: row i of PA = A[perm[i]]
            err = np.abs(PA - L @ U).max() / max(np.abs(A).max(), 1e-300)
            assert err <= 4096 * n * np.finfo(float).eps
            # partial pivoting: |L| <= 1
            assert np.abs(np.tril(LU, -1)).max() <= 1.0
            # the right-looking restatement (the GPU kernels' order) gives the same values
            LU2, perm2, _, sgn2 = oracle_lib.lu_decomp(A, rightlooking=True)
            assert np.array_equal(LU, LU2) and np.array_equal(perm, perm2) and sgn == sgn2
            assert sgn == round(np.linalg.det(PA @ np.linalg.inv(A))) if n <= 40 else True


def test_pivot_rule_first_max(oracle_lib):
    # ties: GSL takes the first maximal |a| in the current row order; NaN never wins
    A = np.array([[1.0, 2.0], [-1.0, 3.0]])
    LU, perm, ipiv, _ = oracle_lib.lu_decomp(A)
    assert ipiv[0] == 0
    A = np.array([[0.0, 1.0], [0.0, 2.0]])      # all-zero column: pivot stays in place
    LU, perm, ipiv, _ = oracle_lib.lu_decomp(A)
    assert ipiv[0] == 0
    A = np.array([[np.nan, 1.0], [2.0, 1.0], [3.0, 5.0]])[:2, :2]
    LU, perm, ipiv, _ = oracle_lib.lu_decomp(A)
    assert ipiv[0] == 1


def test_residual_codes(oracle_lib):
    A = np.eye(3)
    b = np.array([1.0, 0.0, 2.0])
    assert oracle_lib.residual_code(A, b, b.copy()) == 0
    assert oracle_lib.residual_code(A, b, np.array([1.01, 0.0, 2.0])) == 2        # relative, b > 0
    assert oracle_lib.residual_code(A, b, np.array([1.0, 1e-3, 2.0])) == 3        # absolute, b == 0
    # only the first failing row sets the code
    assert oracle_lib.residual_code(A, b, np.array([1.0, 1e-3, 3.0])) == 3
    assert oracle_lib.residual_code(A, b, np.array([1.5, 1e-3, 3.0])) == 2
    # singular
    code, x = oracle_lib.solve_matrix(np.array([[1.0, 0.0], [2.0, 0.0]]), np.array([1.0, 2.0]))
    assert code == 4


def test_inverse_invariants_and_neumann(oracle_lib):
    rng = np.random.default_rng(9)
    for n in (3, 24, 25, 50, 77, 130):
        A = rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)
        code, inv = oracle_lib.invert_matrix(A)
        assert code == 0
        assert np.abs(A @ inv - np.eye(n)).max() <= 1e-12 * np.linalg.cond(A)
        assert np.abs(inv - np.linalg.inv(A)).max() <= 1e-12 * np.linalg.cond(A) * np.abs(inv).max()
    # N = I - Q with row sums of Q <= 1 - eps: N^-1 = sum_k Q^k (entrywise >= 0)
    from python_amd.generators import macro_q
    N = macro_q(np.random.default_rng(4), 1, 40, eps_range=(0.2, 0.5))[0]
    Q = np.eye(40) - N
    S = np.eye(40)
    term = np.eye(40)
    for _ in range(400):
        term = term @ Q
        S += term
    code, inv = oracle_lib.invert_matrix(N)
    assert code == 0
    assert inv.min() >= -1e-15
    assert np.abs(inv - S).max() <= 1e-10 * np.abs(S).max()
    # singular -> GSL_EDOM (1), as gsl_linalg_LU_invx
    assert oracle_lib.invert_matrix(np.array([[1.0, 2.0], [2.0, 4.0]]))[0] == 1


def test_transpose_solve_and_adjoint_identity(oracle_lib):
    rng = np.random.default_rng(13)
    n = 30
    A = rng.standard_normal((n, n)) + 3 * np.eye(n)
    b = rng.standard_normal(n)
    w = rng.random(n)
    LU, perm, _, _ = oracle_lib.lu_decomp(A)
    st, x = oracle_lib.lu_solve(LU, perm, b)
    st, lam = oracle_lib.lu_solve_transpose(LU, perm, w)
    assert st == 0
    assert np.abs(A.T @ lam - w).max() <= 1e-12 * np.abs(w).max() * np.linalg.cond(A)
    # d(w.x)/dA_ij = -lambda_i x_j, checked by central differences
    G = -np.outer(lam, x)
    for (i, j) in [(0, 0), (3, 7), (n - 1, 2), (10, 10)]:
        h = 1e-6 * max(1.0, abs(A[i, j]))
        Ap, Am = A.copy(), A.copy()
        Ap[i, j] += h
        Am[i, j] -= h
        fd = (w @ np.linalg.solve(Ap, b) - w @ np.linalg.solve(Am, b)) / (2 * h)
        assert abs(fd - G[i, j]) <= 1e-6 * max(1.0, abs(G[i, j]))
    # planted mistake: solving A lam = w (forgetting the transpose) breaks the identity
    st, wrong = oracle_lib.lu_solve(LU, perm, w)
    assert np.abs(A.T @ wrong - w).max() > 1e-3


def test_dd_refinement_is_closer_on_ill_conditioned(oracle_lib):
    rng = np.random.default_rng(21)
    n = 12
    A = _hilbert(n)
    xt = rng.standard_normal(n)
    # b computed exactly in rationals then rounded: the true solution of the rounded system is
    # close to xt up to kappa * eps(b); the dd-refined answer must not be worse than plain LU
    b = np.array([float(sum(Fraction(A[i, j]) * Fraction(xt[j]) for j in range(n))) for i in range(n)])
    _, x64 = oracle_lib.solve_matrix(A, b)
    st, xdd = oracle_lib.solve_dd_refined(A, b, 6)
    r64 = np.abs(np.array([float(Fraction(b[i]) - sum(Fraction(A[i, j]) * Fraction(x64[j]) for j in range(n))) for i in range(n)])).max()
    rdd = np.abs(np.array([float(Fraction(b[i]) - sum(Fraction(A[i, j]) * Fraction(xdd[j]) for j in range(n))) for i in range(n)])).max()
    assert rdd <= r64


def test_batched_oracle_matches_single(oracle_lib):
    from python_amd.generators import uneven
    batch = uneven(scale=0.005)
    X, codes = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=4)
    for c in range(batch.ncells):
        a, b = batch.cell(c)
        code, x = oracle_lib.solve_matrix(a, b)
        assert code == codes[c]
        assert np.array_equal(x, X[batch.offV[c]:batch.offV[c] + len(b)])


def test_batched_inverse_matches_single(oracle_lib):
    """ref_invert_batch (the C4 CPU baseline) equals invert_matrix cell by cell, bit for bit,
    on a ragged batch with a singular cell: its output slot stays as the caller left it (zeros),
    as cpu_invert_matrix leaves inverse_out untouched on error (source/matrix_cpu.c:242-245)."""
    from python_amd.generators import macro_q, pack
    rng = np.random.default_rng(11)
    mats = [macro_q(rng, 1, n)[0] for n in (7, 33, 5, 64, 20)]
    mats.insert(2, np.array([[1.0, 2.0], [2.0, 4.0]]))
    batch = pack(mats)
    Inv, codes = oracle_lib.invert_batch(batch.n, batch.offA, batch.A, nthreads=3)
    for c in range(batch.ncells):
        a, _ = batch.cell(c)
        n = a.shape[0]
        code, inv = oracle_lib.invert_matrix(a)
        assert code == codes[c]
        got = Inv[batch.offA[c]:batch.offA[c] + n * n].reshape(n, n)
        if code == 0:
            assert np.array_equal(got, inv)
        else:
            assert code == 1 and not got.any()
    # a row-major / column-major mix-up in the batch loop would transpose a non-symmetric cell
    a, _ = batch.cell(4)
    assert a.shape == (64, 64)
    assert not np.array_equal(Inv[batch.offA[4]:batch.offA[4] + 64 * 64].reshape(64, 64).T,
                              oracle_lib.invert_matrix(a)[1])
