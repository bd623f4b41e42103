"""Structure of the synthetic cell batches (seeded, deterministic)."""
import numpy as np

from python_amd import generators as G


def test_layout_30():
    blocks = G.element_layout(30)
    assert [m for (_, m, _) in blocks] == [2, 3, 7, 8, 9, 1]
    assert [ab for (_, _, ab) in blocks][:5] == [1.0, 0.1, 3e-4, 1e-4, 7e-4]
    for n in (1, 2, 12, 64, 96, 128, 200, 256, 500):
        assert sum(m for (_, m, _) in G.element_layout(n)) == n


def test_ion_structure():
    rng = np.random.default_rng(0)
    A, b = G.ion_matrices(rng, 5, 30)
    cons = [s for (s, m, ab) in G.element_layout(30)]
    for c in range(5):
        colsum = A[c].sum(axis=0)
        # every column sums to zero except through the replaced conservation rows
        other = np.setdiff1d(np.arange(30), cons)
        red = A[c][other].sum(axis=0) + sum(A[c][s] * 0 for s in cons)
        assert np.all(b[c][other] == 0)
        assert np.all(b[c][cons] > 0)
        for (s, m, ab) in G.element_layout(30):
            assert np.all(A[c][s, s:s + m] == 1.0) and A[c][s].sum() == m
        off = A[c] - np.diag(np.diag(A[c]))
        assert np.all(off[other] >= 0)


def test_deterministic_and_counts():
    b1 = G.uneven(scale=0.01)
    b2 = G.uneven(scale=0.01)
    assert np.array_equal(b1.A, b2.A) and np.array_equal(b1.n, b2.n)
    full_counts = dict(zip(G.UNEVEN_SIZES, G.UNEVEN_COUNTS))
    assert sum(full_counts.values()) == 10000
    p = G.plain_small(ncells=20)
    assert p.ncells == 20 and set(p.n.tolist()) == {30}


def test_macro_q_substochastic():
    N = G.macro_q(np.random.default_rng(1), 2, 64)
    for c in range(2):
        Q = np.eye(64) - N[c]
        rs = Q.sum(axis=1)
        assert np.all(np.diag(Q) == 0) and np.all(Q >= 0)
        assert np.all(rs <= 1 - 1e-6 + 1e-12)


def test_te_derivative_matches_finite_difference():
    rng = np.random.default_rng(5)
    d = G.ion_matrices_te(rng, 4, 20)
    T = d["te"]
    h = 1e-6 * T
    Ap, _, _ = G.te_assemble(20, d["blocks"], d["params"], T + h)
    Am, _, _ = G.te_assemble(20, d["blocks"], d["params"], T - h)
    fd = (Ap - Am) / (2 * h)[:, None, None]
    scale = np.abs(d["dA"]).max(axis=(1, 2))[:, None, None] + 1e-300
    assert np.all(np.abs(fd - d["dA"]) <= 1e-5 * scale)


def test_macro_pops_structure():
    """Level-population systems (source/macro_gov.c:366-440): below row 0 every column of the
    rate matrix keeps only the transitions into other levels, so columns sum to minus the
    removed row-0 entries' rates; the oracle's populations are positive and sum to 1."""
    import oracle
    from python_amd.generators import macro_pops_matrices
    A, b = macro_pops_matrices(np.random.default_rng(4), 4, 15)
    for M, rhs in zip(A, b):
        assert np.all(M[0] == 1.0) and rhs[0] == 1.0 and not rhs[1:].any()
        offd = M[1:] - np.diag(np.diag(M))[1:]
        assert np.all(offd >= 0.0)
        assert np.all(np.diag(M)[1:] < 0.0)
        code, x = oracle.solve_matrix(M, rhs)
        assert code == 0
        assert np.all(x > 0.0) and abs(x.sum() - 1.0) <= 1e-12
