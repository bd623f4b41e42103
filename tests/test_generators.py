"""Structure of the synthetic cell batches (seeded, deterministic)."""
import numpy as np

from python_amd import generators as G


def test_layout_30():
    blocks = G.element_layout(30)
    assert [m for (_, m, _) in blocks] == [2, 3, 7, 8, 9, 1]
    assert [ab for (_, _, ab) in blocks][:5] == [1.0, 0.1, 3e-4, 1e-4, 7e-4]
    for n in (1, 2, 12, 64, 96, 128, 200, 256, 500):
        assert sum(m for (_, m, _) in G.element_layout(n)) == n


def _check_rate_structure(A, b, n, auger):
    blocks = G.element_layout(n)
    cons = [s for (s, m, ab) in blocks]
    other = np.setdiff1d(np.arange(n), cons)
    elem = np.zeros(n, dtype=int)
    for e, (s, m, ab) in enumerate(blocks):
        elem[s:s + m] = e
    sites = {col: nj for (col, _r0, nj) in G.auger_sites(blocks)}
    for c in range(A.shape[0]):
        M = A[c]
        assert np.all(b[c][other] == 0)
        assert np.array_equal(b[c][cons], [ab for (_, _, ab) in blocks])
        for (s, m, ab) in blocks:
            assert np.all(M[s, s:s + m] == 1.0) and M[s].sum() == m
        off = M - np.diag(np.diag(M))
        assert np.all(off[other] >= 0)
        # rate conservation: a column's rates out of its ion (the diagonal) equal the rates into
        # the other ions, wherever no conservation row replaced one of them (row j - 1 receives
        # recombination from j; ionization and Auger jumps land below j, never on a first ion)
        colsum = M[other].sum(axis=0)
        for j in range(n):
            if j not in cons and (j - 1) not in cons:
                assert abs(colsum[j]) <= 1e-12 * np.abs(M[other, j]).sum(), (c, j, colsum[j])
        # fill: nothing above the super-diagonal or across elements; below the sub-diagonal only
        # the Auger jumps of ions with >= 3 electrons, at most +10 stages
        i, j = np.nonzero(off)
        i, j = i[~np.isin(i, cons)], j[~np.isin(i, cons)]
        assert np.all(i >= j - 1) and np.all(elem[i] == elem[j])
        deep = i > j + 1
        if not auger:
            assert not deep.any()
        else:
            assert deep.any() or not sites
            for ii, jj in zip(i[deep], j[deep]):
                assert jj in sites and ii - jj <= sites[jj] <= 10


def test_ion_structure():
    rng = np.random.default_rng(0)
    A, b = G.ion_matrices(rng, 5, 30)
    _check_rate_structure(A, b, 30, auger=False)
    A, b = G.ion_matrices(rng, 5, 128, auger=True)
    _check_rate_structure(A, b, 128, auger=True)


def test_auger_sites_follow_reference_rule():
    """source/matrix_ion.c:600-616: n_elec = z - istate + 1 capped at 11, jumps 1 .. n_elec - 1.
    Fe (27 stages, z = 26): the neutral atom has 26 electrons -> 10 jumps; stage k has 26 - k."""
    blocks = G.element_layout(200)
    s_fe, m_fe, _ = [b for b in blocks if b[1] == 27][0]
    fe = {col - s_fe: nj for (col, r0, nj) in G.auger_sites(blocks) if s_fe <= col < s_fe + m_fe}
    assert fe[0] == 10 and fe[15] == 10 and fe[16] == 9 and fe[23] == 2 and 24 not in fe
    # H (2 stages) and He (3) have no ion with three electrons
    assert not [c for (c, r0, nj) in G.auger_sites(G.element_layout(5))]


def test_wide_batch_condition_numbers():
    """The precision batch: entries inside the wide rate range and every cell's 2-norm kappa in
    1e6..1e14, measured again here."""
    b = G.uneven(seed=3, wide=True, scale=0.004)
    lk = b.meta["log10_kappa"]
    assert len(lk) == b.ncells and np.all((lk >= 6.0) & (lk <= 14.0))
    for c in range(b.ncells):
        a, _ = b.cell(c)
        assert abs(np.log10(np.linalg.cond(a)) - lk[c]) < 1e-6
        rates = np.abs(a[a != 1.0])
        assert rates[rates > 0].min() >= 1e-20
    assert sum(b.meta["kappa_hist"].values()) == b.ncells


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
