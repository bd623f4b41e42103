"""The oracle's macro-atom emissivity contraction (oracle.macro_emiss: the reference's
invert_matrix + B = N^-1 R + a^T B order, source/macro_accelerate.c:315-334 and
source/macro_gen_f.c:490-506) pinned by what the mathematics fixes:

* N = I: B = R, so out = a * r * e exactly;
* a 2 x 2 closed form of N^-T a;
* random non-symmetric N = I - Q: out = (N^-T a) r e within 1e-12 relative, and a planted
  transposition (N^-1 a instead of N^-T a) is told apart.
"""
import numpy as np

import oracle


def test_identity_is_exact():
    rng = np.random.default_rng(0)
    n = 9
    r, a, e = rng.random(n), rng.random(n), rng.random(n)
    code, out = oracle.macro_emiss(np.eye(n), r, a, e)
    assert code == 0
    assert np.array_equal(out, a * r * e)


def test_two_by_two_closed_form():
    p, q = 0.3, 0.45
    N = np.array([[1.0, -q], [-p, 1.0]])
    r, a, e = np.array([0.7, 0.55]), np.array([2.0, 3.0]), np.array([0.5, 0.25])
    det = 1.0 - p * q
    y = np.array([a[0] + p * a[1], q * a[0] + a[1]]) / det        # N^-T a
    code, out = oracle.macro_emiss(N, r, a, e)
    assert code == 0
    np.testing.assert_allclose(out, y * r * e, rtol=1e-14)


def test_random_against_transposed_solve():
    from python_amd.generators import macro_emiss_vectors, macro_q
    rng = np.random.default_rng(3)
    for n in (5, 30, 76):
        N = macro_q(rng, 1, n)[0]
        r, a, e = macro_emiss_vectors(rng, N)
        code, out = oracle.macro_emiss(N, r, a, e)
        assert code == 0
        want = np.linalg.solve(N.T, a) * r * e
        assert np.abs(out - want).max() <= 1e-12 * np.abs(want).max()
        wrong = np.linalg.solve(N, a) * r * e
        assert np.abs(out - wrong).max() > 1e-6 * np.abs(want).max()


def test_singular_returns_code():
    N = np.eye(4)
    N[2, :] = 0.0
    code, out = oracle.macro_emiss(N, np.ones(4), np.ones(4), np.ones(4))
    assert code != 0 and out is None
