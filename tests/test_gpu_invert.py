"""GPU checks of invert_matrix (drop-in and batched) against the reference fixtures and the oracle.

Tolerances: the reference's own fixtures at 1e-6 absolute (test_matrix.c:198, but on every
entry, not only row 0) and 2e-4 relative on the full inverse_macro matrix (its matrix.txt is
rounded to 7 digits, kappa ~ 9e3); batched cells within 1e-8 relative Frobenius of the FP64
oracle inverse.  The inverse's LU is the GSL-order elimination, so singular verdicts match.
"""
import numpy as np
import pytest

from tests.conftest import load_fixture

pytestmark = pytest.mark.gpu


def test_dropin_inverse_fixtures(hip, oracle_lib):
    n, M = load_fixture("inverse_small", "matrix.txt")
    _, I = load_fixture("inverse_small", "inverse.txt")
    out = np.zeros(n * n)
    assert hip.invert_matrix(M.copy(), out, n) == 0
    assert np.abs(out - I).max() <= 1e-6
    n, M = load_fixture("inverse_macro", "matrix.txt")
    _, I = load_fixture("inverse_macro", "inverse.txt")
    out = np.zeros(n * n)
    assert hip.invert_matrix(M.copy(), out, n) == 0
    assert np.abs(out[:n] - I[:n]).max() <= 1e-6
    assert np.abs(out - I).max() / np.abs(I).max() < 2e-4
    code, ref = oracle_lib.invert_matrix(M.reshape(n, n))
    assert np.linalg.norm(out - ref.ravel()) / np.linalg.norm(ref) <= 1e-12


def test_dropin_inverse_singular_untouched(hip):
    n = 5
    M = np.arange(n * n, dtype=float)     # rank 2
    out = np.full(n * n, 3.0)
    assert hip.invert_matrix(M, out, n) == 1
    assert np.all(out == 3.0)


def test_batched_inverse_macro_subset(hip, oracle_lib):
    import torch
    from python_amd.generators import macro_inverse, pack
    batch = macro_inverse(ncells=6, sizes=(256, 512, 1024))
    # plus odd sizes around the panel width and an ionization-style matrix
    rng = np.random.default_rng(3)
    extra = [rng.standard_normal((k, k)) + k ** 0.5 * np.eye(k) for k in (1, 7, 9, 33, 100, 141, 200)]
    mats = [batch.cell(c)[0] for c in range(batch.ncells)] + extra
    b2 = pack(mats)
    plan = hip.Plan.for_batch(b2, mode="fp64", graph=False)
    A = torch.from_numpy(b2.A).cuda()
    Inv, codes, _ = plan.invert(A)
    torch.cuda.synchronize()
    Inv = Inv.cpu().numpy()
    codes = codes.cpu().numpy()
    assert np.all(codes == 0)
    for c in range(b2.ncells):
        n = int(b2.n[c])
        a = b2.A[b2.offA[c]:b2.offA[c] + n * n].reshape(n, n)
        code, ref = oracle_lib.invert_matrix(a)
        got = Inv[b2.offA[c]:b2.offA[c] + n * n].reshape(n, n)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print("n", n, "rel frob err", err, flush=True)
        assert err <= 1e-8
        if n >= 256:
            # M-matrix inverse (I - Q)^-1 = sum Q^k is entrywise non-negative
            assert got.min() >= -1e-12 * np.abs(got).max()


def test_batched_inverse_graph_replay(hip):
    import torch
    from python_amd.generators import pack
    rng = np.random.default_rng(8)
    mats = [rng.standard_normal((k, k)) + 3 * np.eye(k) for k in (10, 70, 160)]
    b2 = pack(mats)
    plan = hip.Plan.for_batch(b2, mode="fp64", graph=True)
    A = torch.from_numpy(b2.A).cuda()
    ref = None
    for rep in range(3):
        Inv, codes, _ = plan.invert(A)
        torch.cuda.synchronize()
        got = Inv.cpu().numpy()
        if ref is None:
            ref = got
        assert np.array_equal(got, ref)
        _ = (A * 2).sum().item()
