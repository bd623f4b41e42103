"""GPU checks of the FP64 MFMA Gauss-Jordan inverse (k_gj.hip): orders that are multiples of 64
from 128 on, row diagonally dominant cells swept without pivoting, every other cell handed to the
GSL-order pivoting kernel.  Tolerance: 1e-8 relative Frobenius error against the FP64 oracle's
gsl_linalg_LU_invx restatement; singular cells give code 1 (GSL_EDOM) as the oracle does."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(hip, oracle_lib, b2, graph=False):
    import torch
    plan = hip.Plan.for_batch(b2, mode="fp64", graph=graph)
    A = torch.from_numpy(b2.A).cuda()
    Inv, codes, _ = plan.invert(A)
    torch.cuda.synchronize()
    Inv, codes = Inv.cpu().numpy(), codes.cpu().numpy()
    worst = 0.0
    for c in range(b2.ncells):
        n = int(b2.n[c])
        a = b2.A[b2.offA[c]:b2.offA[c] + n * n].reshape(n, n)
        code, ref = oracle_lib.invert_matrix(a)
        assert codes[c] == code, (c, n, codes[c], code)
        if code:
            continue
        got = Inv[b2.offA[c]:b2.offA[c] + n * n].reshape(n, n)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        worst = max(worst, err)
        assert err <= 1e-8, (c, n, err)
    return plan, A, worst


def test_swept_and_handed_back_cells(hip, oracle_lib):
    """I - Q cells (swept) and dense random cells (not dominant: handed back) at n = 128, 192,
    256, 320, plus a dominant cell with a zero row (a zero column of the swept matrix: code 1)."""
    from python_amd.generators import macro_q, pack
    rng = np.random.default_rng(17)
    mats = []
    for n in (128, 192, 256, 320):
        mats += [macro_q(rng, 1, n)[0] for _ in range(3)]
        mats.append(rng.standard_normal((n, n)) + 2.0 * np.eye(n))
    sing = macro_q(rng, 1, 192)[0]
    sing[57, :] = 0.0
    mats.append(sing)
    b2 = pack(mats)
    _, _, worst = _check(hip, oracle_lib, b2)
    print("worst rel frob err", worst, flush=True)


def test_c4_sizes_and_graph_replays(hip, oracle_lib):
    """Two cells of each C4 order (256, 512, 1024) through a captured graph, replayed three times
    with eager work in between: every replay equal to the first, the first within 1e-8."""
    import torch
    from python_amd.generators import macro_inverse
    batch = macro_inverse(ncells=6, sizes=(256, 512, 1024))
    plan, A, worst = _check(hip, oracle_lib, batch, graph=True)
    print("worst rel frob err", worst, flush=True)
    first = None
    for rep in range(3):
        Inv, codes, _ = plan.invert(A)
        torch.cuda.synchronize()
        got = Inv.cpu().numpy()
        if first is None:
            first = got
        assert np.array_equal(got, first), rep
        torch.empty(1 << 20, device=A.device).normal_()


def test_dropin_invert_swept_order(hip, oracle_lib):
    from python_amd.generators import macro_q
    rng = np.random.default_rng(23)
    N = macro_q(rng, 1, 256)[0]
    out = np.zeros(256 * 256)
    assert hip.invert_matrix(N.ravel().copy(), out, 256) == 0
    code, ref = oracle_lib.invert_matrix(N)
    assert np.linalg.norm(out.reshape(256, 256) - ref) <= 1e-8 * np.linalg.norm(ref)


@pytest.mark.parametrize("sizes", [(128, 256), (512, 1024)])
def test_mixed_inverse_bf16x3_refined(hip, oracle_lib, sizes):
    """HIPLA_MODE_MIXED: the BF16x3 sweep in FP32 plus two FP64 refinement steps, within 1e-8
    of the oracle on I - Q cells (plus one handed-back dense cell per size)."""
    import torch
    from python_amd.generators import macro_q, pack
    rng = np.random.default_rng(sum(sizes))
    mats = []
    for n in sizes:
        mats += [macro_q(rng, 1, n)[0] for _ in range(2)]
        mats.append(rng.standard_normal((n, n)) + 2.0 * np.eye(n))
    b2 = pack(mats)
    plan = hip.Plan.for_batch(b2, mode="mixed", graph=False)
    A = torch.from_numpy(b2.A).cuda()
    Inv, codes, _ = plan.invert(A)
    torch.cuda.synchronize()
    Inv, codes = Inv.cpu().numpy(), codes.cpu().numpy()
    for c in range(b2.ncells):
        n = int(b2.n[c])
        a = b2.A[b2.offA[c]:b2.offA[c] + n * n].reshape(n, n)
        code, ref = oracle_lib.invert_matrix(a)
        assert codes[c] == code
        got = Inv[b2.offA[c]:b2.offA[c] + n * n].reshape(n, n)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print("n", n, "mixed rel frob err", err, flush=True)
        assert err <= 1e-8, (c, n, err)
