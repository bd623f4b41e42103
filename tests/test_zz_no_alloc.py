"""The allocation-free entry point hipla_solve_matrix_noalloc (no counterpart in the reference):
through the C ABI on device arrays, two calls in a row (the second at a smaller order, so the
static scratch array holds the first call's factors when the second starts), each equal to the
drop-in solve_matrix on the same matrix -- values bit for bit, codes equal -- and no growth of
the library's allocation count."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(hip, A, b):
    import torch
    n = b.shape[0]
    dA = torch.from_numpy(np.ascontiguousarray(A).ravel()).cuda()
    dB = torch.from_numpy(b).cuda()
    dX = torch.full((n,), -7.0, dtype=torch.float64, device="cuda")
    dc = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream()
    rc = hip.lib().hipla_solve_matrix_noalloc(ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(dB.data_ptr()), n,
                                             ctypes.c_void_p(dX.data_ptr()), ctypes.c_void_p(dc.data_ptr()),
                                             ctypes.c_void_p(int(s.cuda_stream)))
    assert rc == 0, hip.last_error_detail()
    torch.cuda.synchronize()
    return int(dc.cpu()[0]), dX.cpu().numpy()


@pytest.mark.parametrize("orders", [(200, 30), (1024, 130), (64, 12)])
def test_noalloc_equals_solve_matrix(hip, orders):
    from python_amd.generators import ion_matrices, macro_q
    rng = np.random.default_rng(sum(orders))
    for n in orders:
        if n >= 512:
            A = macro_q(rng, 1, n)[0]
            b = rng.random(n)
        else:
            A, b = ion_matrices(rng, 1, n, auger=True)
            A, b = A[0], b[0]
        x_ref = np.zeros(n)
        code_ref = hip.solve_matrix(A.ravel().copy(), b.copy(), n, x_ref, -1)
        before = hip.alloc_stats()[0]
        code, x = _noalloc(hip, A, b)
        after = hip.alloc_stats()[0]
        print("n", n, "codes", code, code_ref, "allocations", before, after, flush=True)
        assert after == before
        assert code == code_ref
        if code != 4:
            assert np.array_equal(x, x_ref), np.abs(x - x_ref).max()


def test_noalloc_singular_and_limits(hip):
    import torch
    n = 40
    A = np.eye(n)
    A[:, 5] = 0.0
    code, x = _noalloc(hip, A, np.ones(n))
    assert code == 4 and np.all(x == -7.0)          # x not written on 4
    d = torch.zeros(4, dtype=torch.float64, device="cuda")
    rc = hip.lib().hipla_solve_matrix_noalloc(ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(d.data_ptr()), 1025,
                                             ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(d.data_ptr()), None)
    assert rc == 5 and "1024" in hip.last_error_detail()
