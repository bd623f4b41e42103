"""GPU parity of the two FP64 factor kernels for 175 < n <= 256 against the CPU oracle.

A group of up to one cell per CU runs k_lookahead_f64 (one-wave panel, look-ahead schedule), a
larger group k_blocked_f64<256>; a device-side cell count (the ionization iteration, the mixed
solver's FP64 re-solve) launches both, each working only in its own range.  Every route must give
values equal to the oracle's GSL-order solve and codes equal, and keep the factors an adjoint
solve can use.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(n=200, k_ion=290, k_tie=10, seed=41):
    from python_amd.generators import ion_matrices, pack
    rng = np.random.default_rng(seed)
    A, b = ion_matrices(rng, k_ion, n, auger=True)
    mats, rhs = list(A), list(b)
    for _ in range(k_tie):
        T = np.round(rng.standard_normal((n, n)) * 2) / 2    # pivot ties
        T[np.arange(n), np.arange(n)] += 4.0
        mats.append(T)
        rhs.append(rng.standard_normal(n))
    return pack(mats, rhs)


def _check(oracle_lib, batch, X, codes):
    Xo, co = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=8)
    assert np.array_equal(codes, co), np.where(codes != co)[0][:10]
    for c in np.where(co != 4)[0]:
        o, n = batch.offV[c], batch.n[c]
        assert np.array_equal(X[o:o + n], Xo[o:o + n]), (c, n)


@pytest.mark.parametrize("keep", [False, True])
def test_both_kernels_equal_oracle(hip, oracle_lib, keep):
    import torch
    full = _batch()
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    assert full.ncells > ncu
    small = full.subset(range(ncu // 2))
    for batch, want in ((full, "k_blocked_f64<256"), (small, "k_lookahead_f64")):
        A, B = hip.to_device(batch)
        plan = hip.Plan.for_batch(batch, mode="fp64", graph=False, keep_factors=keep)
        names = [g["fp64_kernel"] for g in plan.describe()["groups"]]
        print(batch.ncells, names, flush=True)
        assert len(names) == 1 and names[0].startswith(want), names
        X, codes, _ = plan.solve(A, B)
        torch.cuda.synchronize()
        _check(oracle_lib, batch, X.cpu().numpy(), codes.cpu().numpy())
        if keep:
            # the kept factors and ipiv: the adjoint solve A^T lambda = g through them matches the
            # oracle's transposed solve on its own LU (the adjoint test's bound)
            rng = np.random.default_rng(5)
            G = torch.from_numpy(rng.random(int(batch.offV[-1] + batch.n[-1]))).cuda()
            Lam, _ = plan.adjoint(X, G)
            torch.cuda.synchronize()
            Lam, Gh = Lam.cpu().numpy(), G.cpu().numpy()
            for c in range(0, batch.ncells, 23):
                a, _ = batch.cell(c)
                o, n = batch.offV[c], batch.n[c]
                LU, perm, _, _ = oracle_lib.lu_decomp(a)
                _, lam = oracle_lib.lu_solve_transpose(LU, perm, Gh[o:o + n])
                assert np.abs(Lam[o:o + n] - lam).max() <= 1e-10 * np.abs(lam).max(), c
        plan.close()


def test_ion_iteration_crosses_the_cell_limit(hip, oracle_lib):
    """30% of the uneven ionization workload: the first iterations hand more than one cell per CU
    of order 175 < n <= 256 to the FP64 solve (the blocked kernel runs), later ones fewer (the
    look-ahead kernel runs); results equal the oracle's per-cell iteration bit for bit."""
    from python_amd.generators import ion_iteration_batch
    from python_amd.ion import IonIteration
    batch = ion_iteration_batch(scale=0.3)
    big = int((np.asarray(batch.n) > 175).sum())
    want = oracle_lib.ion_populations_batch(batch, nthreads=8)
    print("cells", batch.ncells, "of n > 175", big, "iterations hist", np.bincount(want[3])[:20].tolist(),
          flush=True)
    assert big > 256
    it = IonIteration(batch)
    dens, ne, st, its, codes = (t.cpu().numpy() for t in it.populations())
    it.close()
    wd, wne, wst, wit, wcodes = want
    assert np.array_equal(st, wst) and np.array_equal(its, wit) and np.array_equal(codes, wcodes)
    ok = wst == 0
    assert np.array_equal(ne[ok], wne[ok])
    assert np.array_equal(dens.view(np.uint64), wd.view(np.uint64))


def _hard_cells(n, seed):
    """A singular matrix (a zero column: code 4), near-singular ones (singular values down to 1e-17:
    the residual check fires with codes 2 or 3 depending on the sign of b), and a regular one."""
    rng = np.random.default_rng(seed)
    mats, rhs = [], []
    A = rng.standard_normal((n, n))
    A[:, n // 3] = 0.0
    mats.append(A)
    rhs.append(rng.standard_normal(n))
    for s in range(4):
        U, _ = np.linalg.qr(rng.standard_normal((n, n)))
        A = (U * np.logspace(0, -17, n)) @ U.T
        b = rng.standard_normal(n)
        mats.append(A)
        rhs.append(np.abs(b) if s % 2 else b)
    A = rng.standard_normal((n, n))
    A[np.arange(n), np.arange(n)] += 2 * np.sqrt(n)
    mats.append(A)
    rhs.append(rng.standard_normal(n))
    return mats, rhs


def test_singular_and_residual_codes_both_kernels(hip, oracle_lib):
    """Codes 4 (x left untouched), 2 and 3 of singular and near-singular cells of order 200, in a
    group small enough for the look-ahead kernel and in one large enough for the blocked kernel
    (the same cells repeated), equal to the oracle's."""
    import torch
    from python_amd.generators import pack
    mats, rhs = _hard_cells(200, 71)
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    reps = ncu // len(mats) + 2
    for batch, want in ((pack(mats, rhs), "k_lookahead_f64"), (pack(mats * reps, rhs * reps), "k_blocked_f64<256")):
        A, B = hip.to_device(batch)
        plan = hip.Plan.for_batch(batch, mode="fp64", graph=False)
        names = [g["fp64_kernel"] for g in plan.describe()["groups"]]
        assert len(names) == 1 and names[0].startswith(want), names
        X = torch.full_like(B, 7.0)
        Xr, codes, _ = plan.solve(A, B, X)
        torch.cuda.synchronize()
        Xh, ch = Xr.cpu().numpy(), codes.cpu().numpy()
        Xo, co = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=8)
        print(batch.ncells, want, "codes", np.bincount(co, minlength=5).tolist(), flush=True)
        assert np.array_equal(ch, co)
        assert {0, 4} <= set(co.tolist()) and (2 in co or 3 in co)
        for c in range(batch.ncells):
            o, n = batch.offV[c], batch.n[c]
            if co[c] == 4:
                assert np.all(Xh[o:o + n] == 7.0), c
            else:
                assert np.array_equal(Xh[o:o + n], Xo[o:o + n]), c
        plan.close()
