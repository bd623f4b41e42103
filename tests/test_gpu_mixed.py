"""GPU checks of the mixed-precision solver (equilibrated FP32 LU + FP64 refinement + FP64 fallback).

Tolerance (the stated mixed-precision bound): per-cell normwise relative error
||x - x_oracle||_inf / ||x_oracle||_inf <= 1e-8 after refinement, where the oracle is the FP64
GSL-order solve.  Cells the solver hands back to FP64 must then equal the oracle exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-8


def _relerr(batch, X, Xo, codes_oracle=None):
    """Per-cell normwise relative error; 0 for cells the oracle reports singular (code 4: x is
    not written by either side)."""
    errs = np.zeros(batch.ncells)
    for c in range(batch.ncells):
        if codes_oracle is not None and codes_oracle[c] == 4:
            continue
        o, n = batch.offV[c], batch.n[c]
        xo = Xo[o:o + n]
        errs[c] = np.abs(X[o:o + n] - xo).max() / max(np.abs(xo).max(), 1e-300)
    return errs


def _cond_est(batch, cells):
    out = {}
    for c in cells:
        a, _ = batch.cell(c)
        out[c] = np.linalg.cond(a)
    return out


@pytest.mark.parametrize("scale", [0.02])
def test_mixed_uneven_matches_oracle(hip, oracle_lib, scale):
    from python_amd.generators import uneven
    batch = uneven(scale=scale)
    print("cells", batch.ncells, flush=True)
    X, codes, stats = hip.solve_batch(batch, mode="mixed")
    Xo, co = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=8)
    err = _relerr(batch, X, Xo, co)
    assert np.array_equal(codes == 4, co == 4)
    fb = stats["fallback"]
    print("max err", err.max(), "fallback", int((fb == 1).sum()), "steps hist",
          np.bincount(stats["refine_steps"][fb == 0]), flush=True)
    bad = np.where(err > TOL)[0]
    if len(bad):
        ce = _cond_est(batch, bad[:10])
        print("over tol:", [(int(c), int(batch.n[c]), float(err[c]), float(ce[c])) for c in bad[:10]], flush=True)
    assert np.all(err <= TOL)
    # fallback cells are the FP64 GSL-order answer
    for c in np.where((fb == 1) & (co != 4))[0]:
        o, n = batch.offV[c], batch.n[c]
        assert np.array_equal(X[o:o + n], Xo[o:o + n])
    assert np.array_equal(codes, co) or np.mean(codes == co) > 0.99


def test_mixed_sizes_sweep(hip, oracle_lib):
    from python_amd.generators import ion_matrices, pack
    rng = np.random.default_rng(17)
    mats, rhs = [], []
    for n in (1, 2, 5, 12, 16, 17, 30, 32, 33, 63, 64, 65, 80, 96, 97, 128, 129, 150, 192, 193, 200, 240, 256):
        A, b = ion_matrices(rng, 2, n)
        mats += [A[0], A[1]]
        rhs += [b[0], b[1]]
        D = rng.standard_normal((n, n)) + 2 * np.sqrt(n) * np.eye(n)
        mats.append(D)
        rhs.append(rng.standard_normal(n))
    batch = pack(mats, rhs)
    X, codes, stats = hip.solve_batch(batch, mode="mixed")
    Xo, co = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=8)
    err = _relerr(batch, X, Xo, co)
    assert np.array_equal(codes == 4, co == 4)
    for c in range(batch.ncells):
        print(int(batch.n[c]), float(err[c]), int(stats["refine_steps"][c]), int(stats["fallback"][c]),
              float(stats["backward_err"][c]), flush=True) if err[c] > TOL else None
    assert np.all(err <= TOL)


def test_mixed_singular_goes_to_fp64(hip, oracle_lib):
    from python_amd.generators import pack
    rng = np.random.default_rng(2)
    mats, rhs = [], []
    for n in (10, 40, 100, 200):
        A = rng.standard_normal((n, n))
        A[:, 3] = 0.0
        mats.append(A)
        rhs.append(rng.standard_normal(n))
    batch = pack(mats, rhs)
    X, codes, stats = hip.solve_batch(batch, mode="mixed")
    assert np.all(codes == 4)
    assert np.all(stats["fallback"] == 1)


def test_mixed_autotuned_fp64_classes_merged(hip, oracle_lib):
    """After autotuning, adjacent size classes routed straight to the same FP64 kernel are solved
    by one launch (also under graph replay); every such cell still equals the FP64 oracle and the
    codes match."""
    import torch
    from python_amd.generators import uneven
    batch = uneven(scale=0.06)        # 600 cells: 30 each of n = 200 and n = 256
    plan = hip.Plan.for_batch(batch, mode="mixed", graph=True)
    A, B = hip.to_device(batch)
    X = torch.zeros_like(B)
    codes = torch.empty(batch.ncells, dtype=torch.int32, device=B.device)
    plan.autotune(A, B, X, codes)
    desc = plan.describe()
    routed = [g for g in desc["groups"] if g["direct_fp64"]]
    print("classes routed to FP64:", [(g["kernel"], g["nmax"], g["count"]) for g in routed], flush=True)
    Xo, co = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=8)
    for rep in range(3):
        X.zero_()
        plan.solve(A, B, X, codes, stats=True)
        torch.cuda.synchronize()
        Xh, ch = X.cpu().numpy(), codes.cpu().numpy()
        assert np.array_equal(ch, co), rep
        err = _relerr(batch, Xh, Xo, co)
        assert err.max() <= TOL, (rep, err.max())
        for g in routed:
            for c in np.where((batch.n == g["nmax"]) & (co != 4))[0]:
                o, n = batch.offV[c], batch.n[c]
                assert np.array_equal(Xh[o:o + n], Xo[o:o + n]), (rep, c)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_forked_back_substitution_graph_replays(hip, oracle_lib, mode):
    """A plan whose largest class runs on the blocked kernel and has smaller classes beside it:
    the largest class's back substitution runs on the plan's second stream (forked after its
    factor kernel, joined at the end), in eager calls and in graph replays.  Every replay equals
    the oracle (FP64 kernels: bit for bit) with codes equal, with eager work and synchronisations
    between the replays."""
    import torch
    from python_amd.generators import ion_matrices, pack
    rng = np.random.default_rng(21)
    mats, rhs = [], []
    for n, count in ((12, 40), (48, 30), (100, 20), (230, 12)):
        A, b = ion_matrices(rng, count, n)
        mats += list(A)
        rhs += list(b)
    batch = pack(mats, rhs)
    Xo, co = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=8)
    A, B = hip.to_device(batch)
    for graph in (False, True):
        plan = hip.Plan.for_batch(batch, mode=mode, graph=graph)
        X = torch.zeros_like(B)
        codes = torch.empty(batch.ncells, dtype=torch.int32, device=B.device)
        if mode == "mixed":
            plan.autotune(A, B, X, codes)
        for rep in range(3):
            X.zero_()
            plan.solve(A, B, X, codes)
            torch.cuda.synchronize()
            Xh, ch = X.cpu().numpy(), codes.cpu().numpy()
            assert np.array_equal(ch, co), (graph, rep)
            if mode == "fp64" or all(g["direct_fp64"] for g in plan.describe()["groups"]):
                assert np.array_equal(Xh, Xo), (graph, rep)
            else:
                err = _relerr(batch, Xh, Xo, co)
                assert err.max() <= TOL, (graph, rep, err.max())
            torch.empty(1 << 20, device=B.device).normal_()      # eager work between replays
        plan.close()


@pytest.mark.parametrize("tuned", [False, True])
def test_wide_range_batch(hip, oracle_lib, tuned):
    """The precision workload: rates inside 1e-20..1e4 and 2-norm condition numbers 1e6..1e14 (all
    seven sizes, 2% of the 10,000-cell batch, seed 3).  Mixed mode untuned (every class on the
    FP32 path) and autotuned: kept cells within 1e-8 of the FP64 oracle, handed-back cells equal to
    it, codes equal to the oracle's; FP64 mode equal to the oracle in every cell."""
    import torch
    from python_amd.generators import uneven
    batch = uneven(seed=3, wide=True, scale=0.02)
    print("cells", batch.ncells, "kappa hist", batch.meta["kappa_hist"], flush=True)
    Xo, co = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=8)
    A, B = hip.to_device(batch)
    plan = hip.Plan.for_batch(batch, mode="mixed", graph=True)
    X = torch.zeros_like(B)
    codes = torch.empty(batch.ncells, dtype=torch.int32, device=B.device)
    if tuned:
        plan.autotune(A, B, X, codes)
    X.zero_()
    _, _, st = plan.solve(A, B, X, codes, stats=True)
    torch.cuda.synchronize()
    Xh, ch, stats = X.cpu().numpy(), codes.cpu().numpy(), hip.stats_to_numpy(st)
    err = _relerr(batch, Xh, Xo, co)
    fb = stats["fallback"]
    print("singular (oracle code 4):", int((co == 4).sum()), "handed back:", int((fb == 1).sum()),
          "FP64 direct:", int((fb == 2).sum()), "max err of kept cells:", float(err[fb == 0].max(initial=0.0)),
          flush=True)
    assert np.all(err <= TOL)
    for c in np.where((fb != 0) & (co != 4))[0]:
        o, n = batch.offV[c], batch.n[c]
        assert np.array_equal(Xh[o:o + n], Xo[o:o + n]), c
    assert np.array_equal(ch, co), np.where(ch != co)[0][:10]
    X64, c64, _ = hip.solve_batch(batch, mode="fp64")
    assert np.array_equal(c64, co)
    for c in np.where(co != 4)[0]:
        o, n = batch.offV[c], batch.n[c]
        assert np.array_equal(X64[o:o + n], Xo[o:o + n]), c
