"""GPU checks of the FP64 kernel shapes the autotuner chooses between (hipla_plan_set_variant,
hipla_plan_autotune): every shape of every kernel family -- one-wave-per-matrix kernels with two
or four waves per workgroup (one or two for n <= 64), register tiles or the blocked kernel for
64 < n <= 175, the blocked kernel with 256 or 512 threads for n <= 256 -- gives values equal to
the oracle's GSL-order solve in every cell, with and without kept factors, and the tuned plans
(FP64 and mixed) stay equal too."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch():
    from python_amd.generators import ion_matrices, pack
    rng = np.random.default_rng(31)
    mats, rhs = [], []
    for n, k in ((12, 40), (30, 40), (64, 30), (96, 12), (128, 12), (160, 8), (200, 6), (256, 6)):
        A, b = ion_matrices(rng, k, n, auger=True)
        mats += list(A)
        rhs += list(b)
    return pack(mats, rhs)


def _equal(oracle_lib, batch, X, codes):
    Xo, co = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=8)
    assert np.array_equal(codes, co), np.where(codes != co)[0][:10]
    for c in np.where(co != 4)[0]:
        o, n = batch.offV[c], batch.n[c]
        assert np.array_equal(X[o:o + n], Xo[o:o + n]), (c, n)


@pytest.mark.parametrize("keep", [False, True])
def test_every_shape_equals_oracle(hip, oracle_lib, keep):
    import torch
    batch = _batch()
    A, B = hip.to_device(batch)
    plan = hip.Plan.for_batch(batch, mode="fp64", graph=False, keep_factors=keep)
    groups = plan.describe()["groups"]
    for v in (0, 1):
        for gi in range(len(groups)):
            try:
                plan.set_variant(gi, v)
            except hip.HiplaError:
                plan.set_variant(gi, 0)       # a family with one shape
        X = torch.zeros_like(B)
        Xr, codes, _ = plan.solve(A, B, X)
        torch.cuda.synchronize()
        print("variant", v, [g["fp64_kernel"] for g in plan.describe()["groups"]], flush=True)
        _equal(oracle_lib, batch, Xr.cpu().numpy(), codes.cpu().numpy())


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_tuned_plans_equal_oracle(hip, oracle_lib, mode):
    import torch
    batch = _batch()
    A, B = hip.to_device(batch)
    plan = hip.Plan.for_batch(batch, mode=mode, graph=True)
    X = torch.zeros_like(B)
    codes = torch.empty(batch.ncells, dtype=torch.int32, device=B.device)
    plan.autotune(A, B, X, codes)
    d = plan.describe()
    print(mode, json.dumps([(g["nmax"], g["direct_fp64"], g["variant"]) for g in d["groups"]]), flush=True)
    X.zero_()
    for _ in range(2):                       # capture, then replay
        Xr, codes, _ = plan.solve(A, B, X, codes)
    torch.cuda.synchronize()
    if mode == "fp64":
        _equal(oracle_lib, batch, Xr.cpu().numpy(), codes.cpu().numpy())
    else:
        # mixed: cells kept in FP32 + refinement are within 1e-8; FP64 cells equal
        Xo, co = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=8)
        Xh = Xr.cpu().numpy()
        for c in np.where(co != 4)[0]:
            o, n = batch.offV[c], batch.n[c]
            err = np.abs(Xh[o:o + n] - Xo[o:o + n]).max() / max(np.abs(Xo[o:o + n]).max(), 1e-300)
            assert err <= 1e-8, (c, n, err)
