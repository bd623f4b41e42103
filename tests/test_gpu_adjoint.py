"""GPU checks of the adjoint (A^T lambda = g with the stored LU) and d(populations)/dT_e.

* lambda against the oracle's transpose solve on the same factors (1e-10 relative);
* dL/dT_e through torch autograd against the FP64 reference gradient -lambda^T (dA/dT) x computed
  on the CPU with the oracle (1e-8 relative to the magnitude of the summed terms), and against
  central finite differences of double-double refined CPU solves (1e-6 relative) on the cells
  whose gradient finite differences can resolve;
* one Newton step on T_e at least halves the temperature error of every cell whose loss
  depends on T_e.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(ncells=64, n=64, seed=5):
    from python_amd.adjoint import IonTe
    from python_amd.generators import ion_matrices_te
    rng = np.random.default_rng(seed)
    d = ion_matrices_te(rng, ncells, n)
    w = rng.random((ncells, n))
    return d, IonTe(d), w


def test_adjoint_lambda_matches_oracle(hip, oracle_lib):
    import torch
    from python_amd.generators import pack
    rng = np.random.default_rng(1)
    mats = [rng.standard_normal((k, k)) + 2 * np.sqrt(k) * np.eye(k) for k in (3, 16, 40, 64, 100, 141, 200, 256)]
    rhs = [rng.standard_normal(k) for k in (3, 16, 40, 64, 100, 141, 200, 256)]
    b = pack(mats, rhs)
    plan = hip.Plan.for_batch(b, mode="fp64", graph=False, keep_factors=True)
    A, B = hip.to_device(b)
    X, codes, _ = plan.solve(A, B)
    G = torch.from_numpy(np.concatenate([rng.random(k) for k in b.n])).cuda()
    Lam, GradA = plan.adjoint(X, G, GradA=torch.empty_like(A))
    torch.cuda.synchronize()
    Lam, GradA, Xh = Lam.cpu().numpy(), GradA.cpu().numpy(), X.cpu().numpy()
    Gh = G.cpu().numpy()
    for c in range(b.ncells):
        a, _ = b.cell(c)
        n = a.shape[0]
        LU, perm, _, _ = oracle_lib.lu_decomp(a)
        st, lam = oracle_lib.lu_solve_transpose(LU, perm, Gh[b.offV[c]:b.offV[c] + n])
        got = Lam[b.offV[c]:b.offV[c] + n]
        assert np.abs(got - lam).max() <= 1e-10 * np.abs(lam).max(), n
        ga = GradA[b.offA[c]:b.offA[c] + n * n].reshape(n, n)
        assert np.allclose(ga, -np.outer(got, Xh[b.offV[c]:b.offV[c] + n]), rtol=0, atol=0)


def test_te_assembly_on_gpu_matches_numpy(hip):
    """hipla_te_assemble against the numpy restatement of the rate model (te_assemble): the same
    entries, up to the last bits of exp and pow (1e-14 relative), the abundance rows and b exact."""
    import torch
    from python_amd.generators import te_assemble
    d, model, w = _model()
    n, ncells = d["A"].shape[1], d["A"].shape[0]
    T = torch.from_numpy(d["te"]).cuda()
    A, b = model.assemble_gpu(T)
    A = A.cpu().numpy().reshape(ncells, n, n)
    b = b.cpu().numpy().reshape(ncells, n)
    Aref, _, bref = te_assemble(n, d["blocks"], d["params"], d["te"])
    assert np.array_equal(b, bref)
    assert np.array_equal(A == 0.0, Aref == 0.0)
    assert np.allclose(A, Aref, rtol=1e-14, atol=0.0)
    for (s, m, ab) in model.cons:
        assert np.array_equal(A[:, s, :], Aref[:, s, :])


@pytest.mark.parametrize("fused", [True, False])
def test_te_gradient_fd_and_reference(hip, oracle_lib, fused):
    """dL/dT_e through the adjoint: fused (GPU assembly, hipla_te_grad from lambda) and through
    torch's chain rule over the dense dL/dA (torch assembly)."""
    import torch
    from python_amd.adjoint import loss_and_grad
    from python_amd.generators import te_assemble
    d, model, w = _model()
    n, ncells = d["A"].shape[1], d["A"].shape[0]
    plan = model.plan()
    T = torch.from_numpy(d["te"]).cuda()
    wt = torch.from_numpy(w).cuda()
    L, g, codes = loss_and_grad(model, plan, T, wt, fused=fused)
    L, g = L.cpu().numpy(), g.cpu().numpy()

    # FP64 reference gradient with the oracle, -lambda^T (dA/dT) x, on every cell.  The sum has
    # cancelling terms (dA's diagonal is minus its column sums), so besides 1e-8 of the result
    # the bound allows 1e-10 (the lambda tolerance above) of the magnitude of the terms,
    # sum |lambda_i| |dA_ij| |x_j|.
    A, dA, b = te_assemble(n, d["blocks"], d["params"], d["te"])
    for c in range(ncells):
        LU, perm, _, _ = oracle_lib.lu_decomp(A[c])
        _, x = oracle_lib.lu_solve(LU, perm, b[c])
        _, lam = oracle_lib.lu_solve_transpose(LU, perm, w[c])
        gref = -lam @ (dA[c] @ x)
        mag = np.abs(lam) @ (np.abs(dA[c]) @ np.abs(x))
        assert abs(gref - g[c]) <= 1e-8 * abs(gref) + 1e-10 * mag, (c, gref, g[c], mag)

    # Central differences with h = 1e-5 T (truncation ~h^2, about 1e-10 relative), each L(T +- h)
    # from double-double refined CPU solves so that rounding in the solve does not enter the
    # difference.  Rounding of A(T +- h) itself still does: the check covers the cells whose
    # relative sensitivity |T dL/dT / L| is at least 1e-2, where that noise stays below 1e-7.
    h = 1e-5 * d["te"]

    def loss_dd(te):
        Ah, _, bh = te_assemble(n, d["blocks"], d["params"], te)
        return np.array([w[c] @ oracle_lib.solve_dd_refined(Ah[c], bh[c])[1] for c in range(ncells)])

    fd = (loss_dd(d["te"] + h) - loss_dd(d["te"] - h)) / (2 * h)
    sens = np.abs(d["te"] * g) / np.maximum(np.abs(L), 1e-300)
    ok = sens >= 1e-2
    rel = np.abs(fd - g)[ok] / np.maximum(np.abs(fd), np.abs(g))[ok]
    print(f"fd check on {ok.sum()} of {ncells} cells: max rel err {rel.max():.2e}", flush=True)
    assert ok.sum() >= ncells // 4
    assert np.all(rel <= 1e-6)


def test_newton_step_on_te(hip):
    import torch
    from python_amd.adjoint import loss_and_grad, newton_te
    d, model, w = _model(ncells=128, seed=9)
    plan = model.plan()
    rng = np.random.default_rng(10)
    Tstar = torch.from_numpy(d["te"]).cuda()
    wt = torch.from_numpy(w).cuda()
    Lstar, _, _ = loss_and_grad(model, plan, Tstar, wt)
    T0 = Tstar * torch.from_numpy(1.0 + rng.uniform(-0.02, 0.02, Tstar.shape[0])).cuda()
    T1, L0, dL = newton_te(model, plan, T0, wt, Lstar)
    e0 = (T0 - Tstar).abs().cpu().numpy()
    e1 = (T1 - Tstar).abs().cpu().numpy()
    # cells whose loss depends on T_e at all (relative sensitivity |T dL/dT / L| >= 1e-2); the
    # others have populations frozen in one ion stage and no temperature signal to follow
    sens = (T0 * dL).abs().cpu().numpy() / np.maximum(L0.abs().cpu().numpy(), 1e-300)
    ok = sens >= 1e-2
    ratio = e1[ok] / e0[ok]
    print("cells", ok.sum(), "worst error ratio", ratio.max(), flush=True)
    assert ok.sum() >= 16
    assert np.all(ratio < 0.5)
