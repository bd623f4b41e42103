"""GPU checks of the library's host-side contracts: drop-in calls stop allocating once every order
has been seen, a changed refinement setting reaches a graph-replayed mixed plan, and the adjoint
refuses stale factors and gives zeros for singular cells."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_dropin_warm_calls_allocate_nothing(hip, oracle_lib):
    """macro_pops-like traffic: solve_matrix at orders 5, 12, 76, 30 interleaved with
    invert_matrix (source/macro_gov.c:366-444, source/matrix_ion.c:276,
    source/macro_accelerate.c:315), 100 calls.  After the first pass the library's allocation
    count stays put, and every result equals the oracle (solve: exactly; inverse: 1e-12)."""
    from python_amd.generators import ion_matrices, macro_q
    rng = np.random.default_rng(21)
    orders = (5, 12, 76, 30)
    systems = {n: ion_matrices(rng, 1, n) for n in orders}
    inverses = {n: macro_q(rng, 1, n)[0] for n in orders}
    counts = []
    for call in range(100):
        n = orders[call % 4]
        if (call // 4) % 2 == 0:
            A, b = systems[n][0][0], systems[n][1][0]
            x = np.zeros(n)
            code = hip.solve_matrix(A.ravel().copy(), b.copy(), n, x, -1)
            co, xo = oracle_lib.solve_matrix(A, b)
            assert code == co and np.array_equal(x, xo), (call, n, code, co)
        else:
            N = inverses[n]
            inv = np.zeros(n * n)
            code = hip.invert_matrix(N.ravel().copy(), inv, n)
            co, ref = oracle_lib.invert_matrix(N)
            assert code == co == 0
            assert np.linalg.norm(inv.reshape(n, n) - ref) <= 1e-12 * np.linalg.norm(ref), (call, n)
        counts.append(hip.alloc_stats()[0])
    print("allocation counts after calls 8, 16, 99:", counts[7], counts[15], counts[99], flush=True)
    # the first eight calls see every order as a solve and as an inverse
    assert counts[99] == counts[7]


def test_set_refinement_reaches_graph_replays(hip):
    """A mixed plan replays its captured graph; set_refinement must drop it so the next solve runs
    with the new step cap (the captured kernel arguments hold the old one)."""
    import torch
    from python_amd.generators import uneven
    batch = uneven(scale=0.01)
    A, B = hip.to_device(batch)
    plan = hip.Plan.for_batch(batch, mode="mixed", graph=True)
    X = torch.zeros_like(B)
    st = torch.zeros(batch.ncells * 24, dtype=torch.uint8, device=B.device)
    for _ in range(2):                           # capture, then replay
        plan.solve(A, B, X, stats=st)
    torch.cuda.synchronize()
    s5 = hip.stats_to_numpy(st).copy()
    assert s5["refine_steps"].max() >= 1
    plan.set_refinement(0)
    st.zero_()
    plan.solve(A, B, X, stats=st)
    torch.cuda.synchronize()
    s0 = hip.stats_to_numpy(st)
    print("steps before", np.bincount(s5["refine_steps"]), "after", np.bincount(s0["refine_steps"]), flush=True)
    assert s0["refine_steps"].max() == 0
    plan.set_refinement(5)
    st.zero_()
    plan.solve(A, B, X, stats=st)
    torch.cuda.synchronize()
    assert np.array_equal(hip.stats_to_numpy(st)["refine_steps"], s5["refine_steps"])


def test_adjoint_stale_factors_and_singular_cells(hip):
    import torch
    from python_amd.adjoint import solve
    from python_amd.generators import pack
    rng = np.random.default_rng(4)
    good = rng.standard_normal((6, 6)) + 6 * np.eye(6)
    sing = good.copy()
    sing[:, 2] = 0.0                                           # an exact zero pivot: code 4
    batch = pack([good, sing, good.T.copy()], [rng.standard_normal(6) for _ in range(3)])
    plan = hip.Plan.for_batch(batch, mode="fp64", graph=False, keep_factors=True)
    A, B = hip.to_device(batch)
    A.requires_grad_(True)
    X, codes = solve(A, B, plan)
    loss = X.sum()
    gA, = torch.autograd.grad(loss, A, retain_graph=True)
    torch.cuda.synchronize()
    assert codes.cpu().tolist() == [0, 4, 0]
    g = gA.cpu().numpy()
    assert np.isfinite(g).all()
    assert not g[36:72].any()                                  # the singular cell's dL/dA
    assert g[:36].any() and g[72:].any()
    # a second forward on the same plan before the first backward: the kept factors are stale
    X1, _ = solve(A, B, plan)
    X2, _ = solve(A, B, plan)
    with pytest.raises(RuntimeError, match="solved again"):
        torch.autograd.grad(X1.sum(), A)
    torch.autograd.grad(X2.sum(), A)


def test_wire_file_through_a_plan(hip, oracle_lib, tmp_path):
    """A 1% subset of the uneven batch written with hipla_batch_write, read back natively, solved
    by a batched FP64 plan: equal to the oracle on the original arrays, codes equal."""
    import torch
    from python_amd import wire
    from python_amd.generators import uneven
    batch = uneven(scale=0.01)
    path = str(tmp_path / "uneven_1pct.hrb")
    wire.write_native(batch, path)
    got = wire.read_native(path)
    assert np.array_equal(got.A.view(np.uint64), batch.A.view(np.uint64))
    plan = hip.Plan.for_batch(got, mode="fp64", graph=True)
    A, B = hip.to_device(got)
    X, codes, _ = plan.solve(A, B, torch.zeros_like(B))
    torch.cuda.synchronize()
    Xo, co = oracle_lib.solve_batch(batch.n, batch.offA, batch.offV, batch.A, batch.B, nthreads=8)
    assert np.array_equal(codes.cpu().numpy(), co)
    Xh = X.cpu().numpy()
    for c in np.where(co != 4)[0]:
        o, n = batch.offV[c], batch.n[c]
        assert np.array_equal(Xh[o:o + n], Xo[o:o + n]), c
