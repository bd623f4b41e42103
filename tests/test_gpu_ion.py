"""GPU parity of the lock-step n_e iteration (hipla_ion_populations) with the oracle's per-cell
restatement of matrix_ion_populations (source/matrix_ion.c:219-382 around
populate_ion_rate_matrix :444-659): per cell the same return value, iteration count, last solve
code, n_e and ion densities, bit for bit (FP64 in the reference's order throughout)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _compare(batch, got, want):
    dens, ne, st, it, codes = (t.cpu().numpy() for t in got)
    wd, wne, wst, wit, wcodes = want
    assert np.array_equal(st, wst), np.where(st != wst)[0][:10]
    assert np.array_equal(it, wit), np.where(it != wit)[0][:10]
    assert np.array_equal(codes, wcodes), np.where(codes != wcodes)[0][:10]
    ok = wst == 0
    assert np.array_equal(ne[ok], wne[ok]), np.where(ne[ok] != wne[ok])[0][:10]
    # densities: new ones for converged cells, the input ones where the cell returned -1
    assert np.array_equal(dens.view(np.uint64), wd.view(np.uint64))


def test_uneven_batch_equals_oracle(hip, oracle_lib):
    """5% of the uneven ionization workload (all seven orders 12..256, Auger fill, charge
    exchange), including cells that hit MAXITERATIONS."""
    from python_amd.generators import ion_iteration_batch
    from python_amd.ion import IonIteration
    batch = ion_iteration_batch(scale=0.05)
    want = oracle_lib.ion_populations_batch(batch, nthreads=8)
    print("cells", batch.ncells, "iterations hist", np.bincount(want[3])[:20].tolist(), "failed",
          int((want[2] != 0).sum()), flush=True)
    it = IonIteration(batch)
    got = it.populations()
    _compare(batch, got, want)
    # a second run on the same context (graph replayed) gives the same answer
    got2 = it.populations()
    _compare(batch, got2, want)
    it.close()


def test_odd_orders_equal_oracle(hip, oracle_lib):
    """Odd orders (13 .. 255): the assembly's row flags and rates in LDS for matrices whose order is
    not a multiple of the grouped reads, through every route, bit for bit."""
    from python_amd.generators import ion_iteration_batch
    from python_amd.ion import IonIteration
    batch = ion_iteration_batch(scale=0.02, seed=3, sizes=(13, 31, 65, 97, 129, 201, 255))
    want = oracle_lib.ion_populations_batch(batch, nthreads=8)
    print("cells", batch.ncells, "orders", sorted(set(np.asarray(batch.n).tolist())), flush=True)
    it = IonIteration(batch)
    _compare(batch, it.populations(), want)
    it.close()


def test_restart_from_converged_and_singular_cells(hip, oracle_lib):
    """Restarting from the converged densities (most cells stop at once), and cells with no
    process into or out of H II (its row is zero: solve code 4, return -1, densities kept)."""
    import torch
    from python_amd.generators import ion_iteration_batch
    from python_amd.ion import IonIteration
    batch = ion_iteration_batch(scale=0.01, seed=11)
    for c in (0, 3):                        # hydrogen is ions 0-1 of every cell: zero H II row
        o = batch.offV[c]
        for k in ("pi", "di", "cxi"):
            batch.coef[k][o] = 0.0
        for k in ("rr", "qr", "dr"):
            batch.coef[k][o + 1] = 0.0
    want = oracle_lib.ion_populations_batch(batch, nthreads=8)
    assert want[2][0] == -1 and want[4][0] == 4
    it = IonIteration(batch)
    got = it.populations()
    _compare(batch, got, want)
    restart = batch.subset(range(batch.ncells))
    restart.density = got[0].cpu().numpy().copy()
    want2 = oracle_lib.ion_populations_batch(restart, nthreads=8)
    got2 = it.populations(torch.from_numpy(restart.density).cuda())
    print("restart iterations hist", np.bincount(want2[3]).tolist(), flush=True)
    _compare(restart, got2, want2)
    it.close()
