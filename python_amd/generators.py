"""Seeded synthetic cell batches for the per-cell dense solve / inverse path.

The reference builds its matrices from atomic physics (source/matrix_ion.c:444-659
``populate_ion_rate_matrix``; source/macro_accelerate.c:30-348 ``calc_matom_matrix``).  These
generators restate their *structure* with closed-form random rates:

* Ionization balance (``matrix_ion.c``): per element a block of ionization stages; off-diagonal
  A[i+1, i] = rate i -> i+1 (photo/collisional ionization), A[i, i+1] = n_e * alpha (recombination
  i+1 -> i); every diagonal is minus its column sum (population conservation of the rate
  equations); then the row of each element's first ion is replaced by the abundance constraint
  (ones over that element's ions, b = abundance), as ``populate_ion_rate_matrix`` does
  (matrix_ion.c:633-655).  b is zero except on those rows.  With ``auger=True`` (the uneven
  batches) inner-shell ionization adds A[i+d, i] for jumps d = 1..10 (matrix_ion.c:600-616).
* Macro-atom inverse (``macro_accelerate.c:269-284``): N = I - Q with Q substochastic.
* Macro-atom level populations (``macro_gov.c:366-440``, ``macro_pops_fill_rate_matrix``
  ``:535-749``): per element a rate matrix over its levels with M[to, from] += rate and
  M[from, from] -= rate for every transition, row 0 replaced by ones, b = e_0.

All randomness comes from ``numpy.random.default_rng(seed)`` so the same seed gives the same
bytes on any machine.  Layout of a batch: cell c is the row-major n_c x n_c matrix at
``A[offA[c]:offA[c]+n_c*n_c]`` and its right-hand side at ``B[offV[c]:offV[c]+n_c]``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# (name, number of ionization stages incl. neutral, abundance relative to H)
ELEMENTS = [
    ("H", 2, 1.0), ("He", 3, 0.1), ("C", 7, 3e-4), ("N", 8, 1e-4), ("O", 9, 7e-4),
    ("Ne", 11, 1e-4), ("Mg", 13, 4e-5), ("Si", 15, 3.5e-5), ("S", 17, 1.5e-5), ("Fe", 27, 3e-5),
    ("Na", 12, 2e-6), ("Al", 14, 3e-6), ("Ar", 19, 4e-6), ("Ca", 21, 2e-6), ("Ni", 29, 2e-6),
]
SLACK_ABUNDANCE = 1e-5


def element_layout(n: int):
    """Split n unknowns into element blocks (start, stages, abundance) plus one slack row.

    Elements are taken in ELEMENTS order (repeating the list if n needs more); the last block
    is truncated so the blocks fill n - 1 rows exactly; the final row is a one-stage slack
    element.  n = 30 gives H(2) He(3) C(7) N(8) O(9) + slack, the layout of the 900-cell case.
    """
    if n < 1:
        raise ValueError("n must be >= 1")
    blocks = []
    s = 0
    room = n - 1
    i = 0
    while room > 0:
        name, m, ab = ELEMENTS[i % len(ELEMENTS)]
        m = min(m, room)
        blocks.append((s, m, ab))
        s += m
        room -= m
        i += 1
    blocks.append((s, 1, SLACK_ABUNDANCE))
    return blocks


def _logu(rng, lo, hi, size):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size))


AUGER_MAX_JUMP = 10     # inner-shell ionization ejects at most 10 extra electrons (n_elec <= 11)
AUGER_SCALE = 1e-2      # inner-shell rates relative to the ion's outer-shell rate range


def auger_sites(blocks):
    """(column, first row, jumps) of every ion with an inner shell that can autoionize: an ion with
    n_elec >= 3 bound electrons, whose inner-shell ionization lands d = 1 .. min(n_elec, 11) - 1
    stages higher (source/matrix_ion.c:600-616: rate_matrix[ion + d][ion] += rate * yield[d-1],
    n_elec = z - istate + 1 capped at 11).  Element blocks of m stages hold z = m - 1, so stage
    k (istate k + 1) has n_elec = m - 1 - k."""
    out = []
    for (s, m, ab) in blocks:
        for k in range(m):
            n_elec = min(m - 1 - k, AUGER_MAX_JUMP + 1)
            if n_elec >= 3:
                out.append((s + k, s + k + 1, n_elec - 1))
    return out


def _add_auger(rng, A, blocks, rates):
    """Add the inner-shell terms: rates[:, t] for site t, spread over its jumps by yields that sum
    to one (a seeded Dirichlet draw), into column `col` below the diagonal."""
    for t, (col, r0, nj) in enumerate(auger_sites(blocks)):
        y = rng.gamma(1.0, 1.0, (A.shape[0], nj))
        y /= y.sum(axis=1, keepdims=True)
        A[:, r0:r0 + nj, col] += rates[:, t:t + 1] * y


def ion_matrices(rng, ncells: int, n: int, ion_range=(1e-12, 1e2), ne_range=(1e6, 1e14),
                 alpha_range=(1e-14, 1e-10), auger: bool = False):
    """ncells ionization-balance matrices of order n: returns (A [ncells,n,n], b [ncells,n]).

    auger=True adds inner-shell (Auger) ionization: every ion with three or more electrons gets
    a rate log-uniform over AUGER_SCALE * ion_range, split over jumps of +1 .. +10 stages, so
    the element blocks fill below the sub-diagonal as the reference's do."""
    A = np.zeros((ncells, n, n))
    b = np.zeros((ncells, n))
    ne = _logu(rng, *ne_range, ncells)
    blocks = element_layout(n)
    for (s, m, ab) in blocks:
        if m > 1:
            k = np.arange(m - 1)
            ion = _logu(rng, *ion_range, (ncells, m - 1))
            rec = ne[:, None] * _logu(rng, *alpha_range, (ncells, m - 1))
            A[:, s + k + 1, s + k] = ion
            A[:, s + k, s + k + 1] = rec
    if auger:
        sites = auger_sites(blocks)
        lo, hi = AUGER_SCALE * ion_range[0], AUGER_SCALE * ion_range[1]
        _add_auger(rng, A, blocks, _logu(rng, lo, hi, (ncells, len(sites))))
    d = np.arange(n)
    A[:, d, d] = -A.sum(axis=1)
    for (s, m, ab) in blocks:
        A[:, s, :] = 0.0
        A[:, s, s:s + m] = 1.0
        b[:, s] = ab
    return A, b


WIDE_LOG_RATES = (-20.0, 4.0)      # log10 of the wide rate range, 1e-20 .. 1e4 s^-1
WIDE_LOG_KAPPA = (6.0, 14.0)       # log10 of the condition-number range of the wide batch


def _contracted_ion_matrix(n, u_ion, u_rec, u_aug, yields, s, m):
    """Ionization matrices whose log10 rates are m + s * (u + 8), clipped to the wide range: u are
    raw log10 draws over WIDE_LOG_RATES (centre -8), s in [0, 1] contracts them toward the
    cell's level m.  Inner-shell rates (u_aug, split over their jumps by ``yields``) sit two
    decades lower.  One row per element is the abundance constraint (ones)."""
    c = u_ion.shape[0]
    A = np.zeros((c, n, n))
    lo, hi = WIDE_LOG_RATES
    j = 0
    blocks = element_layout(n)
    for (st, mm, ab) in blocks:
        if mm > 1:
            k = np.arange(mm - 1)
            li = np.clip(m[:, None] + s[:, None] * (u_ion[:, j:j + mm - 1] + 8.0), lo, hi)
            lr = np.clip(m[:, None] + s[:, None] * (u_rec[:, j:j + mm - 1] + 8.0), lo, hi)
            A[:, st + k + 1, st + k] = 10.0 ** li
            A[:, st + k, st + k + 1] = 10.0 ** lr
            j += mm - 1
    la = np.clip(m[:, None] + s[:, None] * (u_aug + 8.0) + np.log10(AUGER_SCALE), lo, hi)
    for t, (col, r0, nj) in enumerate(auger_sites(blocks)):
        A[:, r0:r0 + nj, col] += (10.0 ** la[:, t:t + 1]) * yields[t]
    d = np.arange(n)
    A[:, d, d] = -A.sum(axis=1)
    for (st, mm, ab) in blocks:
        A[:, st, :] = 0.0
        A[:, st, st:st + mm] = 1.0
    return A


def ion_matrices_conditioned(rng, ncells: int, n: int, log_kappa=WIDE_LOG_KAPPA, iters: int = 10):
    """ncells ionization matrices of order n with rates inside 1e-20..1e4 and a 2-norm condition
    number log-uniform over 10^log_kappa (the wide-range precision workload).

    Raw rates (ionization i -> i+1, recombination n_e alpha, i+1 -> i, and inner-shell ionization
    with Auger jumps of +1 .. +10 stages) are drawn log-uniform over the whole 1e-20..1e4 range; those raw matrices have kappa of 1e17..1e25 and many are
    numerically singular in FP64.  Each cell therefore draws a target kappa and a rate level m
    (log10 uniform over -8..0), and its log-rates are contracted toward m by a factor s found by
    bisection on numpy's 2-norm condition number.  Cells that miss the range are drawn again.
    Returns (A, b, log10 kappa per cell)."""
    blocks = element_layout(n)
    nr = sum(mm - 1 for (_, mm, _) in blocks if mm > 1)
    sites = auger_sites(blocks)
    A = np.zeros((ncells, n, n))
    lk = np.full(ncells, np.nan)
    todo = np.arange(ncells)
    for _round in range(20):
        if len(todo) == 0:
            break
        c = len(todo)
        u_ion = rng.uniform(*WIDE_LOG_RATES, (c, nr))
        u_rec = rng.uniform(*WIDE_LOG_RATES, (c, nr))
        u_aug = rng.uniform(*WIDE_LOG_RATES, (c, len(sites)))
        yields = []
        for (_col, _r0, nj) in sites:
            y = rng.gamma(1.0, 1.0, (c, nj))
            yields.append(y / y.sum(axis=1, keepdims=True))
        m = rng.uniform(-8.0, 0.0, c)

        def build(sv):
            return _contracted_ion_matrix(n, u_ion, u_rec, u_aug, yields, sv, m)

        k0 = np.log10(np.linalg.cond(build(np.zeros(c))))
        t = np.maximum(rng.uniform(*log_kappa, c), k0 + 0.25)
        slo, shi = np.zeros(c), np.ones(c)
        for _ in range(iters):
            s = 0.5 * (slo + shi)
            up = np.log10(np.linalg.cond(build(s))) < t
            slo, shi = np.where(up, s, slo), np.where(up, shi, s)
        Ac = build(0.5 * (slo + shi))
        kc = np.log10(np.linalg.cond(Ac))
        ok = (kc >= log_kappa[0]) & (kc <= log_kappa[1])
        A[todo[ok]] = Ac[ok]
        lk[todo[ok]] = kc[ok]
        todo = todo[~ok]
    if len(todo):
        raise RuntimeError(f"{len(todo)} cells of order {n} missed the condition-number range")
    b = np.zeros((ncells, n))
    for (st, mm, ab) in blocks:
        b[:, st] = ab
    return A, b, lk


def ion_matrices_te(rng, ncells: int, n: int, te=None):
    """Temperature-dependent ionization matrices of order n (the adjoint / Newton case).

    Ionization i -> i+1: C * T^0.5 * exp(-E / kT), E ~ U[5, 500] eV, C log-uniform.
    Recombination i+1 -> i: alpha0 * (T / 1e4)^-0.7, alpha0 log-uniform (n_e folded in).
    Returns a dict with T, the rate parameters, A(T), dA/dT and b.
    """
    if te is None:
        te = _logu(rng, 3e3, 1e7, ncells)
    blocks = element_layout(n)
    params = []
    for (s, m, ab) in blocks:
        if m > 1:
            C = _logu(rng, 1e-8, 1e-2, (ncells, m - 1))
            E = rng.uniform(5.0, 500.0, (ncells, m - 1))
            a0 = _logu(rng, 1e-6, 1e2, (ncells, m - 1))
            params.append((s, m, ab, C, E, a0))
        else:
            params.append((s, m, ab, None, None, None))
    A, dA, b = te_assemble(n, blocks, params, te)
    return {"te": te, "blocks": blocks, "params": params, "A": A, "dA": dA, "b": b}


KB_EV = 8.617333262e-5  # Boltzmann constant, eV / K


def te_assemble(n, blocks, params, te):
    """A(T), dA/dT and b for ion_matrices_te parameters at temperatures te [ncells]."""
    ncells = te.shape[0]
    A = np.zeros((ncells, n, n))
    dA = np.zeros((ncells, n, n))
    b = np.zeros((ncells, n))
    T = te[:, None]
    for (s, m, ab, C, E, a0) in params:
        if m > 1:
            k = np.arange(m - 1)
            x = E / (KB_EV * T)
            ion = C * np.sqrt(T) * np.exp(-x)
            dion = ion * (0.5 / T + x / T)
            rec = a0 * (T / 1e4) ** -0.7
            drec = rec * (-0.7 / T)
            A[:, s + k + 1, s + k] = ion
            A[:, s + k, s + k + 1] = rec
            dA[:, s + k + 1, s + k] = dion
            dA[:, s + k, s + k + 1] = drec
    d = np.arange(n)
    A[:, d, d] = -A.sum(axis=1)
    dA[:, d, d] = -dA.sum(axis=1)
    for (s, m, ab, *_rest) in params:
        A[:, s, :] = 0.0
        dA[:, s, :] = 0.0
        A[:, s, s:s + m] = 1.0
        b[:, s] = ab
    return A, dA, b


def macro_q(rng, ncells: int, n: int, density=(0.10, 0.40), val_range=(1e-8, 1.0), eps_range=(1e-6, 1e-1)):
    """ncells macro-atom matrices N = I - Q of order n (Q substochastic, zero diagonal)."""
    N = np.zeros((ncells, n, n))
    for c in range(ncells):
        dens = rng.uniform(*density, n)
        mask = rng.random((n, n)) < dens[:, None]
        np.fill_diagonal(mask, False)
        Q = np.where(mask, _logu(rng, *val_range, (n, n)), 0.0)
        rs = Q.sum(axis=1)
        eps = _logu(rng, *eps_range, n)
        scale = np.where(rs > 0, (1.0 - eps) / np.where(rs > 0, rs, 1.0), 0.0)
        Q *= scale[:, None]
        N[c] = -Q
        N[c][np.arange(n), np.arange(n)] = 1.0
    return N


def macro_pops_matrices(rng, ncells: int, n: int, rate_range=(1e-6, 1e8), density=0.15):
    """ncells macro-atom level-population systems of order n: returns (A [ncells,n,n], b [ncells,n]).

    Every transition from -> to adds its rate at M[to, from] and takes it from M[from, from]
    (source/macro_gov.c:593-597, 628-629, 661-662, 693-708, 737-738), so every column sums to
    zero; adjacent levels are always linked both ways (bound-bound / bound-free between
    neighbouring levels and stages) and random further pairs are linked at ``density``.  Row 0
    is then replaced by the normalisation (all ones) and b = e_0 (macro_gov.c:405-408, 436)."""
    A = np.zeros((ncells, n, n))
    b = np.zeros((ncells, n))
    b[:, 0] = 1.0
    for c in range(ncells):
        M = np.zeros((n, n))
        link = rng.random((n, n)) < density
        np.fill_diagonal(link, False)
        idx = np.arange(n - 1)
        link[idx, idx + 1] = link[idx + 1, idx] = True
        frm, to = np.nonzero(link)
        rates = _logu(rng, *rate_range, len(frm))
        np.add.at(M, (to, frm), rates)
        np.add.at(M, (frm, frm), -rates)
        M[0, :] = 1.0
        A[c] = M
    return A, b


def macro_emiss_vectors(rng, N):
    """The per-cell vectors of the macro-atom emissivity contraction for one N = I - Q
    (source/macro_accelerate.c:253-298, source/macro_gen_f.c:482-506): diag(R) with N's row sums
    (the normalisation the reference checks, macro_accelerate.c:286-298: R_ii - sum_j N_ij = 0),
    absorbed energies (matom_abs, kpkt_abs last) spread over six decades, emission fractions in
    [0, 1)."""
    n = N.shape[0]
    return N.sum(axis=1), _logu(rng, 1e-6, 1.0, n), rng.random(n)


@dataclass
class RaggedBatch:
    """A ragged batch of square FP64 cells packed row-major back to back."""
    n: np.ndarray                 # int32 [ncells]
    offA: np.ndarray              # int64 [ncells]
    offV: np.ndarray              # int64 [ncells]
    A: np.ndarray                 # float64 [sum n^2]
    B: np.ndarray                 # float64 [sum n]
    meta: dict = field(default_factory=dict)

    @property
    def ncells(self) -> int:
        return int(self.n.shape[0])

    def cell(self, c: int):
        n = int(self.n[c])
        a = self.A[self.offA[c]:self.offA[c] + n * n].reshape(n, n)
        b = self.B[self.offV[c]:self.offV[c] + n]
        return a, b

    def subset(self, cells):
        cells = list(cells)
        mats = [self.cell(c) for c in cells]
        return pack([m[0] for m in mats], [m[1] for m in mats])

    @property
    def bytes_solve(self) -> int:
        """Algorithmic bytes of a solve: A and b in, x out (FP64)."""
        n = self.n.astype(np.int64)
        return int((8 * (n * n + 2 * n)).sum())

    @property
    def flops_solve(self) -> float:
        n = self.n.astype(np.float64)
        return float((2.0 / 3.0 * n ** 3 + 2.0 * n ** 2).sum())


def pack(mats, rhs=None) -> RaggedBatch:
    ns = np.array([m.shape[0] for m in mats], dtype=np.int32)
    offA = np.zeros(len(mats), dtype=np.int64)
    offV = np.zeros(len(mats), dtype=np.int64)
    if len(mats):
        offA[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
        offV[1:] = np.cumsum(ns.astype(np.int64))[:-1]
    A = np.concatenate([np.ascontiguousarray(m, dtype=np.float64).ravel() for m in mats]) if mats else np.zeros(0)
    if rhs is None:
        B = np.zeros(int(ns.sum()))
    else:
        B = np.concatenate([np.ascontiguousarray(r, dtype=np.float64).ravel() for r in rhs])
    return RaggedBatch(ns, offA, offV, A, B)


def pack_groups(groups, order) -> RaggedBatch:
    """Pack {n: (A[k,n,n], b[k,n])} groups into a batch whose cell c has order order[c]."""
    ns = np.asarray(order, dtype=np.int32)
    offA = np.zeros(len(ns), dtype=np.int64)
    offV = np.zeros(len(ns), dtype=np.int64)
    offA[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
    offV[1:] = np.cumsum(ns.astype(np.int64))[:-1]
    A = np.empty(int((ns.astype(np.int64) ** 2).sum()))
    B = np.empty(int(ns.astype(np.int64).sum()))
    cursor = {k: 0 for k in groups}
    for c, n in enumerate(ns):
        Ag, bg = groups[int(n)]
        i = cursor[int(n)]
        cursor[int(n)] = i + 1
        A[offA[c]:offA[c] + n * n] = Ag[i].ravel()
        B[offV[c]:offV[c] + n] = bg[i]
    return RaggedBatch(ns, offA, offV, A, B)


# ------------------------------------------------------------------------------------------
# The workloads
# ------------------------------------------------------------------------------------------
UNEVEN_SIZES = (12, 30, 64, 96, 128, 200, 256)
UNEVEN_COUNTS = (3000, 2500, 1500, 1000, 1000, 500, 500)


def plain_small(ncells: int = 900, seed: int = 1) -> RaggedBatch:
    """900 cells of the n = 30 ionization matrix (H, He, C, N, O + slack), seed 1."""
    rng = np.random.default_rng(seed)
    A, b = ion_matrices(rng, ncells, 30)
    batch = pack_groups({30: (A, b)}, [30] * ncells)
    batch.meta = {"workload": "plain_small", "seed": seed}
    return batch


def uneven(ncells: int = 10000, seed: int = 2, wide: bool = False, scale: float = 1.0) -> RaggedBatch:
    """10,000 cells, n in {12..256} with counts 3000/2500/1500/1000/1000/500/500, shuffled, with
    inner-shell (Auger) fill below each element block's sub-diagonal.

    wide=True is the precision workload (seed 3): rates inside 1e-20..1e4 and condition numbers
    log-uniform over 1e6..1e14 (``ion_matrices_conditioned``); the per-cell log10 kappa is kept in
    ``meta["log10_kappa"]`` (cell order) with its histogram in ``meta["kappa_hist"]``.
    ``scale`` shrinks every count proportionally (for quick tests)."""
    rng = np.random.default_rng(seed)
    counts = [max(1, int(round(c * scale * ncells / 10000))) for c in UNEVEN_COUNTS]
    order = np.concatenate([np.full(c, n, dtype=np.int32) for n, c in zip(UNEVEN_SIZES, counts)])
    rng.shuffle(order)
    groups, kappas = {}, {}
    for n, c in zip(UNEVEN_SIZES, counts):
        if wide:
            A, b, lk = ion_matrices_conditioned(rng, c, n)
            groups[n] = (A, b)
            kappas[n] = lk
        else:
            groups[n] = ion_matrices(rng, c, n, auger=True)
    batch = pack_groups(groups, order)
    batch.meta = {"workload": "uneven_wide" if wide else "uneven", "seed": seed, "counts": dict(zip(UNEVEN_SIZES, counts))}
    if wide:
        lk = np.empty(len(order))
        for n in kappas:
            lk[order == n] = kappas[n]          # pack_groups takes each size's cells in order
        nz = batch.A[batch.A != 0.0]
        rates = np.abs(nz[np.abs(nz) != 1.0])
        batch.meta.update({
            "log10_kappa": lk,
            "kappa_hist": {f"1e{e}..1e{e + 1}": int(((lk >= e) & (lk < e + 1)).sum()) for e in range(6, 14)},
            "log10_kappa_range": [float(lk.min()), float(lk.max())],
            "abs_entry_range": [float(rates.min()), float(rates.max())]})
    return batch


def macro_inverse(ncells: int = 2000, seed: int = 4, sizes=(256, 512, 1024)) -> RaggedBatch:
    """2,000 macro-atom cells N = I - Q, n in {256, 512, 1024} in equal thirds, seed 4."""
    rng = np.random.default_rng(seed)
    per = [ncells // len(sizes) + (1 if i < ncells % len(sizes) else 0) for i in range(len(sizes))]
    order = np.concatenate([np.full(c, n, dtype=np.int32) for n, c in zip(sizes, per)])
    rng.shuffle(order)
    groups = {n: (macro_q(rng, c, n), np.zeros((c, n))) for n, c in zip(sizes, per)}
    batch = pack_groups(groups, order)
    batch.meta = {"workload": "macro_inverse", "seed": seed}
    return batch


# ------------------------------------------------------------------------------------------
# The ionization-balance iteration (source/matrix_ion.c:219-382 around populate_ion_rate_matrix
# :444-659): per-ion rate coefficients instead of assembled matrices
# ------------------------------------------------------------------------------------------
ION_FIRST, ION_MAXST, ION_DI, ION_DR, ION_CXI = 1 << 16, 1 << 17, 1 << 18, 1 << 19, 1 << 20
ION_MAXJUMP = 10                 # inner-shell electron yields per record (prob[0..9])
SLACK_Z = 30                     # the one-stage slack element: not hydrogen, no inner shells
ION_COEF = ("pi", "rr", "di", "dr", "qr", "cxr", "cxi")


@dataclass
class IonBatch:
    """Per-cell ion tables and rate coefficients of the n_e iteration, ragged over cells.

    Cell c has n[c] ions at offV[c] in every per-ion array: ``elem`` (element id, unique within
    the cell), ``bits`` (istate | z << 8 | ION_* flags), the coefficients of ``ION_COEF``
    (pi_rates, rr_rates, di/dr/qrecomb_coeffs, charge-exchange recombination and ionization
    coefficients), ``inner_beg`` / ``inner_cnt`` (the ion's inner-shell records, global indices
    into ``inner_rate`` and ``inner_yield`` [nrec, 10]), and ``density`` (absolute ion densities
    the iteration starts from).  ``offA`` lays out the n x n rate matrices back to back."""
    n: np.ndarray
    offV: np.ndarray
    offA: np.ndarray
    elem: np.ndarray
    bits: np.ndarray
    coef: dict
    inner_beg: np.ndarray
    inner_cnt: np.ndarray
    inner_rate: np.ndarray
    inner_yield: np.ndarray
    density: np.ndarray
    has_cx: int = 1
    meta: dict = field(default_factory=dict)

    @property
    def ncells(self) -> int:
        return int(self.n.shape[0])

    @property
    def totV(self) -> int:
        return int(self.offV[-1] + self.n[-1]) if self.ncells else 0

    def subset(self, cells) -> "IonBatch":
        cells = np.asarray(list(cells), dtype=np.int64)
        ns = self.n[cells].astype(np.int32)
        offV = np.zeros(len(cells), dtype=np.int64)
        offA = np.zeros(len(cells), dtype=np.int64)
        offV[1:] = np.cumsum(ns.astype(np.int64))[:-1]
        offA[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
        idx = np.concatenate([np.arange(self.offV[c], self.offV[c] + self.n[c]) for c in cells])
        beg, cnt = self.inner_beg[idx], self.inner_cnt[idx]
        recs = np.concatenate([np.arange(b, b + k) for b, k in zip(beg, cnt)]).astype(np.int64) \
            if cnt.sum() else np.zeros(0, dtype=np.int64)
        nbeg = np.zeros(len(idx), dtype=np.int32)
        nbeg[1:] = np.cumsum(cnt)[:-1]
        return IonBatch(ns, offV, offA, self.elem[idx].copy(), self.bits[idx].copy(),
                        {k: v[idx].copy() for k, v in self.coef.items()}, nbeg, cnt.copy(),
                        self.inner_rate[recs].copy(), self.inner_yield[recs].copy(), self.density[idx].copy(),
                        self.has_cx, dict(self.meta))


def _ion_table(n):
    """(elem, z, istate, first, last) per ion of the element layout of order n."""
    blocks = element_layout(n)
    elem, z, ist, first, last = [], [], [], [], []
    for i, (s, m, ab) in enumerate(blocks):
        zz = SLACK_Z if i == len(blocks) - 1 else ELEMENTS[i % len(ELEMENTS)][1] - 1
        for k in range(m):
            elem.append(i)
            z.append(zz)
            ist.append(k + 1)
            first.append(k == 0)
            last.append(k == m - 1)
    return (np.array(elem), np.array(z), np.array(ist), np.array(first), np.array(last), blocks)


def ion_iteration_cells(rng, ncells: int, n: int, nh_range=(1e6, 1e12)):
    """Rate coefficients of ncells cells with n ions each (element layout of ``element_layout``).

    Photoionization rates log-uniform over 1e-8..1e1 s^-1, radiative / dielectronic recombination
    coefficients over 1e-14..1e-10 cm^3 s^-1, three-body over 1e-30..1e-26 cm^6 s^-1, collisional
    ionization over 1e-14..1e-9, charge exchange over 1e-12..1e-9 (recombination on every ion
    above hydrogen that is not a neutral; ionization on a fifth of the metal ions below their top
    stage); 80 % of ions carry a collisional-ionization rate and 70 % a dielectronic one.  Ions
    with n_elec = min(z - istate + 1, 11) >= 3 whose Auger jumps stay inside their element get
    one inner-shell record (rate over 1e-10..1e-2, yields summing to one; a tenth get a second
    record).  Hydrogen density log-uniform over nh_range, element densities nh * abundance,
    initial stage fractions a Dirichlet draw.  Returns a dict of per-ion arrays [ncells, n] and
    the records."""
    elem, z, ist, first, last, blocks = _ion_table(n)
    ab = np.array([blocks[e][2] for e in elem])
    bits = ist | (z << 8) | np.where(first, ION_FIRST, 0) | np.where(last, ION_MAXST, 0)
    bits = np.broadcast_to(bits, (ncells, n)).copy()
    bits |= np.where(rng.random((ncells, n)) < 0.8, ION_DI, 0)
    bits |= np.where(rng.random((ncells, n)) < 0.7, ION_DR, 0)
    metal_mid = (z != 1) & (z != SLACK_Z) & ~last
    bits |= np.where(metal_mid[None, :] & (rng.random((ncells, n)) < 0.2), ION_CXI, 0)
    coef = {
        "pi": _logu(rng, 1e-8, 1e1, (ncells, n)),
        "rr": _logu(rng, 1e-14, 1e-10, (ncells, n)),
        "di": _logu(rng, 1e-14, 1e-9, (ncells, n)),
        "dr": _logu(rng, 1e-14, 1e-10, (ncells, n)),
        "qr": _logu(rng, 1e-30, 1e-26, (ncells, n)),
        "cxr": np.where((z != 1) & ~first, 1.0, 0.0)[None, :] * _logu(rng, 1e-12, 1e-9, (ncells, n)),
        "cxi": np.where(bits & ION_CXI, 1.0, 0.0) * _logu(rng, 1e-12, 1e-9, (ncells, n)),
    }
    # inner-shell records: jumps d = 1 .. n_elec - 1 must stay inside the ion's element
    block_end = np.array([blocks[e][0] + blocks[e][1] - 1 for e in elem])
    n_elec = np.minimum(z - ist + 1, 11)
    ok = (n_elec >= 3) & (np.arange(n) + n_elec - 1 <= block_end) & (z != SLACK_Z)
    nrec = np.where(ok[None, :], 1 + (rng.random((ncells, n)) < 0.1), 0).astype(np.int32)
    tot = int(nrec.sum())
    rate = _logu(rng, 1e-10, 1e-2, tot)
    yld = np.zeros((tot, ION_MAXJUMP))
    owner = np.repeat(np.tile(np.arange(n), ncells), nrec.ravel())
    for j in range(tot):
        k = n_elec[owner[j]] - 1
        y = rng.gamma(1.0, 1.0, k)
        yld[j, :k] = y / y.sum()
    nh = _logu(rng, *nh_range, ncells)
    frac = np.zeros((ncells, n))
    for (s, m, _ab) in blocks:
        g = rng.gamma(1.0, 1.0, (ncells, m))
        frac[:, s:s + m] = g / g.sum(axis=1, keepdims=True)
    density = nh[:, None] * ab[None, :] * frac
    return {"elem": np.broadcast_to(elem, (ncells, n)).astype(np.int32), "bits": bits.astype(np.int32),
            "coef": coef, "nrec": nrec, "inner_rate": rate, "inner_yield": yld, "density": density, "nh": nh}


def ion_iteration_batch(ncells: int = 10000, seed: int = 6, scale: float = 1.0, sizes=UNEVEN_SIZES,
                        counts=UNEVEN_COUNTS) -> IonBatch:
    """The lock-step n_e iteration workload: the uneven size mix (n in {12..256}, counts
    3000/2500/1500/1000/1000/500/500 for 10,000 cells, shuffled), coefficients from
    ``ion_iteration_cells``, seed 6.  ``scale`` shrinks every count."""
    rng = np.random.default_rng(seed)
    counts = [max(1, int(round(c * scale * ncells / 10000))) for c in counts]
    order = np.concatenate([np.full(c, n, dtype=np.int32) for n, c in zip(sizes, counts)])
    rng.shuffle(order)
    per = {n: ion_iteration_cells(rng, c, n) for n, c in zip(sizes, counts)}
    return _pack_ion(per, order, {"workload": "ion_iteration", "seed": seed, "counts": dict(zip(sizes, counts))})


def _pack_ion(per, order, meta) -> IonBatch:
    ns = np.asarray(order, dtype=np.int32)
    nc = len(ns)
    offV = np.zeros(nc, dtype=np.int64)
    offA = np.zeros(nc, dtype=np.int64)
    offV[1:] = np.cumsum(ns.astype(np.int64))[:-1]
    offA[1:] = np.cumsum(ns.astype(np.int64) ** 2)[:-1]
    totV = int(ns.astype(np.int64).sum())
    elem = np.empty(totV, dtype=np.int32)
    bits = np.empty(totV, dtype=np.int32)
    coef = {k: np.empty(totV) for k in ION_COEF}
    density = np.empty(totV)
    cnt = np.empty(totV, dtype=np.int32)
    rates, ylds = [], []
    cursor = {n: 0 for n in per}
    # each size's record list is split per cell in the order its cells are taken
    rec_start = {n: np.concatenate([[0], np.cumsum(per[n]["nrec"].sum(axis=1))]) for n in per}
    for c, n in enumerate(ns):
        n = int(n)
        d = per[n]
        i = cursor[n]
        cursor[n] = i + 1
        sl = slice(offV[c], offV[c] + n)
        elem[sl] = d["elem"][i]
        bits[sl] = d["bits"][i]
        for k in ION_COEF:
            coef[k][sl] = d["coef"][k][i]
        density[sl] = d["density"][i]
        cnt[sl] = d["nrec"][i]
        r0, r1 = rec_start[n][i], rec_start[n][i + 1]
        rates.append(d["inner_rate"][r0:r1])
        ylds.append(d["inner_yield"][r0:r1])
    beg = np.zeros(totV, dtype=np.int32)
    beg[1:] = np.cumsum(cnt)[:-1]
    return IonBatch(ns, offV, offA, elem, bits, coef, beg, cnt,
                    np.concatenate(rates) if rates else np.zeros(0),
                    np.concatenate(ylds).reshape(-1, ION_MAXJUMP) if ylds else np.zeros((0, ION_MAXJUMP)),
                    density, 1, meta)
