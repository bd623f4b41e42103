"""CPU oracle for the per-cell dense solve / inverse path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` import
this package, and only as the checker.  The product (``python_amd``, ``libhipla.so``) never
imports, links or calls it.

The C restatement lives in ``hipla_ref.c`` (each function cites the reference file:line it
follows: ``source/matrix_cpu.c`` and GSL 2.6 ``linalg/lu.c``, ``invtri.c``, ``trimult.c`` and
the cblas kernels).  This module loads ``liboracle.so`` built from it by ``oracle/Makefile``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_szp = ctypes.POINTER(ctypes.c_size_t)
_up = ctypes.POINTER(ctypes.c_uint)


class RefIonCell(ctypes.Structure):
    """ref_ion_cell of hipla_ref.h: per-ion tables and coefficients of the n_e iteration."""
    _fields_ = [("elem", _ip), ("bits", _ip),
                ("pi", _dp), ("rr", _dp), ("di", _dp), ("dr", _dp), ("qr", _dp), ("cxr", _dp), ("cxi", _dp),
                ("inner_beg", _ip), ("inner_cnt", _ip), ("inner_rate", _dp), ("inner_yield", _dp),
                ("has_cx", ctypes.c_int)]


def build(force: bool = False) -> str:
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.ref_lu_decomp.argtypes = [_dp, ctypes.c_int, _szp, _up, _ip]
        L.ref_lu_decomp_rightlooking.argtypes = [_dp, ctypes.c_int, _szp, _up, _ip]
        L.ref_lu_solve.argtypes = [_dp, ctypes.c_int, _szp, _dp, _dp]
        L.ref_residual_code.argtypes = [_dp, ctypes.c_int, _dp, _dp]
        L.ref_solve_matrix.argtypes = [_dp, _dp, ctypes.c_int, _dp]
        L.ref_lu_invx.argtypes = [_dp, ctypes.c_int, _szp]
        L.ref_invert_matrix.argtypes = [_dp, _dp, ctypes.c_int]
        L.ref_lu_solve_transpose.argtypes = [_dp, ctypes.c_int, _szp, _dp, _dp]
        L.ref_solve_dd_refined.argtypes = [_dp, _dp, ctypes.c_int, _dp, ctypes.c_int]
        L.ref_solve_batch.argtypes = [ctypes.c_int, _ip, _i64p, _i64p, _dp, _dp, _dp, _ip, ctypes.c_int]
        L.ref_invert_batch.argtypes = [ctypes.c_int, _ip, _i64p, _dp, _dp, _ip, ctypes.c_int]
        L.ref_populate_ion_rate_matrix.argtypes = [ctypes.c_int, ctypes.POINTER(RefIonCell), ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.ref_populate_ion_rate_matrix.restype = None
        L.ref_ion_populations.argtypes = [ctypes.c_int, ctypes.POINTER(RefIonCell), _dp, _dp, _ip, _ip]
        L.ref_ion_populations_batch.argtypes = [ctypes.c_int, _ip, _i64p, ctypes.POINTER(RefIonCell), _dp, _dp,
                                                _ip, _ip, _ip, ctypes.c_int]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def lu_decomp(a, rightlooking: bool = False):
    """GSL LU_decomp: returns (LU, perm, ipiv, signum); perm as gsl_permutation.data."""
    a = _c(a).copy()
    n = a.shape[0]
    perm = np.zeros(n, dtype=np.uintp)
    ipiv = np.zeros(n, dtype=np.uint32)
    s = ctypes.c_int(0)
    f = lib().ref_lu_decomp_rightlooking if rightlooking else lib().ref_lu_decomp
    f(_d(a), n, perm.ctypes.data_as(_szp), ipiv.ctypes.data_as(_up), ctypes.byref(s))
    return a, perm.astype(np.int64), ipiv.astype(np.int64), s.value


def lu_solve(lu, perm, b):
    lu = _c(lu)
    n = lu.shape[0]
    p = np.ascontiguousarray(perm, dtype=np.uintp)
    x = np.zeros(n)
    st = lib().ref_lu_solve(_d(lu), n, p.ctypes.data_as(_szp), _d(_c(b)), _d(x))
    return st, x


def solve_matrix(a, b):
    """cpu_solve_matrix semantics: returns (code, x); x is zeros when code == 4."""
    a = _c(a).copy()
    b = _c(b)
    n = b.shape[0]
    x = np.zeros(n)
    code = lib().ref_solve_matrix(_d(a), _d(b), n, _d(x))
    return code, x


def residual_code(a, b, x):
    a, b, x = _c(a), _c(b), _c(x)
    return lib().ref_residual_code(_d(a), b.shape[0], _d(b), _d(x))


def invert_matrix(m):
    """cpu_invert_matrix semantics: returns (code, inverse or None)."""
    m = _c(m).copy()
    n = m.shape[0]
    inv = np.zeros((n, n))
    code = lib().ref_invert_matrix(_d(m), _d(inv), n)
    return code, (inv if code == 0 else None)


def macro_emiss(N, rdiag, absorbed, emit):
    """The reference's macro-atom emissivities of one cell, in its order: (code, out or None).

    N^-1 by the GSL-order inverse (``invert_matrix``, source/macro_accelerate.c:315), then
    B_ij = N^-1_ij * R_jj (:324-334), then for each absorbing level i in ascending order
    emiss_j += a_i * B_ij over every j (source/macro_gen_f.c:490-499, the k-packet row last,
    :503-506), and finally emiss_j *= emit_j (:504, :507).  ``absorbed`` is
    (matom_abs[0..n-2], kpkt_abs), ``emit`` (level_emit[0..n-2], kpkt_emit); out[n-1] is
    kpkt_emiss."""
    code, inv = invert_matrix(N)
    if code != 0:
        return code, None
    rdiag, a, e = _c(rdiag), _c(absorbed), _c(emit)
    B = inv * rdiag[None, :]
    out = np.zeros(N.shape[0])
    for i in range(N.shape[0]):
        out += a[i] * B[i, :]
    return 0, out * e


def lu_solve_transpose(lu, perm, g):
    lu = _c(lu)
    n = lu.shape[0]
    p = np.ascontiguousarray(perm, dtype=np.uintp)
    lam = np.zeros(n)
    st = lib().ref_lu_solve_transpose(_d(lu), n, p.ctypes.data_as(_szp), _d(_c(g)), _d(lam))
    return st, lam


def solve_dd_refined(a, b, iters: int = 4):
    a, b = _c(a), _c(b)
    n = b.shape[0]
    x = np.zeros(n)
    st = lib().ref_solve_dd_refined(_d(a), _d(b), n, _d(x), iters)
    return st, x


def solve_batch(nsz, offA, offV, A, B, nthreads: int = 1):
    """Ragged batch of cpu_solve_matrix calls; returns (X, codes)."""
    nsz = np.ascontiguousarray(nsz, dtype=np.int32)
    offA = np.ascontiguousarray(offA, dtype=np.int64)
    offV = np.ascontiguousarray(offV, dtype=np.int64)
    A, B = _c(A), _c(B)
    X = np.zeros_like(B)
    codes = np.zeros(len(nsz), dtype=np.int32)
    lib().ref_solve_batch(len(nsz), nsz.ctypes.data_as(_ip), offA.ctypes.data_as(_i64p),
                          offV.ctypes.data_as(_i64p), _d(A), _d(B), _d(X), codes.ctypes.data_as(_ip),
                          int(nthreads))
    return X, codes


def invert_batch(nsz, offA, A, nthreads: int = 1):
    nsz = np.ascontiguousarray(nsz, dtype=np.int32)
    offA = np.ascontiguousarray(offA, dtype=np.int64)
    A = _c(A)
    Inv = np.zeros_like(A)
    codes = np.zeros(len(nsz), dtype=np.int32)
    lib().ref_invert_batch(len(nsz), nsz.ctypes.data_as(_ip), offA.ctypes.data_as(_i64p), _d(A), _d(Inv),
                           codes.ctypes.data_as(_ip), int(nthreads))
    return Inv, codes


def _ion_struct(t, keep):
    """RefIonCell over the arrays of ``t`` (an IonBatch, or a dict with the same fields); the
    contiguous copies are appended to ``keep`` so they outlive the call."""
    def arr(v, dtype):
        a = np.ascontiguousarray(v, dtype=dtype)
        keep.append(a)
        return a.ctypes.data_as(_ip if dtype == np.int32 else _dp)
    get = (lambda k: getattr(t, k)) if not isinstance(t, dict) else (lambda k: t[k])
    coef = get("coef")
    return RefIonCell(arr(get("elem"), np.int32), arr(get("bits"), np.int32),
                      *[arr(coef[k], np.float64) for k in ("pi", "rr", "di", "dr", "qr", "cxr", "cxi")],
                      arr(get("inner_beg"), np.int32), arr(get("inner_cnt"), np.int32),
                      arr(get("inner_rate"), np.float64), arr(np.ravel(get("inner_yield")), np.float64),
                      int(get("has_cx")))


def populate_ion_rate_matrix(cell, xne, nh1, nh2):
    """populate_ion_rate_matrix (source/matrix_ion.c:444-659) of one cell: returns (A, b).
    ``cell`` is a dict of one cell's per-ion arrays (see ``ion_cell``)."""
    keep = []
    st = _ion_struct(cell, keep)
    n = len(cell["elem"])
    a = np.zeros((n, n))
    b = np.zeros(n)
    lib().ref_populate_ion_rate_matrix(n, ctypes.byref(st), float(xne), float(nh1), float(nh2), _d(a), _d(b))
    return a, b


def ion_populations(cell):
    """The n_e iteration of matrix_ion_populations (source/matrix_ion.c:219-382) for one cell:
    returns (status, density, ne, iterations, last solve code); density is the input density when
    status != 0."""
    keep = []
    st = _ion_struct(cell, keep)
    dens = _c(cell["density"]).copy()
    ne = ctypes.c_double(0.0)
    it = ctypes.c_int(0)
    code = ctypes.c_int(0)
    status = lib().ref_ion_populations(len(dens), ctypes.byref(st), _d(dens), ctypes.byref(ne), ctypes.byref(it),
                                       ctypes.byref(code))
    return status, dens, ne.value, it.value, code.value


def ion_cell(batch, c):
    """One cell of an IonBatch as a dict of its own arrays (inner records re-indexed from 0)."""
    o, n = int(batch.offV[c]), int(batch.n[c])
    sl = slice(o, o + n)
    beg, cnt = batch.inner_beg[sl], batch.inner_cnt[sl]
    r0 = int(beg[0]) if n else 0
    r1 = int(beg[-1] + cnt[-1]) if n else 0
    return {"elem": batch.elem[sl], "bits": batch.bits[sl], "coef": {k: v[sl] for k, v in batch.coef.items()},
            "inner_beg": (beg - r0).astype(np.int32), "inner_cnt": cnt, "inner_rate": batch.inner_rate[r0:r1],
            "inner_yield": batch.inner_yield[r0:r1], "density": batch.density[sl], "has_cx": batch.has_cx}


def ion_populations_batch(batch, nthreads: int = 1):
    """The iteration over every cell of an IonBatch: returns (density, ne, status, iters, codes)."""
    keep = []
    st = _ion_struct(batch, keep)
    nsz = np.ascontiguousarray(batch.n, dtype=np.int32)
    offV = np.ascontiguousarray(batch.offV, dtype=np.int64)
    dens = _c(batch.density).copy()
    nc = len(nsz)
    ne = np.zeros(nc)
    status = np.zeros(nc, dtype=np.int32)
    iters = np.zeros(nc, dtype=np.int32)
    codes = np.zeros(nc, dtype=np.int32)
    lib().ref_ion_populations_batch(nc, nsz.ctypes.data_as(_ip), offV.ctypes.data_as(_i64p), ctypes.byref(st),
                                    _d(dens), _d(ne), status.ctypes.data_as(_ip), iters.ctypes.data_as(_ip),
                                    codes.ctypes.data_as(_ip), int(nthreads))
    return dens, ne, status, iters, codes
