"""Host-side mirror of the reference's dense solve / inverse interface on top of libhipla.

Drop-in functions (same names, argument meaning and return codes as source/matrix_cpu.c:273-349):

    code = solve_matrix(a_matrix, b_matrix, size, x_matrix, nplasma)   # x_matrix filled in place
    code = invert_matrix(matrix, inverted_matrix, num_rows)             # inverted_matrix in place
    text = get_matrix_error_string(code)

Batched form (no reference counterpart; the reference loops over cells one call at a time,
source/wind_updates2d.c:89-153 and source/macro_accelerate.c:1052-1061): a :class:`Plan` over a
ragged batch of cells, executed on torch device tensors on torch's current stream.
"""
from __future__ import annotations

import ctypes
import json

import numpy as np

from . import _lib
from ._lib import (CELL_STATS_BYTES, HIPLA_FLAG_GRAPH, HIPLA_FLAG_KEEP_FACTORS, HIPLA_FLAG_NO_FALLBACK,
                   HIPLA_FLAG_PER_CELL, HIPLA_MODE_FP64, HIPLA_MODE_MIXED, HiplaError, check, lib)

_MODES = {"fp64": HIPLA_MODE_FP64, "mixed": HIPLA_MODE_MIXED}


def _f64(a, name):
    arr = np.asarray(a)
    if arr.dtype != np.float64 or not arr.flags["C_CONTIGUOUS"]:
        raise TypeError(f"{name} must be a C-contiguous float64 array")
    return arr


def solve_matrix(a_matrix, b_matrix, size: int, x_matrix, nplasma: int = -1) -> int:
    """solve_matrix(a, b, size, x, nplasma) of source/matrix_cpu.c:309-321 on the GPU.

    Returns 0, 2, 3 (residual warnings, x written), 4 (singular, x untouched) or 5 (library
    failure).  ``a_matrix`` is row-major size x size; it is only read."""
    a = _f64(a_matrix, "a_matrix")
    b = _f64(b_matrix, "b_matrix")
    x = _f64(x_matrix, "x_matrix")
    if a.size < size * size or b.size < size or x.size < size:
        raise ValueError("arrays smaller than size")
    return lib().solve_matrix(a.ctypes.data_as(_lib._dp), b.ctypes.data_as(_lib._dp), int(size),
                              x.ctypes.data_as(_lib._dp), int(nplasma))


def invert_matrix(matrix, inverted_matrix, num_rows: int) -> int:
    """invert_matrix(m, inv, n) of source/matrix_cpu.c:337-349 on the GPU: 0, 1 (singular) or 5."""
    m = _f64(matrix, "matrix")
    inv = _f64(inverted_matrix, "inverted_matrix")
    if m.size < num_rows * num_rows or inv.size < num_rows * num_rows:
        raise ValueError("arrays smaller than num_rows^2")
    return lib().invert_matrix(m.ctypes.data_as(_lib._dp), inv.ctypes.data_as(_lib._dp), int(num_rows))


def get_matrix_error_string(error_code: int) -> str:
    return lib().get_matrix_error_string(int(error_code)).decode()


def last_error_detail() -> str:
    d = lib().hipla_last_error_detail()
    return d.decode() if d else ""


def alloc_stats() -> tuple[int, int]:
    """(device allocations + streams + events made so far, device bytes held now) of libhipla."""
    n, live = ctypes.c_int64(), ctypes.c_int64()
    check(lib().hipla_alloc_stats(ctypes.byref(n), ctypes.byref(live)), "hipla_alloc_stats")
    return n.value, live.value


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else int(t.data_ptr()))


def _stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(int(s.cuda_stream))


class Plan:
    """A ragged batch plan (hipla_plan_create) bound to torch device tensors.

    ``n``, ``offA``, ``offV`` describe the cells (host arrays).  ``mode`` is "fp64" (GSL
    operation order) or "mixed" (equilibrated FP32 LU + FP64 refinement + FP64 fallback).
    """

    def __init__(self, n, offA, offV=None, mode: str = "fp64", graph: bool = True, keep_factors: bool = False,
                 per_cell: bool = False, fallback: bool = True):
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.offA = np.ascontiguousarray(offA, dtype=np.int64)
        self.offV = np.ascontiguousarray(offV if offV is not None else np.zeros_like(self.offA), dtype=np.int64)
        self.ncells = int(self.n.shape[0])
        self.mode = mode
        flags = 0
        if graph:
            flags |= HIPLA_FLAG_GRAPH
        if keep_factors:
            flags |= HIPLA_FLAG_KEEP_FACTORS
        if per_cell:
            flags |= HIPLA_FLAG_PER_CELL
        if not fallback:
            flags |= HIPLA_FLAG_NO_FALLBACK
        self.flags = flags
        h = ctypes.c_void_p()
        check(lib().hipla_plan_create(ctypes.byref(h), self.ncells, self.n.ctypes.data_as(_lib._ip),
                                      self.offA.ctypes.data_as(_lib._i64p), self.offV.ctypes.data_as(_lib._i64p),
                                      _MODES[mode], flags), "hipla_plan_create")
        self._h = h
        self.generation = 0    # solves issued on this plan (the adjoint uses the latest one's factors)
        self.totA = int((self.offA + self.n.astype(np.int64) ** 2).max()) if self.ncells else 0
        self.totV = int((self.offV + self.n).max()) if self.ncells else 0

    @classmethod
    def for_batch(cls, batch, **kw):
        return cls(batch.n, batch.offA, batch.offV, **kw)

    def set_variant(self, group: int, variant: int):
        """FP64 kernel shape of one size-class group (hipla_plan_set_variant)."""
        check(lib().hipla_plan_set_variant(self._h, int(group), int(variant)), "hipla_plan_set_variant")

    def close(self):
        if getattr(self, "_h", None):
            lib().hipla_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_refinement(self, max_refine: int = 5, tol: float = 1e-12):
        check(lib().hipla_plan_set_refinement(self._h, int(max_refine), float(tol)), "hipla_plan_set_refinement")

    def describe(self) -> dict:
        buf = ctypes.create_string_buffer(1 << 16)
        check(lib().hipla_plan_describe(self._h, buf, len(buf)), "hipla_plan_describe")
        return json.loads(buf.value.decode())

    def _check_tensor(self, t, numel, dtype, name):
        import torch
        if t is None:
            raise ValueError(name + " is required")
        if t.device.type != "cuda" or t.dtype != dtype or not t.is_contiguous() or t.numel() < numel:
            raise ValueError(f"{name} must be a contiguous {dtype} CUDA tensor with >= {numel} elements")

    def solve(self, A, B, X=None, codes=None, stats=None, stream=None):
        """Solve every cell; returns (X, codes, stats) device tensors (stats is a uint8 blob of
        hipla_cell_stats records, or None)."""
        import torch
        self._check_tensor(A, self.totA, torch.float64, "A")
        self._check_tensor(B, self.totV, torch.float64, "B")
        if X is None:
            X = torch.empty_like(B)
        if codes is None:
            codes = torch.empty(self.ncells, dtype=torch.int32, device=B.device)
        self._check_tensor(X, self.totV, torch.float64, "X")
        self._check_tensor(codes, self.ncells, torch.int32, "codes")
        if stats is True:
            stats = torch.zeros(self.ncells * CELL_STATS_BYTES, dtype=torch.uint8, device=B.device)
        elif stats is False:
            stats = None
        check(lib().hipla_solve(self._h, _ptr(A), _ptr(B), _ptr(X), _ptr(codes), _ptr(stats), _stream_handle(stream)),
              "hipla_solve")
        self.generation += 1
        return X, codes, stats

    def invert(self, A, Inv=None, codes=None, stats=None, stream=None):
        import torch
        self._check_tensor(A, self.totA, torch.float64, "A")
        if Inv is None:
            Inv = torch.empty_like(A)
        if codes is None:
            codes = torch.empty(self.ncells, dtype=torch.int32, device=A.device)
        if stats is True:
            stats = torch.zeros(self.ncells * CELL_STATS_BYTES, dtype=torch.uint8, device=A.device)
        elif stats is False:
            stats = None
        check(lib().hipla_invert(self._h, _ptr(A), _ptr(Inv), _ptr(codes), _ptr(stats), _stream_handle(stream)),
              "hipla_invert")
        return Inv, codes, stats

    def adjoint(self, X, G, Lam=None, GradA=None, stream=None):
        import torch
        if Lam is None:
            Lam = torch.empty_like(G)
        check(lib().hipla_adjoint(self._h, _ptr(X), _ptr(G), _ptr(Lam), _ptr(GradA), _stream_handle(stream)),
              "hipla_adjoint")
        return Lam, GradA

    def macro_emiss(self, N, rdiag, absorbed, emit, out=None, codes=None, stream=None):
        """Macro-atom emissivities out_j = (N^-T a)_j R_jj e_j per cell, N^-1 never formed
        (hipla_macro_emiss; the reference's invert_matrix + B = N^-1 R + a^T B contraction,
        source/macro_accelerate.c:315-334, source/macro_gen_f.c:490-506).  N has A's layout;
        rdiag, absorbed, emit and out have B's.  The plan needs keep_factors=True, mode "fp64".
        Returns (out, codes); a cell with code 4 (singular N) is not written."""
        import torch
        self._check_tensor(N, self.totA, torch.float64, "N")
        for t, name in ((rdiag, "rdiag"), (absorbed, "absorbed"), (emit, "emit")):
            self._check_tensor(t, self.totV, torch.float64, name)
        if out is None:
            out = torch.zeros_like(absorbed)
        if codes is None:
            codes = torch.empty(self.ncells, dtype=torch.int32, device=absorbed.device)
        self._check_tensor(out, self.totV, torch.float64, "out")
        self._check_tensor(codes, self.ncells, torch.int32, "codes")
        check(lib().hipla_macro_emiss(self._h, _ptr(N), _ptr(rdiag), _ptr(absorbed), _ptr(emit), _ptr(out),
                                      _ptr(codes), _stream_handle(stream)), "hipla_macro_emiss")
        return out, codes

    def time_groups(self, A, B, X, codes, stats=None, reps: int = 5, classes_only: bool = False):
        """Per-group average ms (HIP events on the plan's stream): (groups, fallback pass, whole),
        or the groups alone with ``classes_only`` (nothing else is launched then)."""
        ng = len(self.describe()["groups"])
        out = np.zeros(ng if classes_only else ng + 2)
        check(lib().hipla_plan_time_groups(self._h, _ptr(A), _ptr(B), _ptr(X), _ptr(codes), _ptr(stats), int(reps),
                                           out.ctypes.data_as(_lib._dp), len(out)), "hipla_plan_time_groups")
        return out

    def autotune(self, A, B, X, codes, stream=None):
        check(lib().hipla_plan_autotune(self._h, _ptr(A), _ptr(B), _ptr(X), _ptr(codes), _stream_handle(stream)),
              "hipla_plan_autotune")


def stats_to_numpy(stats):
    """Decode a hipla_cell_stats blob into a numpy structured array."""
    dt = np.dtype([("refine_steps", np.int32), ("fallback", np.int32), ("rel_correction", np.float64),
                   ("backward_err", np.float64)])
    raw = stats.cpu().numpy() if hasattr(stats, "cpu") else np.asarray(stats)
    return raw.view(dt)


def to_device(batch, device="cuda"):
    import torch
    return (torch.from_numpy(batch.A).to(device), torch.from_numpy(batch.B).to(device))


def solve_batch(batch, mode: str = "fp64", graph: bool = True, device="cuda", **kw):
    """Solve a generators.RaggedBatch on the GPU; returns numpy (X, codes, stats-or-None).

    X starts as zeros, so the solution of a singular cell (code 4, x not written) reads 0.
    """
    import torch
    plan = Plan.for_batch(batch, mode=mode, graph=graph, **kw)
    A, B = to_device(batch, device)
    X, codes, stats = plan.solve(A, B, torch.zeros_like(B), stats=(mode == "mixed"))
    torch.cuda.synchronize()
    out = (X.cpu().numpy(), codes.cpu().numpy(), stats_to_numpy(stats) if stats is not None else None)
    plan.close()
    return out
