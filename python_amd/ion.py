"""The ionization balance of a batch of cells in lock step on the GPU.

Host-side mirror of ``matrix_ion_populations`` (source/matrix_ion.c:219-382) for many cells at
once: the reference iterates n_e one cell at a time, assembling the rate matrix with
``populate_ion_rate_matrix`` (:444-659) and calling ``solve_matrix`` (:276) once per iteration.
``IonIteration`` keeps every cell's per-ion coefficients in HBM and runs the whole fixed point
through ``hipla_ion_populations``: GPU assembly, FP64 GSL-order solves of the cells still
iterating, and the reference's floors, damping, convergence test and MAXITERATIONS on the GPU.

Per cell the result is what ``matrix_ion_populations`` leaves behind: the return value (0, or -1
for a singular solve or MAXITERATIONS, with the densities then untouched), the new absolute ion
densities and n_e = get_ne(densities) (source/saha.c:821-832), plus the iteration count and the
last solve_matrix code.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import IonRates, check, lib
from .generators import ION_COEF, IonBatch


class IonIteration:
    """A context over the cells of an :class:`IonBatch` (ion tables and coefficients copied to the
    GPU once).  ``populations()`` runs the iteration from the batch's densities (or the given
    ones) and returns torch tensors (density, ne, status, iters, codes)."""

    def __init__(self, batch: IonBatch, device="cuda"):
        import torch
        self.batch = batch
        self.ncells = batch.ncells
        dev = torch.device(device)

        def t(a, dtype):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(dev)
        self._t = {"elem": t(batch.elem, np.int32), "bits": t(batch.bits, np.int32),
                   "inner_beg": t(batch.inner_beg, np.int32), "inner_cnt": t(batch.inner_cnt, np.int32),
                   # at least one element so the pointers are valid even without records
                   "inner_rate": t(batch.inner_rate if len(batch.inner_rate) else np.zeros(1), np.float64),
                   "inner_yield": t(np.ravel(batch.inner_yield) if batch.inner_yield.size else np.zeros(10),
                                    np.float64)}
        for k in ION_COEF:
            self._t[k] = t(batch.coef[k], np.float64)
        self.rates = IonRates(**{k: ctypes.c_void_p(self._t[k].data_ptr()) for k in
                                 ("elem", "bits", "pi", "rr", "di", "dr", "qr", "cxr", "cxi", "inner_beg",
                                  "inner_cnt", "inner_rate", "inner_yield")}, has_cx=int(batch.has_cx))
        self.density0 = t(batch.density, np.float64)
        n = np.ascontiguousarray(batch.n, dtype=np.int32)
        offV = np.ascontiguousarray(batch.offV, dtype=np.int64)
        h = ctypes.c_void_p()
        check(lib().hipla_ion_create(ctypes.byref(h), self.ncells, n.ctypes.data_as(_lib._ip),
                                     offV.ctypes.data_as(_lib._i64p)), "hipla_ion_create")
        self._h = h
        self.density = torch.empty_like(self.density0)
        self.ne = torch.empty(self.ncells, dtype=torch.float64, device=dev)
        self.status = torch.empty(self.ncells, dtype=torch.int32, device=dev)
        self.iters = torch.empty(self.ncells, dtype=torch.int32, device=dev)
        self.codes = torch.empty(self.ncells, dtype=torch.int32, device=dev)

    def populations(self, density=None, stream=None):
        """Run every cell to convergence (or failure) starting from ``density`` (default: the
        batch's initial densities).  Synchronous on torch's current stream."""
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        self.density.copy_(self.density0 if density is None else density)
        check(lib().hipla_ion_populations(self._h, ctypes.byref(self.rates), ctypes.c_void_p(self.density.data_ptr()),
                                          ctypes.c_void_p(self.ne.data_ptr()), ctypes.c_void_p(self.status.data_ptr()),
                                          ctypes.c_void_p(self.iters.data_ptr()), ctypes.c_void_p(self.codes.data_ptr()),
                                          ctypes.c_void_p(int(s.cuda_stream))), "hipla_ion_populations")
        return self.density, self.ne, self.status, self.iters, self.codes

    def close(self):
        if self._h:
            lib().hipla_ion_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
