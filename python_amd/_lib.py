"""ctypes binding of libhipla.so (the C ABI declared in include/hipla.h).

The library is built in-tree by ``python_amd.build`` (hipcc, gfx950).  Loading fails loudly when
it is missing: there is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhipla.so")

_lib = None

c_int, c_int64, c_double, c_char_p, c_void_p = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_char_p, ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)

HIPLA_MODE_FP64 = 0
HIPLA_MODE_MIXED = 1
HIPLA_FLAG_GRAPH = 1
HIPLA_FLAG_KEEP_FACTORS = 2
HIPLA_FLAG_NO_FALLBACK = 4
HIPLA_FLAG_PER_CELL = 8


class CellStats(ctypes.Structure):
    """hipla_cell_stats (include/hipla.h)."""
    _fields_ = [("refine_steps", ctypes.c_int32), ("fallback", ctypes.c_int32),
                ("rel_correction", ctypes.c_double), ("backward_err", ctypes.c_double)]


CELL_STATS_BYTES = ctypes.sizeof(CellStats)


class IonRates(ctypes.Structure):
    """hipla_ion_rates (include/hipla.h): device pointers of the n_e iteration's per-ion tables."""
    _fields_ = [(k, c_void_p) for k in ("elem", "bits", "pi", "rr", "di", "dr", "qr", "cxr", "cxi", "inner_beg",
                                        "inner_cnt", "inner_rate", "inner_yield")] + [("has_cx", c_int)]

# name -> (restype, argtypes); every function declared in include/hipla.h
SIGNATURES = {
    "solve_matrix": (c_int, [_dp, _dp, c_int, _dp, c_int]),
    "invert_matrix": (c_int, [_dp, _dp, c_int]),
    "get_matrix_error_string": (c_char_p, [c_int]),
    "hipla_init": (c_int, []),
    "hipla_finalize": (c_int, []),
    "hipla_last_error_detail": (c_char_p, []),
    "hipla_alloc_stats": (c_int, [_i64p, _i64p]),
    "hipla_solve_matrix_noalloc": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "hipla_plan_create": (c_int, [ctypes.POINTER(c_void_p), c_int, _ip, _i64p, _i64p, c_int, c_int]),
    "hipla_plan_destroy": (c_int, [c_void_p]),
    "hipla_plan_set_refinement": (c_int, [c_void_p, c_int, c_double]),
    "hipla_solve": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hipla_invert": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hipla_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hipla_macro_emiss": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hipla_batch_write": (c_int, [c_char_p, c_int, _ip, _i64p, _i64p, _dp, _dp]),
    "hipla_batch_info": (c_int, [c_char_p, _ip, _i64p, _i64p, _ip]),
    "hipla_batch_read": (c_int, [c_char_p, _ip, _i64p, _i64p, _dp, _dp]),
    "hipla_ion_create": (c_int, [ctypes.POINTER(c_void_p), c_int, _ip, _i64p]),
    "hipla_ion_destroy": (c_int, [c_void_p]),
    "hipla_ion_populations": (c_int, [c_void_p, ctypes.POINTER(IonRates), c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p]),
    "hipla_plan_autotune": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hipla_plan_describe": (c_int, [c_void_p, ctypes.c_char_p, c_int]),
    "hipla_plan_set_variant": (c_int, [c_void_p, c_int, c_int]),
    "hipla_plan_time_groups": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, _dp, c_int]),
    "hipla_te_assemble": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "hipla_te_grad": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
}


class HiplaError(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HiplaError(f"{LIB_PATH} is missing: build it with `python -m python_amd.build` "
                             "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        detail = lib().hipla_last_error_detail()
        raise HiplaError(f"{what} failed (code {rc}): {detail.decode() if detail else ''}")
