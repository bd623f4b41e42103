"""python_amd -- MI355X-native per-cell dense solve / inverse (libhipla, gfx950 HIP kernels).

Drop-in for the reference's ``solve_matrix`` / ``invert_matrix`` path (source/matrix_cpu.c) with a
batched ragged API, an equilibrated mixed-precision solver and an adjoint for d(x)/d(T_e).
"""
from ._lib import HiplaError, LIB_PATH, lib  # noqa: F401
from .solver import (Plan, alloc_stats, get_matrix_error_string, invert_matrix, last_error_detail,  # noqa: F401
                     solve_batch, solve_matrix, stats_to_numpy, to_device)

__all__ = ["Plan", "alloc_stats", "solve_matrix", "invert_matrix", "get_matrix_error_string", "last_error_detail", "solve_batch",
           "stats_to_numpy", "to_device", "HiplaError", "lib", "LIB_PATH"]
