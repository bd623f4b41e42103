"""The C-ABI library loads and exports every function include/hipla.h declares (no GPU calls)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hipla.h")
LIB = os.path.join(ROOT, "python_amd", "libhipla.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_declares_dropin_and_batched():
    names = declared_functions()
    for must in ("solve_matrix", "invert_matrix", "get_matrix_error_string", "hipla_init", "hipla_finalize",
                 "hipla_plan_create", "hipla_solve", "hipla_invert", "hipla_adjoint"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libhipla.so not built (python -m python_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_ctypes_binding_covers_header():
    from python_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_functions()
    L = _lib.lib()
    for name in _lib.SIGNATURES:
        assert getattr(L, name) is not None
    # pure host call, no device needed
    assert L.get_matrix_error_string(4) == b"Unsolvable matrix! Determinant is zero. Defaulting to no change"
    assert L.get_matrix_error_string(2) == b"some matrix rows failing relative error check"
