"""Binary ragged-batch files: one layout for the C harness, the torch front end and any exchange.

The format (declared with ``hipla_batch_write`` in include/hipla.h, written and read natively by
python_amd/csrc/hipla_wire.hip) keeps a whole ragged batch bit-exact::

    64-byte header: b"HIPLARB1", uint32 version 1, uint32 flags (bit 0: B present),
                    int64 ncells, int64 totA, int64 totV, 24 zero bytes
    int32 n[ncells], zero padding to a multiple of 8 bytes
    int64 offA[ncells], int64 offV[ncells]
    float64 A[totA] (cell c row-major at offA[c]), float64 B[totV] (cell c at offV[c]) if present

all little-endian.  The reference's tests exchange one matrix at a time as "%e" text (the order
n, then the row-major values, 7 significant digits; source/tests/test_matrix.c:52-62,113-125);
``from_text`` reads that format into a one-cell batch.

``save`` / ``load`` are numpy restatements of the format; ``write_native`` / ``read_native`` go
through the library's C functions (the same file either way).
"""
from __future__ import annotations

import ctypes
import struct

import numpy as np

from ._lib import _dp, _i64p, _ip, check, lib
from .generators import RaggedBatch

MAGIC = b"HIPLARB1"
VERSION = 1
_HEADER = struct.Struct("<8sIIqqq24x")


MAX_ORDER = 65536        # the largest n a plan accepts (hipla_plan_create)
_I64_MAX = np.iinfo(np.int64).max


def _sizes(batch):
    """(totA, totV) of a batch; refuses n outside 1..MAX_ORDER and offsets whose extent would not
    fit in int64 (compared with the room left, so nothing wraps)."""
    n = np.asarray(batch.n, dtype=np.int64)
    if len(n) == 0:
        return 0, 0
    offA = np.asarray(batch.offA, dtype=np.int64)
    offV = np.asarray(batch.offV, dtype=np.int64)
    if (n < 1).any() or (n > MAX_ORDER).any() or (offA < 0).any() or (offV < 0).any():
        raise ValueError(f"every cell needs 1 <= n <= {MAX_ORDER} and non-negative offsets")
    if (offA > _I64_MAX - n * n).any() or (offV > _I64_MAX - n).any():
        raise ValueError("a cell's offset puts its extent beyond int64")
    return int((offA + n * n).max()), int((offV + n).max())


def _extents_ok(n, offA, offV, totA, totV):
    """Every cell inside [0, totA) x [0, totV), with no sum that can wrap."""
    n = n.astype(np.int64)
    return not ((n < 1).any() or (n > MAX_ORDER).any() or (offA < 0).any() or (offV < 0).any()
                or (n * n > totA).any() or (offA > totA - n * n).any() or (n > totV).any()
                or (offV > totV - n).any())


def save(batch: RaggedBatch, path: str, with_b: bool = True) -> None:
    totA, totV = _sizes(batch)
    nc = len(batch.n)
    with open(path, "wb") as f:
        f.write(_HEADER.pack(MAGIC, VERSION, 1 if with_b else 0, nc, totA, totV))
        f.write(np.ascontiguousarray(batch.n, dtype="<i4").tobytes())
        f.write(b"\0" * ((-4 * nc) % 8))
        f.write(np.ascontiguousarray(batch.offA, dtype="<i8").tobytes())
        f.write(np.ascontiguousarray(batch.offV, dtype="<i8").tobytes())
        f.write(np.ascontiguousarray(batch.A[:totA], dtype="<f8").tobytes())
        if with_b:
            f.write(np.ascontiguousarray(batch.B[:totV], dtype="<f8").tobytes())


def load(path: str) -> RaggedBatch:
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < _HEADER.size:
        raise ValueError(f"{path}: not a ragged-batch file")
    magic, version, flags, nc, totA, totV = _HEADER.unpack_from(raw)
    if magic != MAGIC or version != VERSION or min(nc, totA, totV) < 0:
        raise ValueError(f"{path}: not a version-{VERSION} ragged-batch file")
    pos = _HEADER.size
    npad = 4 * nc + (-4 * nc) % 8
    expect = pos + npad + 16 * nc + 8 * totA + (8 * totV if flags & 1 else 0)
    if len(raw) != expect:
        raise ValueError(f"{path}: {len(raw)} bytes, the header implies {expect}")
    n = np.frombuffer(raw, "<i4", nc, pos).astype(np.int32)
    pos += npad
    offA = np.frombuffer(raw, "<i8", nc, pos).astype(np.int64)
    offV = np.frombuffer(raw, "<i8", nc, pos + 8 * nc).astype(np.int64)
    pos += 16 * nc
    A = np.frombuffer(raw, "<f8", totA, pos).astype(np.float64)
    pos += 8 * totA
    B = np.frombuffer(raw, "<f8", totV, pos).astype(np.float64) if flags & 1 else np.zeros(totV)
    batch = RaggedBatch(n, offA, offV, A, B)
    if nc and (not _extents_ok(n, offA, offV, totA, totV) or _sizes(batch) != (totA, totV)):
        raise ValueError(f"{path}: a cell lies outside the batch")
    return batch


def write_native(batch: RaggedBatch, path: str, with_b: bool = True) -> None:
    n = np.ascontiguousarray(batch.n, dtype=np.int32)
    oa = np.ascontiguousarray(batch.offA, dtype=np.int64)
    ov = np.ascontiguousarray(batch.offV, dtype=np.int64)
    A = np.ascontiguousarray(batch.A, dtype=np.float64)
    B = np.ascontiguousarray(batch.B, dtype=np.float64)
    check(lib().hipla_batch_write(path.encode(), len(n), n.ctypes.data_as(_ip), oa.ctypes.data_as(_i64p),
                                  ov.ctypes.data_as(_i64p), A.ctypes.data_as(_dp),
                                  B.ctypes.data_as(_dp) if with_b else None), "hipla_batch_write")


def read_native(path: str) -> RaggedBatch:
    nc, totA, totV, hb = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
    check(lib().hipla_batch_info(path.encode(), ctypes.byref(nc), ctypes.byref(totA), ctypes.byref(totV),
                                 ctypes.byref(hb)), "hipla_batch_info")
    n = np.zeros(nc.value, dtype=np.int32)
    oa = np.zeros(nc.value, dtype=np.int64)
    ov = np.zeros(nc.value, dtype=np.int64)
    A = np.zeros(totA.value)
    B = np.zeros(totV.value)
    check(lib().hipla_batch_read(path.encode(), n.ctypes.data_as(_ip), oa.ctypes.data_as(_i64p),
                                 ov.ctypes.data_as(_i64p), A.ctypes.data_as(_dp), B.ctypes.data_as(_dp)),
          "hipla_batch_read")
    return RaggedBatch(n, oa, ov, A, B)


def from_text(a_path: str, b_path: str | None = None) -> RaggedBatch:
    """A one-cell batch from the reference's text fixtures: the order n, then n*n row-major
    values (and, for b, the order n then n values)."""
    with open(a_path) as f:
        va = f.read().split()
    n = int(va[0])
    A = np.array(va[1:1 + n * n], dtype=np.float64)
    if A.size != n * n:
        raise ValueError(f"{a_path}: {A.size} values for n = {n}")
    B = np.zeros(n)
    if b_path:
        with open(b_path) as f:
            vb = f.read().split()
        if int(vb[0]) != n:
            raise ValueError(f"{b_path}: order {vb[0]}, matrix order {n}")
        B = np.array(vb[1:1 + n], dtype=np.float64)
    return RaggedBatch(np.array([n], dtype=np.int32), np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64), A, B)
