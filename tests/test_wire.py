"""The binary ragged-batch format (python_amd/wire.py, hipla_batch_* in libhipla.so): the numpy
and the native writers produce the same bytes, either reader returns the batch bit for bit
(signed zeros, infinities and NaN payloads included), and damaged files are refused."""
import os
import struct

import numpy as np
import pytest

from conftest import FIXTURES


def _batch(rng):
    from python_amd.generators import pack
    mats = [rng.standard_normal((k, k)) for k in (1, 3, 12, 30, 7)]
    rhs = [rng.standard_normal(k) for k in (1, 3, 12, 30, 7)]
    b = pack(mats, rhs)
    b.A[5] = -0.0
    b.A[6] = np.inf
    b.A[7] = np.frombuffer(struct.pack("<Q", 0x7FF8000000000123), "<f8")[0]   # NaN with a payload
    return b


def _same(a, b):
    assert np.array_equal(a.n, b.n) and np.array_equal(a.offA, b.offA) and np.array_equal(a.offV, b.offV)
    assert a.A.view(np.uint64).tolist() == b.A.view(np.uint64).tolist()
    assert a.B.view(np.uint64).tolist() == b.B.view(np.uint64).tolist()


def test_numpy_and_native_agree(tmp_path):
    from python_amd import wire
    b = _batch(np.random.default_rng(0))
    p1, p2 = str(tmp_path / "np.hrb"), str(tmp_path / "c.hrb")
    wire.save(b, p1)
    wire.write_native(b, p2)
    assert open(p1, "rb").read() == open(p2, "rb").read()
    for reader in (wire.load, wire.read_native):
        _same(reader(p1), b)


def test_without_rhs_and_empty(tmp_path):
    from python_amd import wire
    from python_amd.generators import RaggedBatch
    b = _batch(np.random.default_rng(1))
    p = str(tmp_path / "a.hrb")
    wire.write_native(b, p, with_b=False)
    got = wire.load(p)
    assert np.array_equal(got.A.view(np.uint64), b.A.view(np.uint64)) and not got.B.any()
    e = RaggedBatch(np.zeros(0, np.int32), np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0), np.zeros(0))
    wire.save(e, p)
    assert os.path.getsize(p) == 64
    assert len(wire.read_native(p).n) == 0


def test_damaged_files_are_refused(tmp_path):
    from python_amd import HiplaError, wire
    b = _batch(np.random.default_rng(2))
    p = str(tmp_path / "b.hrb")
    wire.save(b, p)
    raw = open(p, "rb").read()
    bad = {
        "magic": b"XIPLARB1" + raw[8:],
        "version": raw[:8] + struct.pack("<I", 2) + raw[12:],
        "truncated": raw[:-8],
        "trailing": raw + b"\0" * 8,
    }
    # a cell whose rows run past A: the last cell's offA moved beyond totA - n^2
    nc = len(b.n)
    off_pos = 64 + 4 * nc + (-4 * nc) % 8 + 8 * (nc - 1)
    bad["extent"] = raw[:off_pos] + struct.pack("<q", int(b.offA[-1]) + 1) + raw[off_pos + 8:]
    for name, data in bad.items():
        q = str(tmp_path / (name + ".hrb"))
        open(q, "wb").write(data)
        with pytest.raises(ValueError):
            wire.load(q)
        with pytest.raises(HiplaError):
            wire.read_native(q)


def test_reference_text_fixture_round_trip(tmp_path):
    from python_amd import wire
    from conftest import load_fixture
    d = os.path.join(FIXTURES, "matrix_ion")
    b = wire.from_text(os.path.join(d, "A.txt"), os.path.join(d, "b.txt"))
    n, a = load_fixture("matrix_ion", "A.txt")
    _, rhs = load_fixture("matrix_ion", "b.txt")
    assert b.n[0] == n and np.array_equal(b.A, a) and np.array_equal(b.B, rhs[:n] if rhs.size > n else rhs)
    p = str(tmp_path / "ion.hrb")
    wire.write_native(b, p)
    _same(wire.load(p), b)


def test_offsets_near_int64_max_are_refused(tmp_path):
    """An offset whose extent offA + n^2 wraps past INT64_MAX must not pass the extent checks of
    either reader, nor hipla_plan_create (which checks before it touches the device)."""
    import ctypes
    from python_amd import HiplaError, lib, wire
    b = _batch(np.random.default_rng(3))
    p = str(tmp_path / "w.hrb")
    wire.save(b, p)
    raw = open(p, "rb").read()
    nc = len(b.n)
    off_pos = 64 + 4 * nc + (-4 * nc) % 8 + 8 * (nc - 1)          # offA of the last cell
    q = str(tmp_path / "wrap.hrb")
    open(q, "wb").write(raw[:off_pos] + struct.pack("<q", 2 ** 63 - 2) + raw[off_pos + 8:])
    with pytest.raises(ValueError):
        wire.load(q)
    with pytest.raises(HiplaError):
        wire.read_native(q)
    # an order above the plans' limit is refused too
    n_pos = 64 + 4 * (nc - 1)
    q2 = str(tmp_path / "big.hrb")
    open(q2, "wb").write(raw[:n_pos] + struct.pack("<i", 70000) + raw[n_pos + 4:])
    with pytest.raises(ValueError):
        wire.load(q2)
    with pytest.raises(HiplaError):
        wire.read_native(q2)
    # the plan refuses the same offsets before any device work
    n = np.array([3, 4], dtype=np.int32)
    offA = np.array([0, 2 ** 63 - 2], dtype=np.int64)
    offV = np.array([0, 3], dtype=np.int64)
    h = ctypes.c_void_p()
    rc = lib().hipla_plan_create(ctypes.byref(h), 2, n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                 offA.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 offV.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 0, 0)
    assert rc == 5 and b"cell 1" in lib().hipla_last_error_detail()
