"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

ORACLE — test infrastructure only: tests, the smoke check and bench.py's
CPU-baseline leg import this module as the checker. The product (qz_amd)
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)


def _declare(L):
    L.orc_adler32.restype = C.c_uint32
    L.orc_adler32.argtypes = [u8p, C.c_int64]
    L.orc_suffix_array.restype = C.c_int
    L.orc_suffix_array.argtypes = [u8p, i32p, C.c_int32]
    L.orc_bwt_sample_rate.restype = C.c_int32
    L.orc_bwt_sample_rate.argtypes = [C.c_int32]
    L.orc_bwt_aux.restype = C.c_int32
    L.orc_bwt_aux.argtypes = [u8p, u8p, C.c_int32, C.c_int32, i32p]
    L.orc_bsc_bwt_encode.restype = C.c_int32
    L.orc_bsc_bwt_encode.argtypes = [u8p, C.c_int32, u8p, i32p]
    L.orc_unbwt.restype = C.c_int32
    L.orc_unbwt.argtypes = [u8p, u8p, C.c_int32, C.c_int32]
    L.orc_qlfc_transform.restype = C.c_int32
    L.orc_qlfc_transform.argtypes = [u8p, C.c_int32, u8p, u8p, i32p, u8p, i32p]
    L.orc_qlfc_inverse.restype = C.c_int32
    L.orc_qlfc_inverse.argtypes = [u8p, C.c_int32, u8p, i32p, C.c_int32, u8p, C.c_int32]
    L.orc_num_segments.restype = C.c_int32
    L.orc_num_segments.argtypes = [C.c_int32]
    L.orc_split_segments.restype = None
    L.orc_split_segments.argtypes = [u8p, C.c_int32, C.c_int32, i32p, i32p]


def _u8(a: np.ndarray):
    return a.ctypes.data_as(u8p)


def _i32(a: np.ndarray):
    return a.ctypes.data_as(i32p)


def as_u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8)
    return np.frombuffer(bytes(data), dtype=np.uint8).copy()


def adler32(data) -> int:
    a = as_u8(data)
    return int(lib().orc_adler32(_u8(a), a.size))


def suffix_array(data) -> np.ndarray:
    a = as_u8(data)
    sa = np.empty(max(a.size, 1), dtype=np.int32)
    rc = lib().orc_suffix_array(_u8(a), _i32(sa), a.size)
    assert rc == 0, rc
    return sa[: a.size]


def sample_rate(n: int) -> int:
    return int(lib().orc_bwt_sample_rate(n))


def bwt_aux(data):
    """(U, primary, I) with libsais bwt_aux conventions; I[k] = 1 + rank(k*r)."""
    a = as_u8(data)
    n = a.size
    r = sample_rate(n)
    U = np.empty(n, dtype=np.uint8)
    I = np.zeros((n - 1) // r + 1, dtype=np.int32)
    p = lib().orc_bwt_aux(_u8(a), _u8(U), n, r, _i32(I))
    assert p >= 1, p
    return U, int(p), I


def bsc_bwt_encode(data):
    """libbsc bsc_bwt_encode: (U, primary, trailer indexes)."""
    a = as_u8(data).copy()
    cnt = np.zeros(1, dtype=np.uint8)
    idx = np.zeros(256, dtype=np.int32)
    p = lib().orc_bsc_bwt_encode(_u8(a), a.size, _u8(cnt), _i32(idx))
    assert p >= 1, p
    return a, int(p), idx[: int(cnt[0])].copy()


def unbwt(U, primary: int) -> np.ndarray:
    u = as_u8(U)
    T = np.empty(u.size, dtype=np.uint8)
    rc = lib().orc_unbwt(_u8(u), _u8(T), u.size, primary)
    assert rc == 0, rc
    return T


def qlfc_transform(data):
    """(run symbols, run ranks, run lengths, table)."""
    a = as_u8(data)
    n = a.size
    sym = np.empty(max(n, 1), dtype=np.uint8)
    rank = np.empty(max(n, 1), dtype=np.uint8)
    ln = np.empty(max(n, 1), dtype=np.int32)
    table = np.empty(256, dtype=np.uint8)
    nsym = np.zeros(1, dtype=np.int32)
    runs = lib().orc_qlfc_transform(_u8(a), n, _u8(sym), _u8(rank), _i32(ln), _u8(table), _i32(nsym))
    return sym[:runs].copy(), rank[:runs].copy(), ln[:runs].copy(), table[: int(nsym[0])].copy()


def qlfc_inverse(table, ranks, lengths, n: int) -> np.ndarray:
    t = as_u8(table)
    r = as_u8(ranks)
    ln = np.ascontiguousarray(lengths, dtype=np.int32)
    out = np.empty(max(n, 1), dtype=np.uint8)
    rc = lib().orc_qlfc_inverse(_u8(t), t.size, _u8(r), _i32(ln), r.size, _u8(out), n)
    assert rc == 0, rc
    return out[:n]


def segments(bwt):
    """libbsc coder segments of one block: list of (start, size)."""
    a = as_u8(bwt)
    nb = int(lib().orc_num_segments(a.size))
    st = np.zeros(8, dtype=np.int32)
    sz = np.zeros(8, dtype=np.int32)
    lib().orc_split_segments(_u8(a), a.size, nb, _i32(st), _i32(sz))
    return list(zip(st[:nb].tolist(), sz[:nb].tolist()))
