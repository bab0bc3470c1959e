"""Python face of the GPU BWT stages (libbsc bwt.h equivalents and batches).

Mirrors libbsc's bsc_bwt_encode / bsc_bwt_decode semantics as qz reaches
them inside bsc_compress / bsc_decompress
(crates/qz-lib/src/compression/bsc.rs:172-183, 421-429).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._native import lib, u8p, i32p, f32p


class BscError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what} failed with error code: {code}")
        self.code = code


def _u8(a):
    return a.ctypes.data_as(u8p)


def _i32(a):
    return a.ctypes.data_as(i32p)


def sample_rate(n: int) -> int:
    m = n // 8
    for s in (1, 2, 4, 8, 16):
        m |= m >> s
    return (m >> 1) + 1


def encode_batch(blocks):
    """Forward BWT of several blocks in one device pass.

    Returns (bwt bytes per block, primary per block, aux index arrays where
    aux[b][k] = 1 + rank of suffix k*r)."""
    arrs = [np.ascontiguousarray(np.frombuffer(b, dtype=np.uint8) if isinstance(b, (bytes, bytearray)) else b,
                                 dtype=np.uint8) for b in blocks]
    sizes = np.array([a.size for a in arrs], dtype=np.int32)
    buf = np.concatenate(arrs) if arrs else np.zeros(0, np.uint8)
    out = np.empty_like(buf)
    prim = np.zeros(len(arrs), dtype=np.int32)
    aux = np.zeros((len(arrs), 256), dtype=np.int32)
    rc = lib().qzbsc_bwt_encode_batch(_u8(buf), _u8(out), _i32(sizes), len(arrs), _i32(prim), _i32(aux))
    if rc != 0:
        raise BscError(rc, "qzbsc_bwt_encode_batch")
    res, o = [], 0
    for b, a in enumerate(arrs):
        res.append(out[o:o + a.size])
        o += a.size
    auxl = [aux[b, : (a.size - 1) // sample_rate(a.size) + 1].copy() for b, a in enumerate(arrs)]
    return res, prim.tolist(), auxl


def decode_batch(bwts, primaries):
    arrs = [np.ascontiguousarray(np.frombuffer(b, dtype=np.uint8) if isinstance(b, (bytes, bytearray)) else b,
                                 dtype=np.uint8) for b in bwts]
    sizes = np.array([a.size for a in arrs], dtype=np.int32)
    buf = np.concatenate(arrs)
    out = np.empty_like(buf)
    prim = np.ascontiguousarray(primaries, dtype=np.int32)
    rc = lib().qzbsc_bwt_decode_batch(_u8(buf), _u8(out), _i32(sizes), len(arrs), _i32(prim))
    if rc != 0:
        raise BscError(rc, "qzbsc_bwt_decode_batch")
    res, o = [], 0
    for a in arrs:
        res.append(out[o:o + a.size])
        o += a.size
    return res


def bwt_encode(block):
    """libbsc bsc_bwt_encode: (bwt, primary, trailer indexes I[k+1]-1)."""
    a = np.array(np.frombuffer(block, dtype=np.uint8) if isinstance(block, (bytes, bytearray)) else block,
                 dtype=np.uint8, copy=True)
    cnt = np.zeros(1, dtype=np.uint8)
    idx = np.zeros(256, dtype=np.int32)
    p = lib().qzbsc_bwt_encode(_u8(a), a.size, _u8(cnt), _i32(idx), 0)
    if p < 0:
        raise BscError(p, "qzbsc_bwt_encode")
    return a, p, idx[: int(cnt[0])].copy()


def bwt_decode(bwt, primary: int):
    a = np.array(np.frombuffer(bwt, dtype=np.uint8) if isinstance(bwt, (bytes, bytearray)) else bwt,
                 dtype=np.uint8, copy=True)
    idx = np.zeros(256, dtype=np.int32)
    rc = lib().qzbsc_bwt_decode(_u8(a), a.size, primary, 0, _i32(idx), 0)
    if rc != 0:
        raise BscError(rc, "qzbsc_bwt_decode")
    return a


def dev_encode_batch(d_in_ptr: int, d_out_ptr: int, sizes):
    """Device-resident forward BWT; returns (primaries, aux, gpu_ms)."""
    sz = np.ascontiguousarray(sizes, dtype=np.int32)
    prim = np.zeros(sz.size, dtype=np.int32)
    aux = np.zeros((sz.size, 256), dtype=np.int32)
    ms = C.c_float(0.0)
    rc = lib().qzbsc_dev_bwt_encode_batch(C.c_void_p(d_in_ptr), C.c_void_p(d_out_ptr), _i32(sz), sz.size,
                                          _i32(prim), _i32(aux), C.byref(ms))
    if rc != 0:
        raise BscError(rc, "qzbsc_dev_bwt_encode_batch")
    return prim, aux, ms.value


def dev_decode_batch(d_in_ptr: int, d_out_ptr: int, sizes, primaries):
    sz = np.ascontiguousarray(sizes, dtype=np.int32)
    prim = np.ascontiguousarray(primaries, dtype=np.int32)
    ms = C.c_float(0.0)
    rc = lib().qzbsc_dev_bwt_decode_batch(C.c_void_p(d_in_ptr), C.c_void_p(d_out_ptr), _i32(sz), sz.size,
                                          _i32(prim), C.byref(ms))
    if rc != 0:
        raise BscError(rc, "qzbsc_dev_bwt_decode_batch")
    return ms.value


_KNOBS = ((0, "round0_digit_bits", "element"), (1, "large_group_digit_bits", "element"),
          (2, "fused_finish_shape", "element"), (3, "sort_small_lds_bytes", "small-group slot"))
_SHAPES = {0: "512 threads, suffixes in LDS", 5: "1024 threads, suffixes in LDS"}


def tuning():
    """Launch-shape tuner state: per knob the chosen value (None while
    undecided), the timed trials and the mean ns per unit of each candidate."""
    res = {}
    for knob, name, unit in _KNOBS:
        ch, nc = C.c_int(0), C.c_int(0)
        cand, ns = (C.c_int * 4)(), (C.c_double * 4)()
        trials = lib().qzbsc_tune_knob(knob, C.byref(ch), C.byref(nc), cand, ns)
        label = (lambda v: _SHAPES.get(v, str(v))) if knob == 2 else str
        res[name] = {"choice": None if ch.value < 0 else (_SHAPES.get(ch.value, ch.value) if knob == 2 else ch.value),
                     "trials": trials, "unit": unit,
                     "ns_per_unit": {label(cand[k]): round(ns[k], 5) for k in range(nc.value) if ns[k] > 0}}
    return res


def graph_stats():
    """(captures, replays) of HIP graphs in this process."""
    c, r = C.c_ulonglong(0), C.c_ulonglong(0)
    lib().qzbsc_graph_stats(C.byref(c), C.byref(r))
    return c.value, r.value
