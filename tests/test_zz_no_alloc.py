"""The no-allocation inverse BWT (qzbsc_dev_bwt_decode_batch_noalloc)
against the allocating device entry (qzbsc_dev_bwt_decode_batch) on the same
device buffers: two calls in a row, the second batch smaller, both outputs
equal to each other and to the original blocks; batches beyond the limits
are refused with LIBBSC_NOT_ENOUGH_MEMORY."""
import ctypes as C

import numpy as np
import pytest

import qz_amd
from qz_amd import bwt, synth
from qz_amd._native import i32p

pytestmark = pytest.mark.gpu


def _hip_runtime():
    with open("/proc/self/maps") as f:
        path = next(line.split()[-1] for line in f if "libamdhip64" in line)
    return C.CDLL(path)


def _run(hip, L, blocks):
    sizes = np.array([b.size for b in blocks], np.int32)
    out, prim, _ = bwt.encode_batch(blocks)  # BWTs and primaries
    flat = np.ascontiguousarray(np.concatenate(out))
    prim = np.array(prim, np.int32)
    n = flat.size
    print("no-alloc batch:", sizes.tolist(), flush=True)
    bufs = [C.c_void_p(), C.c_void_p(), C.c_void_p()]
    for d in bufs:
        assert hip.hipMalloc(C.byref(d), C.c_size_t(n + 64)) == 0
    try:
        d_in, d_a, d_b = (x.value for x in bufs)
        assert hip.hipMemcpy(C.c_void_p(d_in), flat.ctypes.data_as(C.c_void_p), C.c_size_t(n), 1) == 0
        bwt.dev_decode_batch(d_in, d_a, sizes, prim)
        rc = L.qzbsc_dev_bwt_decode_batch_noalloc(C.c_void_p(d_in), C.c_void_p(d_b), sizes.ctypes.data_as(i32p),
                                                  sizes.size, prim.ctypes.data_as(i32p), None)
        assert rc == 0, rc
        a, b = np.empty(n, np.uint8), np.empty(n, np.uint8)
        assert hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(d_a), C.c_size_t(n), 2) == 0
        assert hip.hipMemcpy(b.ctypes.data_as(C.c_void_p), C.c_void_p(d_b), C.c_size_t(n), 2) == 0
        assert np.array_equal(a, b), "no-alloc output differs from the allocating entry's"
        assert np.array_equal(b, np.concatenate(blocks)), "no-alloc output is not the original"
    finally:
        for d in bufs:
            hip.hipFree(d)


def test_noalloc_inverse_matches_allocating_entry(gpu):
    L = qz_amd.lib()
    hip = _hip_runtime()
    first = [synth.sequence_stream(200_000, 20_000, 150, 3, 0.002, 0.0001)[:3_000_000],
             synth.header_stream(60_000, 4)[:2_500_000], synth.quality_stream(15_000, 150, 5)[:2_000_017]]
    assert sum(b.size for b in first) <= 8 << 20
    _run(hip, L, first)
    second = [first[1][:400_001], first[0][:123_457]]  # smaller, after the larger one
    _run(hip, L, second)


def test_noalloc_limits(gpu):
    L = qz_amd.lib()
    one = np.zeros(1, np.int32)
    big = np.array([(8 << 20) + 1], np.int32)
    many = np.full(65, 100, np.int32)
    pr = np.ones(65, np.int32)
    # the limits are checked before any device pointer is touched
    dummy = C.c_void_p(16)
    assert L.qzbsc_dev_bwt_decode_batch_noalloc(dummy, dummy, big.ctypes.data_as(i32p), 1,
                                                pr.ctypes.data_as(i32p), None) == -2
    assert L.qzbsc_dev_bwt_decode_batch_noalloc(dummy, dummy, many.ctypes.data_as(i32p), 65,
                                                pr.ctypes.data_as(i32p), None) == -2
    assert L.qzbsc_dev_bwt_decode_batch_noalloc(dummy, dummy, one.ctypes.data_as(i32p), 1,
                                                pr.ctypes.data_as(i32p), None) == -1  # empty block
