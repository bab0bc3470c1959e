"""Batched uneven blocks (sequence / header / quality mix) in one device
pass: every block bit-exact against the oracle, and identical to the same
block transformed alone."""
import numpy as np
import pytest

from oracle import oracle as O
from qz_amd import bwt, synth

pytestmark = pytest.mark.gpu


def test_uneven_mixed_batch_matches_oracle_and_alone():
    blocks, kinds = synth.mixed_blocks(24, seed=9, lo=2_000, hi=1_500_000)
    blocks = blocks + [np.full(30_000, 65, np.uint8), np.frombuffer(b"ACGT" * 20, np.uint8)]
    print("sizes", [b.size for b in blocks], flush=True)
    out, prim, aux = bwt.encode_batch(blocks)
    for i, blk in enumerate(blocks):
        u, p, a = O.bwt_aux(blk)
        assert np.array_equal(out[i], u) and prim[i] == p and np.array_equal(aux[i], a), f"block {i}"
    # same result alone as in the mix
    for i in (0, 5, len(blocks) - 1):
        o1, p1, a1 = bwt.encode_batch([blocks[i]])
        assert np.array_equal(o1[0], out[i]) and p1[0] == prim[i] and np.array_equal(a1[0], aux[i])
    back = bwt.decode_batch(out, prim)
    for i, blk in enumerate(blocks):
        assert np.array_equal(back[i], blk), f"round trip {i}"


def test_alphabet_class_split_matches_oracle(gpu, capfd, monkeypatch):
    """A batch that mixes reads, names and binned qualities (each kind
    above the 4 MiB class minimum, interleaved, plus tiny blocks) is sorted
    as one sub-batch per alphabet class: every block against the oracle,
    through the host-buffer entry and the device-resident entry at odd
    offsets, and the same bytes with the split turned off."""
    import ctypes as C
    from test_gpu_bwt import _hip_runtime
    seq = synth.sequence_stream(2_000_000, 60_000, 150, 21, 0.002, 0.0001)
    names = synth.header_stream(120_000, 22)
    quals = synth.quality_stream(60_000, 150, 23)
    rng = np.random.default_rng(24)
    blocks = []
    for i in range(3):
        for src in (seq, names, quals):
            n = 2_000_000 + 7_919 * i
            o = int(rng.integers(0, src.size - n))
            blocks.append(np.ascontiguousarray(src[o:o + n]))
    blocks.insert(4, rng.integers(0, 256, 70_001, dtype=np.uint8))
    blocks.insert(1, np.frombuffer(b"GATTACA" * 71, np.uint8).copy())
    sizes = np.array([b.size for b in blocks], np.int32)
    print("sizes", sizes.tolist(), flush=True)
    monkeypatch.setenv("QZBSC_TRACE", "1")
    out, prim, aux = bwt.encode_batch(blocks)
    err = capfd.readouterr().err
    monkeypatch.delenv("QZBSC_TRACE")
    assert err.count("qzbsc: sub-batch of") >= 3, err[-2000:]
    for i, blk in enumerate(blocks):
        u, p, a = O.bwt_aux(blk)
        assert np.array_equal(out[i], u) and prim[i] == p and np.array_equal(aux[i], a), f"block {i}"
    # device-resident entry, input and output at odd offsets
    hip = _hip_runtime()
    flat = np.ascontiguousarray(np.concatenate(blocks))
    n = flat.size
    bufs = [C.c_void_p(), C.c_void_p()]
    for d in bufs:
        assert hip.hipMalloc(C.byref(d), C.c_size_t(n + 64)) == 0
    try:
        d_in, d_out = bufs[0].value + 5, bufs[1].value + 2
        assert hip.hipMemcpy(C.c_void_p(d_in), flat.ctypes.data_as(C.c_void_p), C.c_size_t(n), 1) == 0
        p2, a2, _ = bwt.dev_encode_batch(d_in, d_out, sizes)
        got = np.empty(n, np.uint8)
        assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), C.c_void_p(d_out), C.c_size_t(n), 2) == 0
    finally:
        for d in bufs:
            hip.hipFree(d)
    assert np.array_equal(got, np.concatenate(out)) and p2.tolist() == list(prim)
    assert all(np.array_equal(a2[i][:len(aux[i])], aux[i]) for i in range(len(blocks)))
    # one batch, no split: the same bytes
    monkeypatch.setenv("QZBSC_CLASSES", "0")
    o3, p3, a3 = bwt.encode_batch(blocks)
    monkeypatch.delenv("QZBSC_CLASSES")
    assert all(np.array_equal(o3[i], out[i]) for i in range(len(blocks))) and list(p3) == list(prim)
    back = bwt.decode_batch(out, prim)
    assert all(np.array_equal(back[i], blk) for i, blk in enumerate(blocks))
