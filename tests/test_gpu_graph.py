"""HIP graph replay of the fixed launch sequences (forward round 0, inverse
BWT): with new data of the same batch shape on every call, and other GPU work
and synchronisation between calls, every replay returns exactly what the
oracle and the eager launch sequence return."""
import numpy as np
import pytest

from oracle import oracle as O
from qz_amd import synth
from qz_amd._native import lib, u8p, i32p

pytestmark = pytest.mark.gpu

SIZES = [300_000, 123_457, 700_001, 50_000]  # 1.17 MB: below the tuner's size, so round 0 is graph-eligible


def _blocks(it):
    seq = synth.sequence_stream(200_000, 3_000, 150, 100 + it, 0.002, 0.0001)
    hdr = synth.header_stream(4_000, 200 + it)
    qual = synth.quality_stream(6_000, 150, 300 + it)
    rng = np.random.default_rng(400 + it)
    src = [seq, hdr, qual, rng.integers(0, 256, SIZES[3], dtype=np.uint8)]
    return [np.ascontiguousarray(s[:n]) for s, n in zip(src, SIZES)]


def _other_gpu_work(bwt):
    """The device QLFC front end (another engine of the same slot and stream),
    synchronised before it returns."""
    b = np.ascontiguousarray(bwt, np.uint8)
    n = b.size
    nseg = np.zeros(1, np.int32)
    st, sz, sr, nsym = (np.zeros(8, np.int32) for _ in range(4))
    tables = np.zeros(8 * 256, np.uint8)
    sym, rank, ln = np.zeros(n, np.uint8), np.zeros(n, np.uint8), np.zeros(n, np.int32)
    R = lib().qzbsc_qlfc_front(b.ctypes.data_as(u8p), n, nseg.ctypes.data_as(i32p), st.ctypes.data_as(i32p),
                               sz.ctypes.data_as(i32p), sr.ctypes.data_as(i32p), tables.ctypes.data_as(u8p),
                               nsym.ctypes.data_as(i32p), sym.ctypes.data_as(u8p), rank.ctypes.data_as(u8p),
                               ln.ctypes.data_as(i32p))
    assert R > 0 and int(ln[:R].sum()) == n


def test_graph_replays_match_oracle_and_eager(gpu, monkeypatch):
    monkeypatch.delenv("QZBSC_GRAPH", raising=False)
    c0, r0 = gpu.bwt.graph_stats()
    for it in range(5):
        blocks = _blocks(it)
        assert [b.size for b in blocks] == SIZES
        print(f"iteration {it}", flush=True)
        out, prim, aux = gpu.bwt.encode_batch(blocks)
        back = gpu.bwt.decode_batch(out, prim)
        _other_gpu_work(out[2])  # unrelated eager GPU work between replays
        for b, blk in enumerate(blocks):
            u, p, a = O.bwt_aux(blk)
            assert np.array_equal(out[b], u), f"iteration {it} block {b} BWT"
            assert prim[b] == p and np.array_equal(aux[b], a), f"iteration {it} block {b} indexes"
            assert np.array_equal(back[b], blk), f"iteration {it} block {b} round trip"
    c1, r1 = gpu.bwt.graph_stats()
    # iteration 0 runs eagerly, 1 captures, 2..4 replay forward round 0 and the inverse
    assert c1 - c0 >= 2 and r1 - r0 >= 6, (c1 - c0, r1 - r0)
    monkeypatch.setenv("QZBSC_GRAPH", "0")
    out_e, prim_e, _ = gpu.bwt.encode_batch(blocks)
    back_e = gpu.bwt.decode_batch(out_e, prim_e)
    for b in range(len(blocks)):
        assert np.array_equal(out_e[b], out[b]) and prim_e[b] == prim[b]
        assert np.array_equal(back_e[b], back[b])
    assert gpu.bwt.graph_stats()[1] == r1, "QZBSC_GRAPH=0 must not replay"
