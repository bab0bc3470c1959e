"""GPU QLFC front end (segment split + run/rank transform) against the
oracle, bit-exact."""
import numpy as np
import pytest

from oracle import oracle as O
from qz_amd import synth
from qz_amd._native import lib, u8p, i32p

pytestmark = pytest.mark.gpu


def gpu_front(bwt):
    b = np.ascontiguousarray(bwt, np.uint8)
    n = b.size
    nseg = np.zeros(1, np.int32)
    st, sz, sr = (np.zeros(8, np.int32) for _ in range(3))
    tables = np.zeros(8 * 256, np.uint8)
    nsym = np.zeros(8, np.int32)
    sym, rank = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    ln = np.zeros(n, np.int32)
    R = lib().qzbsc_qlfc_front(b.ctypes.data_as(u8p), n, nseg.ctypes.data_as(i32p), st.ctypes.data_as(i32p),
                               sz.ctypes.data_as(i32p), sr.ctypes.data_as(i32p), tables.ctypes.data_as(u8p),
                               nsym.ctypes.data_as(i32p), sym.ctypes.data_as(u8p), rank.ctypes.data_as(u8p),
                               ln.ctypes.data_as(i32p))
    assert R >= 0, R
    k = int(nseg[0])
    return k, st[:k], sz[:k], sr[:k], tables.reshape(8, 256), nsym[:k], sym[:R], rank[:R], ln[:R]


def _blocks():
    rng = np.random.default_rng(17)
    out = [synth.reads_15mb_block(5)[:3_000_000], synth.header_stream(60_000, 6), synth.quality_stream(20_000, 150, 7),
           rng.integers(0, 256, 700_000).astype(np.uint8), np.full(600_000, 7, np.uint8),
           np.concatenate([np.full(900_000, 65, np.uint8), np.full(10, 66, np.uint8)]),
           rng.integers(0, 4, 40_000).astype(np.uint8), np.frombuffer(b"AB" * 50, np.uint8)]
    return out


def test_front_end_matches_oracle():
    for data in _blocks():
        u, _, _ = O.bwt_aux(data)
        print(f"n={u.size}", flush=True)
        k, st, sz, sr, tables, nsym, sym, rank, ln = gpu_front(u)
        segs = O.segments(u)
        assert [(int(a), int(b)) for a, b in zip(st, sz)] == segs
        r = 0
        for s, (a, b) in enumerate(segs):
            osym, orank, olen, otab = O.qlfc_transform(u[a:a + b])
            nr = int(sr[s])
            assert nr == osym.size
            assert np.array_equal(sym[r:r + nr], osym)
            assert np.array_equal(rank[r:r + nr], orank)
            assert np.array_equal(ln[r:r + nr], olen)
            assert int(nsym[s]) == otab.size and np.array_equal(tables[s, :otab.size], otab)
            r += nr
        assert r == sym.size
