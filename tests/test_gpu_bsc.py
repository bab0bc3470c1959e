"""libbsc drop-in ABI on the GPU: round trips, block fields against the
oracle, error behaviour on corrupt / truncated blocks (the reference's
integration tests require an Err on both, crates/qz-lib/tests/
integration_test.rs:470-549), determinism."""
import struct
import zlib

import numpy as np
import pytest

from oracle import oracle as O
from qz_amd import bsc, bwt, synth
from qz_amd.bwt import BscError

pytestmark = pytest.mark.gpu

# BWT sorter | this library's coder id << 5 (libbsc's own blocks read 0x41)
MODE_QZ = bsc.LIBBSC_BLOCKSORTER_BWT | (bsc.QZBSC_CODER_QZ_ADAPTIVE << 5)


def _inputs():
    rng = np.random.default_rng(11)
    return [
        b"ACGTACGTACGTACGTACGTACGTACGTACGT",            # bsc.rs:446 test_roundtrip_simple
        b"ACGTACGTNNNACGTACGTACGTACGTNNACGTACGTACGTACGTACGT"
        b"ACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGTACGT",  # bsc.rs:454
        b"I" * 53 + b"HHHHHHHHHHGGGGGGGGFFFFFFFFFFEEEEEEEEDDDDDDDDDCCCCCCC",  # bsc.rs:463
        bytes(29), b"A" * 70_000, rng.integers(0, 256, 70_000).astype(np.uint8).tobytes(),
        synth.sequence_stream(50_000, 3_000, 150, 12, 0.002, 0.0001).tobytes(),
        synth.header_stream(30_000, 13).tobytes(),
        synth.quality_stream(8_000, 150, 14).tobytes(),
    ]


def test_roundtrip_and_fields():
    for data in _inputs():
        print(f"n={len(data)}", flush=True)
        c = bsc.compress_adaptive_no_lzp(data)
        assert bsc.decompress(c) == data
        size, n, mode, prim, a_data, a_body, a_hdr = struct.unpack_from("<iiiiIII", c, 0)
        assert size == len(c) and n == len(data)
        assert a_data == zlib.adler32(data) and a_body == zlib.adler32(c[28:]) and a_hdr == zlib.adler32(c[:24])
        if mode == 0:
            assert c[28:] == data
            continue
        assert mode == MODE_QZ
        _, p, idx = O.bsc_bwt_encode(data)
        assert prim == p
        nidx = c[-1]
        want = idx.tolist() if len(data) >= 64 * 1024 else []
        assert nidx == len(want)
        got = list(struct.unpack_from(f"<{nidx}i", c, len(c) - 1 - 4 * nidx)) if nidx else []
        assert got == want


def test_ratio_and_mt_identity():
    data = synth.reads_15mb_block(3)[:3_000_000].tobytes()
    a = bsc.compress_adaptive_no_lzp(data)
    b = bsc.compress_adaptive_mt_no_lzp(data)
    assert a == b
    assert len(a) < len(data) // 3
    assert bsc.decompress_mt(a) == data


def test_corrupt_and_truncated_blocks_fail():
    data = synth.quality_stream(5_000, 150, 21).tobytes()
    c = bytearray(bsc.compress_adaptive_no_lzp(data))
    for pos in (10, 40, len(c) // 2, len(c) - 2):
        bad = bytearray(c)
        bad[pos] ^= 0xFF
        with pytest.raises(BscError):
            bsc.decompress(bytes(bad))
    with pytest.raises(BscError) as e:
        bsc.decompress(bytes(c[:-10]))
    assert e.value.code == -5
    with pytest.raises(BscError):
        bsc.decompress(bytes(c[:20]))


def test_parallel_framing_roundtrip():
    data = synth.quality_stream(2_000, 150, 31).tobytes()
    c = bsc.compress_parallel_adaptive_no_lzp(data)
    assert struct.unpack_from("<I", c, 0)[0] == 1
    assert bsc.decompress_parallel(c) == data
    assert bsc.compress_parallel_adaptive_no_lzp(b"") == b"" and bsc.decompress_parallel(b"") == b""


def test_lzp_and_coder_requests_roundtrip():
    """qz's other parameter sets (bsc.rs:146-150, 194-196, 212-215): LZP
    16/128 and other LZP settings, and the static-QLFC request, all round
    trip; a block LZP shrinks records both LZP parameters in its mode word,
    and every coded block records the adaptive coder."""
    rep = (synth.header_stream(3_000, 21).tobytes() * 6)[:700_000]  # long repeats: LZP 16/128 finds them
    cases = _inputs() + [rep]
    for data in cases:
        for fn in (bsc.compress, bsc.compress_adaptive, bsc.compress_adaptive_mt):
            blk = fn(data)
            assert bsc.decompress(blk) == bytes(data)
            mode = struct.unpack_from("<i", blk, 8)[0]
            if mode:
                assert mode & 0xff == MODE_QZ
    blk = bsc.compress_adaptive(rep)
    mode = struct.unpack_from("<i", blk, 8)[0]
    assert mode == MODE_QZ | (128 << 8) | (16 << 16)
    assert len(blk) < len(bsc.compress_adaptive_no_lzp(rep))
    for hs, ml in ((10, 4), (20, 32), (28, 255)):
        blk = bsc.compress_with_params(rep, hs, ml, 1, 2)
        assert bsc.decompress(blk) == rep
        assert struct.unpack_from("<i", blk, 8)[0] >> 8 in (0, ml | (hs << 8))
    framed = bsc.compress_parallel_adaptive(rep)
    assert struct.unpack_from("<I", framed, 0)[0] == 1


def test_lzp_block_corruption_fails():
    rep = (synth.header_stream(3_000, 22).tobytes() * 6)[:700_000]
    blk = bytearray(bsc.compress_adaptive(rep))
    assert struct.unpack_from("<i", blk, 8)[0] >> 8
    for pos in (30, len(blk) // 2, len(blk) - 2):
        bad = bytearray(blk)
        bad[pos] ^= 0x10
        with pytest.raises(BscError):
            bsc.decompress(bytes(bad))


def test_large_block_mt_roundtrip():
    """qz's large-block MT mode (compress_adaptive_mt_no_lzp, ultra.rs:877-887
    uses blocks up to 750 MB): one 400 MB block through the drop-in ABI. Its
    BWT, primary index and trailer indexes equal the oracle's (SA-IS on the
    host, run in a thread beside the GPU calls)."""
    import threading
    data = synth.reads_50mbp_stream(2_700_000, seed=31)[:400_000_000]
    assert data.size == 400_000_000
    ref = {}
    th = threading.Thread(target=lambda: ref.update(zip(("u", "p", "idx"), O.bsc_bwt_encode(data))))
    th.start()
    print("large block: oracle started, GPU compress", flush=True)
    blk = bsc.compress_adaptive_mt_no_lzp(data)
    size, n, mode, prim = struct.unpack_from("<iiii", blk, 0)
    assert size == len(blk) and n == data.size and mode == MODE_QZ
    nidx = blk[-1]
    assert nidx == (n - 1) // O.sample_rate(n)
    trailer = list(struct.unpack_from(f"<{nidx}i", blk, len(blk) - 1 - 4 * nidx))
    assert len(blk) < data.size // 3
    back = bsc.decompress(blk)
    assert back == data.tobytes()
    del back
    out, prims, _ = bwt.encode_batch([data])
    th.join()
    print("large block: oracle done", flush=True)
    assert prim == ref["p"] == prims[0]
    assert trailer == ref["idx"].tolist()
    assert np.array_equal(out[0], ref["u"])


def test_ultra_block_750mb_roundtrip():
    """ultra's largest sequence block (750 MB, ultra.rs:44-55, 877-887)
    through the drop-in ABI: header fields, sample count, round trip."""
    data = synth.reads_50mbp_stream(5_000_000, seed=33)[:750_000_000]
    assert data.size == 750_000_000
    blk = bsc.compress_adaptive_mt_no_lzp(data)
    size, n, mode = struct.unpack_from("<iii", blk, 0)
    assert size == len(blk) and n == data.size and mode == MODE_QZ
    assert blk[-1] == (n - 1) // O.sample_rate(n)
    assert bsc.decompress(blk) == data.tobytes()
