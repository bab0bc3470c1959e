"""GPU forward/inverse BWT parity against the oracle (bit-exact)."""
import random

import numpy as np
import pytest

from oracle import oracle as O
from qz_amd import synth

pytestmark = pytest.mark.gpu


def _cases():
    rng = random.Random(42)
    cases = [b"ab" * 20, b"A" * 1000, b"\x00" * 64, bytes(range(256)) * 3, b"banana" * 11, b"mississippi" * 9]
    for _ in range(20):
        n = rng.randint(29, 5000)
        sigma = rng.choice([1, 2, 4, 5, 40, 256])
        cases.append(bytes(rng.randrange(sigma) for _ in range(n)))
    for p in (1, 2, 3, 5, 8):
        base = bytes(rng.randrange(4) + 65 for _ in range(p)) * (6000 // p)
        b = bytearray(base)
        b[rng.randrange(len(b))] = 90
        cases.append(bytes(b))
    return cases


def _check(blocks, gpu):
    out, prim, aux = gpu.bwt.encode_batch(blocks)
    for b, blk in enumerate(blocks):
        u, p, a = O.bwt_aux(blk)
        assert np.array_equal(out[b], u), f"block {b} (n={len(blk)}) BWT differs"
        assert prim[b] == p, f"block {b} primary {prim[b]} != {p}"
        assert np.array_equal(aux[b], a), f"block {b} aux differs"
    back = gpu.bwt.decode_batch(out, prim)
    for b, blk in enumerate(blocks):
        assert np.array_equal(back[b], np.frombuffer(bytes(blk), np.uint8)), f"block {b} round trip"


def test_bwt_small_cases_one_by_one(gpu):
    for blk in _cases():
        print(f"case n={len(blk)}", flush=True)
        _check([blk], gpu)


def test_bwt_small_cases_batched(gpu):
    _check(_cases(), gpu)


def test_bwt_libbsc_wrappers(gpu):
    blk = synth.sequence_stream(20_000, 2_000, 150, 3, 0.001).tobytes()
    u, p, idx = gpu.bwt.bwt_encode(blk)
    ou, op, oidx = O.bsc_bwt_encode(blk)
    assert np.array_equal(u, ou) and p == op and np.array_equal(idx, oidx)
    assert gpu.bwt.bwt_decode(u, p).tobytes() == blk


def test_bwt_mixed_streams(gpu):
    rng = np.random.default_rng(7)
    blocks = [
        synth.sequence_stream(200_000, 4_000, 150, 5, 0.002, 0.0001),
        synth.header_stream(5_000, 6),
        synth.quality_stream(3_000, 150, 7),
        rng.integers(0, 256, size=100_000, dtype=np.uint8),
    ]
    _check(blocks, gpu)


def test_bwt_reads_15mb_block(gpu):
    """One 15,000,000-byte sequence block (1 Mbp genome, 100K x 150 bp)."""
    blk = synth.reads_15mb_block(seed=1)
    assert blk.size == 15_000_000
    _check([blk], gpu)


def test_digit_widths_give_identical_order(gpu, monkeypatch):
    """8-, 9- and 11-bit radix digits (the tuner's choices) sort alike, on
    round 0 and on groups larger than 4096 (headers give large groups)."""
    blocks = [synth.header_stream(60_000, 31), synth.quality_stream(20_000, 150, 32)]
    res = []
    for bits in ("8", "9", "11"):
        monkeypatch.setenv("QZBSC_RADIX_BITS", bits)
        out, prim, aux = gpu.bwt.encode_batch(blocks)
        res.append((out, prim))
    for b, blk in enumerate(blocks):
        u, p, _ = O.bwt_aux(blk)
        for out, prim in res:
            assert np.array_equal(out[b], u) and prim[b] == p


def test_pathological_repeats(gpu):
    """Deep suffix sorting: one-symbol and short-period blocks need ~log2(n)
    doubling rounds with every suffix in one large group."""
    rng = np.random.default_rng(41)
    blocks = [np.full(2_000_000, 71, np.uint8),
              np.tile(np.frombuffer(b"ACGTTGCA", np.uint8), 250_000),
              np.tile(rng.integers(0, 4, 1000).astype(np.uint8), 1_500),
              np.concatenate([np.full(1_000_000, 0, np.uint8), np.full(1, 1, np.uint8),
                              np.full(1_000_000, 0, np.uint8)])]
    for blk in blocks:
        print(f"n={blk.size}", flush=True)
        _check([blk], gpu)
    _check(blocks, gpu)


@pytest.mark.parametrize("mode", ["1", "0"])
def test_round0_key_modes_give_the_same_bwt(gpu, monkeypatch, mode):
    """Run-length round-0 keys (per-group depths) and packed keys give the
    oracle's BWT on every kind of block: forced on for reads, names and
    random bytes, forced off for run-heavy qualities, alone and batched."""
    monkeypatch.setenv("QZBSC_RLE", mode)
    rng = np.random.default_rng(77)
    blocks = [synth.quality_stream(20_000, 150, 33), synth.header_stream(20_000, 34),
              synth.sequence_stream(200_000, 4_000, 150, 35, 0.002, 0.0001),
              rng.integers(0, 256, size=50_000, dtype=np.uint8),
              np.full(300_000, 70, np.uint8),
              np.repeat(rng.integers(0, 4, 20_000).astype(np.uint8), rng.integers(1, 90, 20_000))[:800_000]]
    print("key mode", mode, [b.size for b in blocks], flush=True)
    _check(_cases(), gpu)
    _check(blocks, gpu)
    for blk in blocks[:2]:
        _check([blk], gpu)


def _fuse_batches():
    """Batches whose whole alphabet fits the 4-bit text (sigma + 1 <= 16), the
    case where round 0 also does the first doubling round for small groups
    inside one radix tile. The radix (sigma + 1) ** k0 reaches 2 ** 32 for
    sigma 1, 3 and 15, so the largest code is 0xffffffff there."""
    rng = np.random.default_rng(88)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    unit = rng.integers(0, 4, 50_000)
    dna = [synth.sequence_stream(60_000, 9_000, 150, 36, 0.002, 0.0001),
           acgt[rng.integers(0, 4, 5_000)], np.frombuffer(b"ACGT" * 20, np.uint8),
           acgt[np.tile(unit, 40)],                 # 40 copies: fused groups of 40, some across tiles
           acgt[np.tile(unit[:10_000], 64)],        # 64 copies: the largest fused size
           acgt[np.tile(unit[:7_000], 65)],         # 65 copies: just past it
           np.frombuffer(b"A", np.uint8), np.frombuffer(b"CA", np.uint8)]
    acg = np.frombuffer(b"ACG", np.uint8)
    three = [acg[rng.integers(0, 3, 200_000)], acg[np.tile(rng.integers(0, 3, 3_000), 50)], acg[rng.integers(0, 3, 17)]]
    fifteen = np.frombuffer(b"ABCDEFGHIJKLMNO", np.uint8)
    hexa = [fifteen[rng.integers(0, 15, 300_000)], fifteen[np.tile(rng.integers(0, 15, 2_000), 30)]]
    ones = [np.full(100_000, 65, np.uint8), np.full(5_000, 65, np.uint8), np.full(70, 65, np.uint8)]
    return {"dna": dna, "sigma3": three, "sigma15": hexa, "sigma1": ones}


@pytest.mark.parametrize("fuse", ["1", "0"])
@pytest.mark.parametrize("kind", ["dna", "sigma3", "sigma15", "sigma1"])
def test_fused_first_round_matches_oracle(gpu, monkeypatch, fuse, kind):
    """Round 0's finish with the first doubling round fused in (and, for
    comparison, without): the oracle's BWT, primary and samples on every
    block, batched and alone. A one-symbol block would take run-length keys,
    which never fuse, so packed keys are forced for it."""
    monkeypatch.setenv("QZBSC_FUSE", fuse)
    if kind == "sigma1":
        monkeypatch.setenv("QZBSC_RLE", "0")
    blocks = _fuse_batches()[kind]
    print("fuse", fuse, kind, [b.size for b in blocks], flush=True)
    _check(blocks, gpu)
    _check([blocks[0]], gpu)


def _hip_runtime():
    """The HIP runtime that the library has already loaded in this process."""
    import ctypes as C
    with open("/proc/self/maps") as f:
        path = next(line.split()[-1] for line in f if "libamdhip64" in line)
    return C.CDLL(path)


def test_device_entry_at_unaligned_pointers(gpu):
    """The device-resident batch entry with input and output at odd byte
    offsets (the round-0 and packing kernels take 16- and 8-byte loads only
    when the text is aligned): forward against the oracle, inverse back to
    the input. One batch of at most 15 symbols (4-bit text for the output
    gathers) and one with a full byte alphabet."""
    import ctypes as C
    from qz_amd import bwt
    hip = _hip_runtime()
    rng = np.random.default_rng(91)
    low = [synth.sequence_stream(50_000, 300, 150, 5, 0.002, 0.001)[:40_003], synth.quality_stream(300, 150, 6)[:30_001]]
    high = [rng.integers(0, 256, 25_005, dtype=np.uint8), synth.header_stream(500, 7)[:20_007]]
    for batch in (low, high):
        sizes = np.array([b.size for b in batch], np.int32)
        flat = np.ascontiguousarray(np.concatenate(batch), dtype=np.uint8)
        n = flat.size
        for off_in, off_out in ((1, 3), (8, 5), (3, 0)):
            print("sigma", np.unique(flat).size, "offsets", off_in, off_out, "sizes", sizes.tolist(), flush=True)
            bufs = [C.c_void_p(), C.c_void_p(), C.c_void_p()]
            for d in bufs:
                assert hip.hipMalloc(C.byref(d), C.c_size_t(n + 64)) == 0
            try:
                d_in, d_out, d_back = (bufs[0].value + off_in, bufs[1].value + off_out, bufs[2].value + off_in)
                assert hip.hipMemcpy(C.c_void_p(d_in), flat.ctypes.data_as(C.c_void_p), C.c_size_t(n), 1) == 0
                prim, aux, _ = bwt.dev_encode_batch(d_in, d_out, sizes)
                out = np.empty(n, np.uint8)
                assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(d_out), C.c_size_t(n), 2) == 0
                pos = 0
                for b, blk in enumerate(batch):
                    u, p, a = O.bwt_aux(blk)
                    assert np.array_equal(out[pos:pos + blk.size], u), f"block {b} BWT differs"
                    assert prim[b] == p, f"block {b} primary {prim[b]} != {p}"
                    assert np.array_equal(aux[b][:len(a)], a), f"block {b} aux differs"
                    pos += blk.size
                bwt.dev_decode_batch(d_out, d_back, sizes, prim)
                back = np.empty(n, np.uint8)
                assert hip.hipMemcpy(back.ctypes.data_as(C.c_void_p), C.c_void_p(d_back), C.c_size_t(n), 2) == 0
                assert np.array_equal(back, flat), "inverse round trip differs"
            finally:
                for d in bufs:
                    hip.hipFree(d)


def test_forward_workspace_per_input_byte(gpu):
    """The forward workspace measured on the device (free memory before and
    after the first call of a trimmed pool): at most 32 bytes per input byte
    for reads (packed keys), 40 with run-length keys (qualities: the depth
    arrays are allocated for such batches only); never above the footprint
    the admission model charges, up to 16 MiB of allocation rounding over
    the workspace's two dozen buffers. The result is still the oracle's."""
    from qz_amd import serve
    lib = gpu.lib()
    n = 64 << 20
    cases = [("reads", synth.reads_50mbp_stream(500_000, seed=9)[:n], 32.0),
             ("qualities", synth.quality_stream(450_000, 150, 9)[:n], 40.0)]
    for name, blk, limit in cases:
        lib.qzbsc_trim()
        free0, _ = serve.device_memory()
        out, prim, _ = gpu.bwt.encode_batch([blk])
        free1, _ = serve.device_memory()
        per = (free0 - free1) / blk.size
        model = serve.footprint(0, blk.size, 1) / blk.size
        print(f"{name}: {per:.2f} B per input byte resident, model {model:.2f}", flush=True)
        assert per <= limit and free0 - free1 <= serve.footprint(0, blk.size, 1) + (16 << 20)
        if name == "reads":
            u, p, _ = O.bwt_aux(blk[:4 << 20])
            o2, p2, _ = gpu.bwt.encode_batch([blk[:4 << 20]])
            assert np.array_equal(o2[0], u) and p2[0] == p


def test_tuner_knobs_settle_and_never_change_results(gpu, monkeypatch):
    """The launch-shape tuner: after a few calls of >= 4M suffixes the
    round-0 digit width and the fused finish's tile shape are chosen among
    their candidates, every knob reports only its candidates, and each
    forced candidate (digit width, tile shape, small-sort LDS budget) gives
    the bytes the oracle gives."""
    from qz_amd import bwt
    blocks = [synth.sequence_stream(3_000_000, 40_000, 150, 31 + k, 0.002, 0.0001)[:5_000_000] for k in range(2)]
    for _ in range(4):
        out, prim, aux = bwt.encode_batch(blocks)
    t = bwt.tuning()
    print(t, flush=True)
    assert t["round0_digit_bits"]["choice"] in (8, 11)
    assert t["fused_finish_shape"]["choice"] in ("512 threads, suffixes in LDS", "1024 threads, suffixes in LDS")
    assert set(t["sort_small_lds_bytes"]["ns_per_unit"]) <= {"16384", "0"}
    want = [O.bwt_aux(b) for b in blocks]
    for i, (u, p, a) in enumerate(want):
        assert np.array_equal(out[i], u) and prim[i] == p and np.array_equal(aux[i], a)
    for env, val in (("QZBSC_FUSE_SHAPE", "1024s"), ("QZBSC_FUSE_SHAPE", "512s"), ("QZBSC_LDS_SORT_SMALL", "0"),
                     ("QZBSC_LDS_SORT_SMALL", "16384"), ("QZBSC_RADIX_BITS", "11")):
        monkeypatch.setenv(env, val)
        o2, p2, a2 = bwt.encode_batch(blocks)
        monkeypatch.delenv(env)
        for i, (u, p, a) in enumerate(want):
            assert np.array_equal(o2[i], u) and p2[i] == p and np.array_equal(a2[i], a), f"{env}={val} block {i}"


def test_chunked_workspace_matches_oracle(gpu, monkeypatch):
    """The sort workspace holds one chunk of whole blocks; a larger batch is
    sorted chunk by chunk. With a 4 MiB chunk forced on a fresh engine: a
    24 MB batch of reads (contiguous chunks) and a mixed batch (alphabet
    classes, each cut into chunks) equal the oracle block by block, the
    device holds about a third of the unchunked workspace for the reads,
    and the inverse gives the blocks back."""
    from qz_amd import bwt, serve
    lib = gpu.lib()
    reads = synth.reads_50mbp_stream(200_000, seed=41)
    seq = [np.ascontiguousarray(reads[i * 2_000_000:(i + 1) * 2_000_000 - 7 * i]) for i in range(12)]
    mixed, _ = synth.mixed_blocks(24, seed=42, lo=200_000, hi=1_800_000)
    monkeypatch.setenv("QZBSC_CHUNK_MIB", "4")
    try:
        for name, blocks in (("reads", seq), ("mixed", mixed)):
            n = sum(b.size for b in blocks)
            lib.qzbsc_trim()
            free0, _ = serve.device_memory()
            monkeypatch.setenv("QZBSC_TRACE", "1")
            out, prim, aux = bwt.encode_batch(blocks)
            monkeypatch.delenv("QZBSC_TRACE")
            free1, _ = serve.device_memory()
            per = (free0 - free1) / n
            print(f"{name}: {len(blocks)} blocks, {n} B, {per:.2f} B per input byte resident", flush=True)
            if name == "reads":
                assert per < 16.0, per  # 31 without chunks
            for i, blk in enumerate(blocks):
                u, p, a = O.bwt_aux(blk)
                assert np.array_equal(out[i], u) and prim[i] == p and np.array_equal(aux[i], a), f"{name} block {i}"
            back = bwt.decode_batch(out, prim)
            assert all(np.array_equal(back[i], b) for i, b in enumerate(blocks))
    finally:
        monkeypatch.delenv("QZBSC_CHUNK_MIB")
        lib.qzbsc_trim()


def test_six_bit_char_mode_names(gpu, capfd, monkeypatch):
    """Names (about 40 symbols) carry their preceding symbol in 6-bit codes
    in the SA entries (block-local positions below 2^26), so their BWT
    output is a streaming pass: the sub-batches report char mode 6 and every
    block equals the oracle."""
    from qz_amd import bwt
    names = synth.header_stream(160_000, 43)
    blocks = [np.ascontiguousarray(names[i * 2_500_001:(i + 1) * 2_500_001]) for i in range(4)]
    monkeypatch.setenv("QZBSC_CHUNK_MIB", "6")
    monkeypatch.setenv("QZBSC_TRACE", "1")
    gpu.lib().qzbsc_trim()
    try:
        out, prim, aux = bwt.encode_batch(blocks)
    finally:
        monkeypatch.delenv("QZBSC_TRACE")
        monkeypatch.delenv("QZBSC_CHUNK_MIB")
        gpu.lib().qzbsc_trim()
    err = capfd.readouterr().err
    assert "char mode 6" in err, err[-1500:]
    for i, blk in enumerate(blocks):
        u, p, a = O.bwt_aux(blk)
        assert np.array_equal(out[i], u) and prim[i] == p and np.array_equal(aux[i], a), f"block {i}"


def test_char_mode_limits(gpu):
    """Edges of the char-mode SA entries: 63 distinct symbols take 6-bit
    codes, 64 fall back to global positions; a name block of 2^26 - 1 bytes
    (the largest whose block-local positions leave 6 bits free) and one of
    2^26 bytes (which cannot) both equal the oracle."""
    from qz_amd import bwt
    rng = np.random.default_rng(44)
    for sigma in (63, 64):
        alpha = rng.choice(256, size=sigma, replace=False).astype(np.uint8)
        blk = alpha[np.minimum(rng.geometric(0.05, size=1_000_003) - 1, sigma - 1)]
        assert np.unique(blk).size == sigma
        out, prim, aux = bwt.encode_batch([blk, blk[:300_001].copy()])
        for i, b in enumerate((blk, blk[:300_001])):
            u, p, a = O.bwt_aux(b)
            assert np.array_equal(out[i], u) and prim[i] == p and np.array_equal(aux[i], a), f"sigma {sigma} block {i}"
    names = synth.header_stream(1_100_000, 45)
    for n in ((1 << 26) - 1, 1 << 26):
        blk = np.ascontiguousarray(names[:n])
        print("name block", n, flush=True)
        out, prim, aux = bwt.encode_batch([blk])
        u, p, a = O.bwt_aux(blk)
        assert np.array_equal(out[0], u) and prim[0] == p and np.array_equal(aux[0], a), f"name block {n}"
