"""Batch block forms and the .qz v2 archive on the GPU.

* qzbsc_compress_batch gives, block for block, the bytes bsc_compress gives
  (default and LZP parameter sets, stored and tiny blocks included), and
  qzbsc_decompress_batch returns what per-block bsc_decompress returns;
* FASTQ -> archive -> FASTQ is the identity for constant and variable read
  lengths, with streams of many blocks, decoded in one batch call or in
  calls of 8 blocks per stream;
* every block of an archive stream decodes alone to its part of the stream;
* a corrupted block, a truncated archive and the libbsc archive under
  tests/golden (whose QLFC payloads this library's coder does not read)
  raise instead of returning data;
* batches of 60 MB, coded and decoded over several device sub-batches, give
  the per-block bytes;
* segments stored raw (bytes rebuilt from the device runs) in a block of
  reads and random bytes;
* quality_ctx archives (qz's choice from 100,000 reads on) round-trip, in
  one or several blobs, next to the BSC-quality archive of the same text.

No torch here: the library must initialise the device itself.
"""
import struct

import numpy as np
import pytest

from qz_amd import archive as A
from qz_amd import bsc, synth
from qz_amd.bwt import BscError

pytestmark = pytest.mark.gpu


def _blocks():
    rng = np.random.default_rng(4)
    return [
        synth.sequence_stream(200_000, 2_000, 150, 7, 0.002),            # 300 KB of reads
        synth.quality_stream(2_000, 150, 8),                               # qualities
        synth.header_stream(5_000, 9),                                     # names
        rng.integers(0, 256, 70_000, dtype=np.uint8),                      # incompressible: stored
        np.frombuffer(b"ACGTACGTAC", np.uint8),                            # <= 28 bytes: stored
        np.frombuffer(b"G", np.uint8),
        np.frombuffer(b"ACGT" * 40_000, np.uint8),                         # long repeats
    ]


def test_batch_forms_match_per_block_calls():
    blocks = _blocks()
    print("batch forms:", [b.size for b in blocks], flush=True)
    one = [bsc.compress_adaptive_no_lzp(b) for b in blocks]
    many = bsc.compress_batch(blocks)
    assert many == one
    lz_one = [bsc.compress_adaptive(b) for b in blocks]
    lz_many = bsc.compress_batch(blocks, 16, 128)
    assert lz_many == lz_one
    for comp in (many, lz_many):
        back = bsc.decompress_batch(comp)
        assert back == [b.tobytes() for b in blocks]
        assert back == [bsc.decompress(c) for c in comp]


@pytest.mark.parametrize("variable", [False, True])
def test_archive_roundtrip(variable):
    text = synth.fastq(30_000, 150, 21, variable=variable)
    print("archive round trip: variable", variable, len(text), "bytes", flush=True)
    arc = A.compress_records(A.parse_fastq(text), block_size=256 * 1024)
    h = A.parse_archive_header(arc)
    assert h.num_reads == 30_000
    assert (h.const_seq_len, h.const_qual_len) == ((0, 0) if variable else (150, 150))
    spans = A.stream_offsets(h)
    nblocks = [len(A.stream_blocks(arc, o, n)) for o, n in spans]
    assert max(nblocks) > A.DECOMPRESS_BATCH_SIZE  # several batch calls per stream with batch=8
    assert A.decompress_archive(arc) == text
    assert A.decompress_archive(arc, batch=A.DECOMPRESS_BATCH_SIZE) == text
    # each block alone gives its 256 KiB piece of the stream
    cs, cq = A.detect_const_lengths(A.parse_fastq(text))
    streams = A.records_to_streams(A.parse_fastq(text), cs, cq)
    for (o, n), raw in zip(spans, streams):
        blocks = A.stream_blocks(arc, o, n)
        pieces = [bsc.decompress(bytes(b)) for b in blocks]
        assert b"".join(pieces) == raw
        assert all(len(p) == 256 * 1024 for p in pieces[:-1])


def test_archive_corruption_and_truncation_fail():
    text = synth.fastq(5_000, 150, 22)
    arc = bytearray(A.compress_fastq(text))
    h = A.parse_archive_header(bytes(arc))
    o, n = A.stream_offsets(h)[1]
    bad = bytearray(arc)
    bad[o + 8 + 28 + 100] ^= 0x10  # inside the first sequence block's payload
    with pytest.raises(BscError):
        A.decompress_archive(bytes(bad))
    with pytest.raises(A.ArchiveError):
        A.decompress_archive(bytes(arc[:-10]))
    bad = bytearray(arc)
    struct.pack_into("<Q", bad, 8 + 12, h.num_reads + 1)  # more reads than the streams hold
    with pytest.raises(A.ArchiveError):
        A.decompress_archive(bytes(bad))


def test_libbsc_archive_is_rejected():
    from pathlib import Path
    g = (Path(__file__).resolve().parent / "golden" / "ref_ne50k.qz").read_bytes()
    with pytest.raises(BscError):
        A.decompress_archive(g)


def test_archive_edge_cases():
    """No records, one record, empty reads, a single long name."""
    for text in (b"", b"@r\nACGT\n+\nIIII\n", b"@a\n\n+\n\n@b\nA\n+\n#\n",
                 b"@" + b"n" * 5000 + b"\n" + b"ACGT" * 500 + b"\n+\n" + b"F" * 2000 + b"\n"):
        arc = A.compress_fastq(text)
        assert A.decompress_archive(arc) == text
        assert A.parse_archive_header(arc).num_reads == text.count(b"\n") // 4


def test_batch_forms_many_tiny_blocks():
    """300 blocks of 1..299 bytes over 1..4 symbols: stored and coded blocks
    mixed in one batch, each equal to its per-block call."""
    rng = np.random.default_rng(12)
    blocks = [rng.integers(65, 65 + int(rng.integers(1, 5)), int(rng.integers(1, 300)), dtype=np.uint8)
              for _ in range(300)]
    print("tiny blocks:", len(blocks), flush=True)
    comp = bsc.compress_batch(blocks)
    assert comp == [bsc.compress_adaptive_no_lzp(b) for b in blocks]
    assert bsc.decompress_batch(comp) == [b.tobytes() for b in blocks]


def test_batch_forms_across_device_sub_batches():
    """60 MB of blocks: the batch calls split them into several device
    sub-batches (32 MB at least, so here one 20 MB block each, small blocks
    joining the current one), overlapping host coding with device sorting.
    The bytes still equal the per-block calls and decode back, without and
    with LZP."""
    rng = np.random.default_rng(14)
    blocks = [
        synth.reads_50mbp_stream(140_000, 5)[:20_000_000],
        rng.integers(65, 69, 5_000, dtype=np.uint8),
        synth.quality_stream(140_000, 150, 6)[:20_000_000],
        np.frombuffer(b"ACGT", np.uint8),
        synth.header_stream(600_000, 7)[:20_000_000],
    ]
    print("sub-batches:", [b.size for b in blocks], flush=True)
    comp = bsc.compress_batch(blocks)
    assert comp == [bsc.compress_adaptive_no_lzp(b) for b in blocks]
    assert bsc.decompress_batch(comp) == [b.tobytes() for b in blocks]
    print("sub-batches with LZP", flush=True)
    lz = bsc.compress_batch(blocks, 16, 128)
    assert lz == [bsc.compress_adaptive(b) for b in blocks]
    assert bsc.decompress_batch(lz) == [b.tobytes() for b in blocks]


def test_raw_segments_rebuilt_from_runs():
    """A block of reads followed by random bytes: some coder segments do not
    shrink and are stored raw, their bytes rebuilt from the device front
    end's runs. The block must hold at least one raw segment, equal the
    per-block call and decode back, alone and in a batch."""
    rng = np.random.default_rng(16)
    blk = np.concatenate([synth.sequence_stream(200_000, 8_000, 150, 9, 0.002)[:1_200_000],
                          rng.integers(0, 256, 1_000_000, dtype=np.uint8)])
    comp = bsc.compress_adaptive_no_lzp(blk)
    nb = comp[28]
    segs = [struct.unpack_from("<ii", comp, 29 + 8 * s) for s in range(nb)]
    print("segments (size, coded):", segs, flush=True)
    assert nb > 1 and any(i == o for i, o in segs) and any(o < i for i, o in segs)
    assert bsc.compress_batch([blk, blk[:300_000]]) == [comp, bsc.compress_adaptive_no_lzp(blk[:300_000])]
    assert bsc.decompress(comp) == blk.tobytes()


def test_trim_releases_workspaces_and_pinned_buffers():
    """qzbsc_trim frees the idle engines' workspaces and the cached pinned
    download buffers; later calls allocate again and give the same bytes."""
    from qz_amd._native import lib
    blocks = [synth.sequence_stream(300_000, 20_000, 150, 17, 0.002), synth.quality_stream(20_000, 150, 18)]
    print("trim:", [b.size for b in blocks], flush=True)
    comp = bsc.compress_batch(blocks)
    one = bsc.compress_adaptive_no_lzp(blocks[0])
    freed = lib().qzbsc_trim()
    assert freed > 0
    assert lib().qzbsc_trim() == 0  # nothing left to free
    assert bsc.compress_batch(blocks) == comp
    assert bsc.compress_adaptive_no_lzp(blocks[0]) == one
    assert bsc.decompress_batch(comp) == [b.tobytes() for b in blocks]


@pytest.mark.parametrize("variable", [False, True])
def test_quality_ctx_archive_auto_rule(variable):
    """From MIN_READS_QUALITY_CTX records on, "auto" writes quality_ctx
    qualities as qz does (compress_impl.rs:203-206), in 500,000-record blobs
    by default; the archive decodes back, as does the BSC-quality archive of
    the same text."""
    text = synth.fastq(A.MIN_READS_QUALITY_CTX, 150, 23, variable=variable)
    print("quality_ctx archive: variable", variable, len(text), "bytes", flush=True)
    arc = A.compress_fastq(text)
    h = A.parse_archive_header(arc)
    assert h.quality_compressor == A.COMPRESSOR_QUALITY_CTX and h.const_qual_len == 0
    assert h.const_seq_len == (0 if variable else 150)
    assert A.decompress_archive(arc) == text
    small = A.compress_fastq(text, quality_ctx_block_size=30_000)  # four blobs, decoded in parallel
    o, n = A.stream_offsets(A.parse_archive_header(small))[2]
    assert len(A.stream_blocks(small, o, n)) == 4
    assert A.decompress_archive(small) == text
    b = A.compress_fastq(text, quality_compressor="bsc")
    assert A.parse_archive_header(b).quality_compressor == A.COMPRESSOR_BSC
    assert A.decompress_archive(b) == text
    print("sizes: quality_ctx", len(arc), "bsc", len(b), flush=True)


def test_concurrent_batch_calls_share_the_engines():
    """Six batch calls at once, each of two device sub-batches, with four
    engines: the calls wait for engines, each keeping one for its device
    work only, and all finish with the bytes of a lone call."""
    from concurrent.futures import ThreadPoolExecutor
    blocks = [synth.sequence_stream(2_000_000, 80_000, 150, 30 + k, 0.002)[:12_000_000] for k in range(3)]
    print("concurrent batch calls:", [b.size for b in blocks], flush=True)
    want = bsc.compress_batch(blocks)
    with ThreadPoolExecutor(6) as ex:
        got = list(ex.map(lambda _: bsc.compress_batch(blocks), range(6)))
    assert all(g == want for g in got)
    assert bsc.decompress_batch(want) == [b.tobytes() for b in blocks]
