"""The .qz v2 container (qz_amd.archive) on the host, no GPU needed.

Pins: the libbsc archive under tests/golden (a real qz archive, 50,000 reads
of 101 bp): its 68 header bytes must be what archive_header writes for its
field values, its three streams must frame as one block each, and each
block's data size must be what records_to_streams gives for 50,000 records
of 101 bases and 101 qualities (5,050,000 and 50,000 x 89 bytes: the 7-bit
packing length). Varints and the bit packing are checked on hand-derived
vectors; the stream layout on hand-assembled bytes.
"""
import struct
from pathlib import Path

import numpy as np
import pytest

from qz_amd import archive as A
from qz_amd import bsc

GOLD = Path(__file__).resolve().parent / "golden" / "ref_ne50k.qz"


def gold():
    return GOLD.read_bytes()


def test_fixture_header_fields_and_bytes():
    g = gold()
    h = A.parse_archive_header(g)
    assert (h.version, h.header_size, h.encoding_type, h.flags) == (2, 68, 0, 0x02)
    assert (h.quality_binning, h.quality_compressor, h.sequence_compressor, h.header_compressor) == (0, 1, 1, 1)
    assert h.num_reads == 50_000
    assert (h.headers_len, h.sequences_len, h.nmasks_len, h.qualities_len) == (203_692, 1_199_090, 0, 1_222_078)
    assert (h.const_seq_len, h.const_qual_len) == (101, 101)
    assert h.header_size + h.headers_len + h.sequences_len + h.qualities_len == len(g)
    mine = A.archive_header(h.num_reads, h.headers_len, h.sequences_len, h.qualities_len, h.const_seq_len,
                            h.const_qual_len)
    assert mine == g[:68]


def test_fixture_stream_framing_and_sizes():
    g = gold()
    h = A.parse_archive_header(g)
    spans = A.stream_offsets(h)
    blocks = [A.stream_blocks(g, o, n) for o, n in spans]
    assert [len(b) for b in blocks] == [1, 1, 1]
    assert [len(b[0]) for b in blocks] == [203_684, 1_199_082, 1_222_070]
    data_sizes = [struct.unpack_from("<i", bytes(b[0]), 4)[0] for b in blocks]
    # header stream: not derivable without decoding; the other two are
    assert data_sizes[0] == 2_904_234
    assert data_sizes[1] == h.num_reads * h.const_seq_len
    assert data_sizes[2] == h.num_reads * A.packed_len(h.const_qual_len) == 4_450_000


def test_varint_vectors():
    for v, enc in [(0, b"\x00"), (1, b"\x01"), (127, b"\x7f"), (128, b"\x80\x01"), (300, b"\xac\x02"),
                   (16384, b"\x80\x80\x01"), (2**32, b"\x80\x80\x80\x80\x10")]:
        out = bytearray()
        A.write_varint(out, v)
        assert bytes(out) == enc
        assert A.read_varint(enc + b"zz", 0) == (v, len(enc))
    with pytest.raises(A.ArchiveError):
        A.read_varint(b"\x80\x80", 0)


def test_quality_packing_vectors():
    # phred 40 ('I') then 2 ('#'): 40 | 2 << 7 = 296 -> bytes 0x28, 0x01
    assert A.pack_qualities(b"I#") == b"\x28\x01"
    # eight codes of 127 fill exactly 7 bytes of ones
    assert A.pack_qualities(bytes([160] * 8)) == b"\xff" * 7
    # one code: 7 bits in one byte; phred 0 -> 0
    assert A.pack_qualities(b"!") == b"\x00"
    assert A.pack_qualities(b"\"") == b"\x01"
    assert A.packed_len(101) == 89 and A.packed_len(8) == 7 and A.packed_len(9) == 8
    rng = np.random.default_rng(5)
    for n in (1, 7, 8, 9, 101, 150, 251):
        q = bytes(rng.integers(33, 161, n, dtype=np.uint8))
        p = A.pack_qualities(q)
        assert len(p) == A.packed_len(n)
        assert A.unpack_qualities(p, n) == q
    # out of the 7-bit range: clamped (columnar.rs encode = phred.min(127); ASCII below '!' saturates at 0)
    assert A.unpack_qualities(A.pack_qualities(bytes([200, 20])), 2) == bytes([160, 33])


def _records(n, rng, const=True):
    recs = []
    for i in range(n):
        L = 12 if const else int(rng.integers(0, 20))
        seq = bytes(rng.choice(list(b"ACGTN"), L).astype(np.uint8))
        qual = bytes(rng.integers(35, 75, L, dtype=np.uint8))
        recs.append(A.FastqRecord(b"@r%d/%d" % (i, L), seq, qual))
    return recs


def test_stream_layout_by_hand():
    recs = [A.FastqRecord(b"@a", b"ACGT", b"II#!"), A.FastqRecord(b"@bb", b"GG", b"#I")]
    cs, cq = A.detect_const_lengths(recs)
    assert (cs, cq) == (0, 0)
    h, s, q = A.records_to_streams(recs, cs, cq)
    assert h == b"\x02@a\x03@bb"
    assert s == b"\x04ACGT\x02GG"
    assert q == b"\x04" + A.pack_qualities(b"II#!") + b"\x02" + A.pack_qualities(b"#I")
    recs2 = [A.FastqRecord(b"@a", b"ACGT", b"II#!"), A.FastqRecord(b"@b", b"TTTT", b"!!!!")]
    assert A.detect_const_lengths(recs2) == (4, 4)
    h, s, q = A.records_to_streams(recs2, 4, 4)
    assert s == b"ACGTTTTT"
    assert q == A.pack_qualities(b"II#!") + A.pack_qualities(b"!!!!")


@pytest.mark.parametrize("const", [True, False])
def test_streams_to_records_roundtrip(const):
    rng = np.random.default_rng(9)
    recs = _records(300, rng, const)
    cs, cq = A.detect_const_lengths(recs)
    hs, ss, qs = A.records_to_streams(recs, cs, cq)
    hdr = A.parse_archive_header(A.archive_header(len(recs), len(hs), len(ss), len(qs), cs, cq) +
                                 hs + ss + qs)
    back = A.streams_to_records(hdr, hs, ss, qs)
    assert back == recs
    text = A.format_fastq(recs)
    assert A.parse_fastq(text) == recs


def test_header_and_framing_errors():
    g = gold()
    with pytest.raises(A.ArchiveError):
        A.parse_archive_header(b"XZ" + g[2:])
    with pytest.raises(A.ArchiveError):
        A.parse_archive_header(g[:2] + b"\x03" + g[3:])
    with pytest.raises(A.ArchiveError):
        A.parse_archive_header(g[:-1])  # streams run past the end
    h = A.parse_archive_header(g)
    o, n = A.stream_offsets(h)[1]
    with pytest.raises(A.ArchiveError):
        A.stream_blocks(g, o, n - 1)    # framing does not fill the stated length
    bad = bytearray(g)
    struct.pack_into("<I", bad, o + 4, 65 * 1024 * 1024)
    with pytest.raises(A.ArchiveError):
        A.stream_blocks(bytes(bad), o, n)
    with pytest.raises(A.ArchiveError):
        A.parse_fastq(b"@a\nAC\n+\n")


def test_unsupported_features_are_rejected():
    g = bytearray(gold())
    g[8] = 4  # encoding type 4 (sequence hints)
    with pytest.raises(A.ArchiveError):
        A.decompress_streams(bytes(g))
    g = bytearray(gold())
    g[8 + 3] = 4  # quality_ctx quality stream
    with pytest.raises(A.ArchiveError):
        A.decompress_streams(bytes(g))


def test_stored_blocks_batch_without_device():
    """Blocks of <= 28 bytes are stored: the batch forms need no device."""
    blocks = [b"", b"A", b"ACGT" * 7]
    comp = bsc.compress_batch([np.frombuffer(b, np.uint8) for b in blocks])
    assert [len(c) for c in comp] == [28 + len(b) for b in blocks]
    assert bsc.decompress_batch(comp[1:]) == blocks[1:]


def _odd_records(rng, n):
    """Long names (2-byte varints), empty reads, out-of-range quality bytes."""
    recs = []
    for i in range(n):
        name = b"@" + b"x" * int(rng.integers(0, 300)) + b"%d" % i
        L = int(rng.integers(0, 200))
        seq = bytes(rng.choice(list(b"ACGTN"), L).astype(np.uint8))
        qual = bytes(rng.integers(33, 161, L, dtype=np.uint8))
        recs.append(A.FastqRecord(name, seq, qual))
    return recs


@pytest.mark.parametrize("kind", ["const", "variable", "odd"])
def test_native_conversions_match_record_functions(kind):
    rng = np.random.default_rng(13)
    recs = _odd_records(rng, 400) if kind == "odd" else _records(500, rng, kind == "const")
    text = A.format_fastq(recs)
    cs, cq = A.detect_const_lengths(recs)
    h, s, q = A.records_to_streams(recs, cs, cq)
    assert A.fastq_to_streams(text) == (len(recs), cs, cq, h, s, q)
    hdr = A.parse_archive_header(A.archive_header(len(recs), len(h), len(s), len(q), cs, cq) + h + s + q)
    assert A.streams_to_fastq(hdr, h, s, q) == text
    # one record more than the streams hold
    hdr.num_reads += 1
    with pytest.raises(A.ArchiveError):
        A.streams_to_fastq(hdr, h, s, q)


def test_native_fastq_edge_cases():
    assert A.fastq_to_streams(b"") == (0, 0, 0, b"", b"", b"")
    # no records: nothing to assemble, with or without an output buffer
    hdr = A.parse_archive_header(A.archive_header(0, 0, 0, 0))
    assert A.streams_to_fastq(hdr, b"", b"", b"") == b""
    assert A.streams_to_fastq(hdr, np.zeros(0, np.uint8), np.zeros(0, np.uint8), np.zeros(0, np.uint8)) == b""
    # no trailing newline
    assert A.fastq_to_streams(b"@a\nAC\n+\nII")[:3] == (1, 2, 2)
    # qz's reader rules (io/fastq.rs:141-197): a missing line, no '+' line,
    # or a quality of another length than the sequence fail ...
    for bad in (b"@a\nAC\n+\n", b"@a\nAC\n-\nII\n", b"@a\nAC\n+\nI\n", b"@a\nAC\n+\nIII\n"):
        with pytest.raises(A.ArchiveError):
            A.fastq_to_streams(bad)
        with pytest.raises(A.ArchiveError):
            A.parse_fastq(bad)
    # ... an id line is taken as it is, with or without '@' ...
    assert A.fastq_to_streams(b"a\nAC\n+\nII\n")[3] == b"\x01a"
    assert A.parse_fastq(b"\nAC\n+x\nII\n") == [A.FastqRecord(b"", b"AC", b"II")]
    # ... and trailing '\r's are dropped: CRLF text gives the LF streams
    lf = b"@r1 x\nACGT\n+\nIIII\n@r2\nGG\n+r2\n#I\n"
    crlf = lf.replace(b"\n", b"\r\n")
    assert A.fastq_to_streams(crlf) == A.fastq_to_streams(lf)
    assert A.parse_fastq(crlf) == A.parse_fastq(lf)
    assert A.fastq_to_streams(b"@a\r\r\nAC\r\n+\r\nII\r\r")[3] == b"\x02@a"


@pytest.mark.parametrize("kind", ["quality_at_signs", "sequence_at_signs", "crlf"])
def test_native_conversions_parallel_stretches(kind):
    """Texts over 4 MB are cut at record starts and converted in parallel.
    Quality lines beginning with '@' (phred 31) must not be taken for record
    starts. Sequence lines beginning with '@' above quality lines beginning
    with '+' do look like record starts; the cut is then not confirmed and
    the text is parsed in one piece. CRLF text gives the streams of its LF
    form. All give the per-record streams."""
    rng = np.random.default_rng(15)
    recs = []
    for i in range(40_000):
        L = int(rng.integers(60, 200))
        seq = bytes(rng.choice(list(b"ACGTN"), L).astype(np.uint8))
        qual = bytes(rng.integers(35, 75, L, dtype=np.uint8))
        if kind == "quality_at_signs" or i % 3:
            qual = b"@" + qual[1:]
        if kind == "sequence_at_signs":
            seq, qual = b"@" + seq[1:], b"+" + qual[1:]
        recs.append(A.FastqRecord(b"@read%d" % i, seq, qual))
    text = A.format_fastq(recs)
    assert len(text) > 8 << 20
    cs, cq = A.detect_const_lengths(recs)
    h, s, q = A.records_to_streams(recs, cs, cq)
    src = text.replace(b"\n", b"\r\n") if kind == "crlf" else text  # CRLF reads as LF
    assert A.fastq_to_streams(src) == (len(recs), cs, cq, h, s, q)
    hdr = A.parse_archive_header(A.archive_header(len(recs), len(h), len(s), len(q), cs, cq) + h + s + q)
    assert A.streams_to_fastq(hdr, h, s, q) == text
    # a malformed record deep inside a stretch still fails the whole text
    bad = text[:len(text) // 2] + b"\n" + text[len(text) // 2:]
    with pytest.raises(A.ArchiveError):
        A.fastq_to_streams(bad)


def test_quality_ctx_archive_small():
    """quality_compressor="quality_ctx": the header records compressor 4 and
    no constant quality length; the quality stream is [u32 blobs][u32 len,
    blob]... of quality_ctx blobs (compress_impl.rs:100-126, write_blocks),
    one per quality_ctx_block_size records. Streams this small are stored
    BSC blocks, so no device is needed."""
    from qz_amd import quality_ctx as QC
    recs = [A.FastqRecord(b"@a", b"ACGT", b"II#!"), A.FastqRecord(b"@b", b"GGA", b"#I5"),
            A.FastqRecord(b"@c", b"TTTT", b"FFFF")]
    text = A.format_fastq(recs)
    for per in (A.QUALITY_CTX_BLOCK_SIZE, 2, 1):
        arc = A.compress_fastq(text, quality_compressor="quality_ctx", quality_ctx_block_size=per)
        h = A.parse_archive_header(arc)
        assert (h.quality_compressor, h.const_qual_len) == (A.COMPRESSOR_QUALITY_CTX, 0)
        o, n = A.stream_offsets(h)[2]
        blobs = [bytes(b) for b in A.stream_blocks(arc, o, n)]
        spans = [(i, min(len(recs), i + per)) for i in range(0, len(recs), per)]
        assert blobs == [QC.compress_qualities_ctx([r.quality for r in recs[a:b]], [r.sequence for r in recs[a:b]])
                         for a, b in spans]
        assert A.decompress_archive(arc) == text
        with pytest.raises(A.ArchiveError):
            A.decompress_streams(arc)
    # "auto" keeps BSC below MIN_READS_QUALITY_CTX records
    assert A.parse_archive_header(A.compress_fastq(text)).quality_compressor == A.COMPRESSOR_BSC
    # a blob count that does not cover the reads fails
    arc = bytearray(A.compress_fastq(text, quality_compressor="quality_ctx", quality_ctx_block_size=1))
    h = A.parse_archive_header(bytes(arc))
    struct.pack_into("<Q", arc, 8 + 12, 2)  # two reads stated, three blobs of one read
    with pytest.raises(A.ArchiveError):
        A.decompress_archive(bytes(arc))
