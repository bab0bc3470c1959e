"""The .qz v2 archive around the BSC streams: qz's default container,
written and read with this library's block coder.

An archive is an 8-byte prefix (magic "QZ", version 2, a reserved byte, the
u32 header size), a 52-byte header body (+8 when constant lengths are
recorded), then the header, sequence, N-mask and quality streams, each in
the BSC multi-block framing [u32 num_blocks][u32 len, block]... (an empty
stream takes no bytes at all).

Writer, following crates/qz-lib/src/compression:
- records_to_streams (mod.rs:280-320): header stream = varint(len) + id per
  record (the id keeps its '@'); sequence stream = the bases, each prefixed
  by varint(len) unless all reads share one length; quality stream = each
  record's qualities as 7-bit Phred codes packed LSB first and flushed to a
  byte boundary per record (columnar.rs:103-134, QualityBinning::None),
  varint(len)-prefixed unless all share one length;
- detect_const_lengths (mod.rs:329-354);
- compress_stream_to_bsc_blocks (mod.rs:388-413): 25 MiB pieces, each one
  adaptive-QLFC BSC block without LZP;
- write_archive_header (mod.rs:430-482) and the stream lengths
  (compress_impl.rs:757-760: 4 + the framed blocks, 0 when no block).

Reader, following decompress_impl.rs: the header parse of
decompress_streaming_bsc (275-335), the stream walk of
stream_decompressor_inner (72-135: blocks in batches of 8, a block claiming
more than 64 MB is rejected), the offset scans (444-496) and the record
assembly of reconstruct_records_parallel (504-636: "id\\nseq\\n+\\nqual\\n").

This build covers the container qz writes by default: encoding 0, lossless
qualities (binning code 0), BSC for the sequence and header streams
(compressor code 1), and for the quality stream either BSC over 7-bit packed
qualities (code 1) or quality_ctx blobs (code 4), which qz itself picks from
MIN_READS_QUALITY_CTX records on (compress_impl.rs:203-206). Archives using
other encodings or codecs raise ArchiveError.

Blocks go through the batch C-ABI (bsc.compress_batch /
bsc.decompress_batch), where qz fans out per-block libbsc calls over rayon:
the blocks of all three streams share one device batch when compressing,
and when decoding each batch takes 8 blocks of every stream (qz's three
decompressor threads, 8 blocks each). The FASTQ <-> streams conversions run
natively (qzbsc_fastq_to_streams / qzbsc_streams_to_fastq); the per-record
Python functions below state the same layout and check them in the tests.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

from . import bsc
from ._native import lib, u8p

ARCHIVE_MAGIC = b"QZ"
ARCHIVE_VERSION = 2
V2_PREFIX_SIZE = 8
BODY_SIZE = 52                # 9 + 2 + 1 + 40 (mod.rs:446)
CONST_LEN_SIZE = 8
DECOMPRESS_BATCH_SIZE = 8     # mod.rs:30
MAX_BLOCK_BYTES = 64 * 1024 * 1024  # decompress_impl.rs:105
BITS_PER_QUAL = 7             # QualityBinning::None (columnar.rs:34)

ENCODING_PLAIN = 0
BINNING_LOSSLESS = 0
COMPRESSOR_BSC = 1
COMPRESSOR_QUALITY_CTX = 4       # QualityCompressor::QualityCtx (mod.rs:109)
MIN_READS_QUALITY_CTX = 100_000  # mod.rs:34
QUALITY_CTX_BLOCK_SIZE = 500_000  # records per quality_ctx blob (cli.rs:132)


class ArchiveError(ValueError):
    """A malformed archive, or one using a feature this build does not read."""


@dataclass
class FastqRecord:
    id: bytes          # header line including the leading '@'
    sequence: bytes
    quality: bytes


@dataclass
class ArchiveHeader:
    version: int
    header_size: int
    encoding_type: int
    flags: int
    quality_binning: int
    quality_compressor: int
    sequence_compressor: int
    header_compressor: int
    num_reads: int
    headers_len: int
    sequences_len: int
    nmasks_len: int
    qualities_len: int
    const_seq_len: int
    const_qual_len: int

    @property
    def data_offset(self) -> int:
        return self.header_size


# ---- varints and quality packing -------------------------------------------

def write_varint(out: bytearray, value: int) -> None:
    """LEB128, 7 bits per byte, least significant first (dna_utils.rs:285-291)."""
    while value >= 0x80:
        out.append((value & 0x7F) | 0x80)
        value >>= 7
    out.append(value)


def read_varint(buf, pos: int):
    """Returns (value, next position); raises ArchiveError when truncated."""
    value = shift = 0
    while True:
        if pos >= len(buf):
            raise ArchiveError("truncated varint")
        b = buf[pos]
        pos += 1
        value |= (b & 0x7F) << shift
        if not b & 0x80:
            return value, pos
        shift += 7
        if shift > 63:
            raise ArchiveError("varint too long")


def packed_len(nqual: int) -> int:
    return (nqual * BITS_PER_QUAL + 7) // 8


def _codes_to_packed(codes: np.ndarray) -> np.ndarray:
    """(N, L) 7-bit codes -> (N, packed_len(L)) bytes, LSB first per record."""
    n, length = codes.shape
    bits = ((codes[:, :, None] >> np.arange(BITS_PER_QUAL, dtype=np.uint8)) & 1).reshape(n, length * BITS_PER_QUAL)
    pad = packed_len(length) * 8 - length * BITS_PER_QUAL
    if pad:
        bits = np.concatenate([bits, np.zeros((n, pad), np.uint8)], axis=1)
    return np.packbits(bits, axis=1, bitorder="little")


def _packed_to_codes(packed: np.ndarray, length: int) -> np.ndarray:
    """(N, packed_len(L)) bytes -> (N, L) 7-bit codes."""
    n = packed.shape[0]
    bits = np.unpackbits(packed, axis=1, bitorder="little")[:, :length * BITS_PER_QUAL]
    bits = bits.reshape(n, length, BITS_PER_QUAL)
    return (bits << np.arange(BITS_PER_QUAL, dtype=np.uint8)).sum(axis=2, dtype=np.uint16).astype(np.uint8)


def pack_qualities(qual: bytes) -> bytes:
    """One record: Phred+33 ASCII -> min(q - 33, 127) (0 below '!') as 7-bit
    codes, packed LSB first (columnar.rs:103-134 with QualityBinning::None,
    encode = phred.min(127))."""
    if not qual:
        return b""
    q = np.frombuffer(qual, dtype=np.uint8).astype(np.int16)
    codes = np.clip(q - 33, 0, 127).astype(np.uint8)
    return _codes_to_packed(codes[None, :])[0].tobytes()


def unpack_qualities(packed: bytes, nqual: int) -> bytes:
    """Inverse of pack_qualities: codes + 33 (columnar.rs:167-207)."""
    if nqual == 0:
        return b""
    a = np.frombuffer(packed, dtype=np.uint8)
    if a.size < packed_len(nqual):
        raise ArchiveError("truncated quality record")
    return (_packed_to_codes(a[None, :packed_len(nqual)], nqual)[0] + 33).tobytes()


# ---- FASTQ text ---------------------------------------------------------------

def parse_fastq(data: bytes) -> list:
    """Four-line FASTQ records as qz's FastqReader reads them (io/fastq.rs:
    141-197): the id line as it is, the third line starting with '+',
    trailing '\r's dropped from each line, the quality as long as the
    sequence."""
    lines = data.split(b"\n")
    if lines and lines[-1] == b"":
        lines.pop()
    if len(lines) % 4:
        raise ArchiveError("FASTQ: line count is not a multiple of 4")
    recs = []
    for i in range(0, len(lines), 4):
        h, s, p, q = lines[i:i + 4]
        if not p.startswith(b"+"):
            raise ArchiveError(f"FASTQ: record {i // 4} has no '+' separator line")
        h, s, q = h.rstrip(b"\r"), s.rstrip(b"\r"), q.rstrip(b"\r")
        if len(q) != len(s):
            raise ArchiveError(f"FASTQ: record {i // 4}: sequence length {len(s)} != quality length {len(q)}")
        recs.append(FastqRecord(h, s, q))
    return recs


def format_fastq(records) -> bytes:
    """decompress_impl.rs:726-748: id, sequence, '+', quality, one per line."""
    parts = []
    for r in records:
        parts += [r.id, b"\n", r.sequence, b"\n+\n", r.quality, b"\n"]
    return b"".join(parts)


# ---- writer -----------------------------------------------------------------

def detect_const_lengths(records):
    """mod.rs:329-354: (const_seq_len, const_qual_len), 0 meaning variable."""
    if not records:
        return 0, 0
    sl = len(records[0].sequence)
    ql = len(records[0].quality)
    cs = sl if all(len(r.sequence) == sl for r in records) else 0
    cq = ql if ql and all(len(r.quality) == ql for r in records) else 0
    return cs, cq


def records_to_streams(records, const_seq_len: int, const_qual_len: int):
    """mod.rs:280-320 for encoding 0 with lossless qualities: returns the
    header, sequence and quality streams."""
    h = bytearray()
    for r in records:
        write_varint(h, len(r.id))
        h += r.id
    if const_seq_len:
        s = b"".join(r.sequence for r in records)
    else:
        sb = bytearray()
        for r in records:
            write_varint(sb, len(r.sequence))
            sb += r.sequence
        s = bytes(sb)
    if const_qual_len:
        if records:
            q = np.frombuffer(b"".join(r.quality for r in records), dtype=np.uint8)
            codes = np.clip(q.astype(np.int16) - 33, 0, 127).astype(np.uint8).reshape(len(records), const_qual_len)
            qs = _codes_to_packed(codes).tobytes()
        else:
            qs = b""
    else:
        qb = bytearray()
        for r in records:
            write_varint(qb, len(r.quality))
            qb += pack_qualities(r.quality)
        qs = bytes(qb)
    return bytes(h), s, qs


def archive_header(num_reads: int, headers_len: int, sequences_len: int, qualities_len: int,
                   const_seq_len: int = 0, const_qual_len: int = 0,
                   quality_compressor: int = COMPRESSOR_BSC) -> bytes:
    """mod.rs:430-482 for encoding 0, lossless qualities, BSC for the header
    and sequence streams and BSC or quality_ctx for the quality stream:
    prefix, body, and the constant lengths when either is set."""
    has_const = const_seq_len > 0 or const_qual_len > 0
    header_size = V2_PREFIX_SIZE + BODY_SIZE + (CONST_LEN_SIZE if has_const else 0)
    out = bytearray(ARCHIVE_MAGIC)
    out += bytes([ARCHIVE_VERSION, 0])
    out += struct.pack("<I", header_size)
    out += bytes([ENCODING_PLAIN, 0x02 if has_const else 0x00, BINNING_LOSSLESS, quality_compressor, COMPRESSOR_BSC,
                  COMPRESSOR_BSC, 0, 0, 0])  # quality model / delta / dictionary disabled
    out += struct.pack("<H", 0)              # template prefix length
    out += bytes([0])                        # no comment
    out += struct.pack("<QQQQQ", num_reads, headers_len, sequences_len, 0, qualities_len)
    if has_const:
        out += struct.pack("<II", const_seq_len, const_qual_len)
    return bytes(out)


def compress_streams(streams, block_size: int = bsc.BSC_BLOCK_SIZE) -> list:
    """Each stream in the multi-block framing (b"" for an empty stream). The
    blocks of all streams go through one batch call: device sub-batches, and
    the host coder on all of them at once."""
    buf = np.frombuffer(b"".join(streams), dtype=np.uint8)
    return _compress_stream_buffer(buf, [len(x) for x in streams], block_size)


def _compress_stream_buffer(buf: np.ndarray, lens, block_size: int) -> list:
    """compress_streams for streams laid end to end in buf (lengths lens)."""
    sizes, owner = [], []
    for k, n in enumerate(lens):
        for i in range(0, n, block_size):
            sizes.append(min(block_size, n - i))
            owner.append(k)
    blocks = bsc.compress_batch_buffer(buf, sizes) if sizes else []
    out = []
    for k in range(len(lens)):
        mine = [b for b, o in zip(blocks, owner) if o == k]
        if not mine:
            out.append(b"")
            continue
        parts = [struct.pack("<I", len(mine))]
        for b in mine:
            parts += [struct.pack("<I", len(b)), b]
        out.append(b"".join(parts))
    return out


def compress_stream(data: bytes, block_size: int = bsc.BSC_BLOCK_SIZE) -> bytes:
    """One stream in the multi-block framing; b"" for an empty stream."""
    return compress_streams([data], block_size)[0]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(u8p)


def fastq_to_stream_buffer(fastq: bytes):
    """Native FASTQ scan (qzbsc_fastq_to_streams) into one uint8 array that
    holds the header, sequence and quality streams end to end:
    (num_reads, const_seq_len, const_qual_len, array, (hl, sl, ql))."""
    import ctypes as C
    src = np.frombuffer(fastq, dtype=np.uint8)
    info = (C.c_longlong * 6)()
    rc = lib().qzbsc_fastq_to_streams(_ptr(src), src.size, None, None, None, info)
    if rc < 0:
        raise ArchiveError("malformed FASTQ (a record needs four lines, '@' and '+')")
    n, cs, cq, hl, sl, ql = (int(v) for v in info)
    buf = np.empty(hl + sl + ql + 1, np.uint8)  # + 1: three valid pointers when streams are empty
    rc = lib().qzbsc_fastq_to_streams(_ptr(src), src.size, _ptr(buf), _ptr(buf[hl:]), _ptr(buf[hl + sl:]), info)
    if rc < 0:
        raise ArchiveError("malformed FASTQ")
    return n, cs, cq, buf[:hl + sl + ql], (hl, sl, ql)


def fastq_to_streams(fastq: bytes):
    """FASTQ text -> (num_reads, const_seq_len, const_qual_len, header,
    sequence, quality streams) by the native scan (qzbsc_fastq_to_streams);
    the same streams as records_to_streams(parse_fastq(...))."""
    n, cs, cq, buf, (hl, sl, ql) = fastq_to_stream_buffer(fastq)
    return n, cs, cq, buf[:hl].tobytes(), buf[hl:hl + sl].tobytes(), buf[hl + sl:].tobytes()


def compress_records(records, block_size: int = bsc.BSC_BLOCK_SIZE) -> bytes:
    cs, cq = detect_const_lengths(records)
    hz, sz, qz = compress_streams(records_to_streams(records, cs, cq), block_size)
    return archive_header(len(records), len(hz), len(sz), len(qz), cs, cq) + hz + sz + qz


def _threads(n: int) -> int:
    import os
    return max(1, min(n, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)))


def _fastq_columns(fastq: bytes, n: int):
    """(sequences, qualities) of the n records of valid FASTQ text, as lists
    of bytes (the text was validated by the native scan)."""
    lines = fastq.split(b"\n")
    return [x.rstrip(b"\r") for x in lines[1:4 * n:4]], [x.rstrip(b"\r") for x in lines[3:4 * n:4]]


def quality_ctx_stream(qualities, sequences, block_size: int = QUALITY_CTX_BLOCK_SIZE) -> bytes:
    """The quality stream of a quality_ctx archive (compress_impl.rs:100-126,
    written by write_blocks like the BSC streams): [u32 blobs][u32 len, blob]...,
    one quality_ctx blob per block_size records, compressed in parallel."""
    from concurrent.futures import ThreadPoolExecutor
    from . import quality_ctx as QC
    n = len(qualities)
    if n == 0:
        return b""
    spans = [(i, min(n, i + block_size)) for i in range(0, n, block_size)]
    with ThreadPoolExecutor(_threads(len(spans))) as ex:
        blobs = list(ex.map(lambda sp: QC.compress_qualities_ctx(qualities[sp[0]:sp[1]], sequences[sp[0]:sp[1]]),
                            spans))
    parts = [struct.pack("<I", len(blobs))]
    for b in blobs:
        parts += [struct.pack("<I", len(b)), b]
    return b"".join(parts)


def compress_fastq(fastq: bytes, block_size: int = bsc.BSC_BLOCK_SIZE, quality_compressor: str = "auto",
                   quality_ctx_block_size: int = QUALITY_CTX_BLOCK_SIZE) -> bytes:
    """FASTQ text -> .qz v2 archive bytes. quality_compressor: "bsc" (7-bit
    packed qualities in BSC blocks), "quality_ctx", or "auto", qz's rule for
    lossless qualities (compress_impl.rs:203-206): quality_ctx from
    MIN_READS_QUALITY_CTX records on. quality_ctx archives record no
    constant quality length (compress_impl.rs:210-215)."""
    if quality_compressor not in ("auto", "bsc", "quality_ctx"):
        raise ValueError(f"unknown quality compressor {quality_compressor!r}")
    n, cs, cq, buf, lens = fastq_to_stream_buffer(fastq)
    ctx = quality_compressor == "quality_ctx" or (quality_compressor == "auto" and n >= MIN_READS_QUALITY_CTX)
    if not ctx:
        hz, sz, qz = _compress_stream_buffer(buf, lens, block_size)
        return archive_header(n, len(hz), len(sz), len(qz), cs, cq) + hz + sz + qz
    hl, sl, _ = lens
    hz, sz = _compress_stream_buffer(buf[:hl + sl], [hl, sl], block_size)
    seqs, quals = _fastq_columns(fastq, n)
    qz = quality_ctx_stream(quals, seqs, quality_ctx_block_size)
    return archive_header(n, len(hz), len(sz), len(qz), cs, 0, COMPRESSOR_QUALITY_CTX) + hz + sz + qz


# ---- reader -----------------------------------------------------------------

def parse_archive_header(buf) -> ArchiveHeader:
    """decompress_impl.rs:285-335, plus the checks that the header and the
    streams fit the buffer."""
    mv = memoryview(buf)
    if len(mv) < V2_PREFIX_SIZE + BODY_SIZE:
        raise ArchiveError("archive shorter than its header")
    if bytes(mv[0:2]) != ARCHIVE_MAGIC:
        raise ArchiveError("Not a QZ archive (missing magic bytes)")
    version = mv[2]
    if version > ARCHIVE_VERSION:
        raise ArchiveError(f"Archive version {version} is newer than this build supports (max {ARCHIVE_VERSION})")
    (header_size,) = struct.unpack_from("<I", mv, 4)
    b = V2_PREFIX_SIZE
    flags = mv[b + 1]
    num_reads, hl, sl, nl, ql = struct.unpack_from("<QQQQQ", mv, b + 12)
    cs = cq = 0
    if flags & 0x02:
        if len(mv) < b + BODY_SIZE + CONST_LEN_SIZE:
            raise ArchiveError("archive shorter than its constant-length fields")
        cs, cq = struct.unpack_from("<II", mv, b + BODY_SIZE)
    h = ArchiveHeader(version=version, header_size=header_size, encoding_type=mv[b], flags=flags,
                      quality_binning=mv[b + 2], quality_compressor=mv[b + 3], sequence_compressor=mv[b + 4],
                      header_compressor=mv[b + 5], num_reads=num_reads, headers_len=hl, sequences_len=sl,
                      nmasks_len=nl, qualities_len=ql, const_seq_len=cs, const_qual_len=cq)
    if header_size < b + BODY_SIZE or header_size > len(mv):
        raise ArchiveError(f"bad header size {header_size}")
    if header_size + hl + sl + nl + ql > len(mv):
        raise ArchiveError("archive truncated: streams extend past the end")
    return h


def stream_offsets(h: ArchiveHeader):
    """(offset, length) of the header, sequence and quality streams
    (decompress_impl.rs:355-359: N-masks lie between sequences and qualities)."""
    ho = h.data_offset
    so = ho + h.headers_len
    qo = so + h.sequences_len + h.nmasks_len
    return (ho, h.headers_len), (so, h.sequences_len), (qo, h.qualities_len)


def stream_blocks(buf, offset: int, length: int) -> list:
    """The compressed blocks of one framed stream (stream_decompressor_inner,
    decompress_impl.rs:81-113), checked to fill exactly `length` bytes."""
    if length == 0:
        return []
    mv = memoryview(buf)
    end = offset + length
    if length < 4 or end > len(mv):
        raise ArchiveError("truncated stream")
    (nb,) = struct.unpack_from("<I", mv, offset)
    p = offset + 4
    blocks = []
    for _ in range(nb):
        if p + 4 > end:
            raise ArchiveError("truncated block length")
        (ln,) = struct.unpack_from("<I", mv, p)
        p += 4
        if ln > MAX_BLOCK_BYTES:
            raise ArchiveError(f"BSC block claims {ln} bytes (max 64 MB) - archive may be corrupt")
        if p + ln > end:
            raise ArchiveError("truncated block data")
        blocks.append(mv[p:p + ln])
        p += ln
    if p != end:
        raise ArchiveError("stream length does not match its blocks")
    return blocks


def decompress_stream(buf, offset: int, length: int, batch: int = DECOMPRESS_BATCH_SIZE) -> bytes:
    """All blocks of one stream, decoded `batch` blocks per device batch."""
    blocks = stream_blocks(buf, offset, length)
    out = []
    for i in range(0, len(blocks), batch):
        out += bsc.decompress_batch(blocks[i:i + batch])
    return b"".join(out)


def _check_supported(h: ArchiveHeader) -> None:
    if h.encoding_type != ENCODING_PLAIN:
        raise ArchiveError(f"encoding type {h.encoding_type} is not read by this build")
    if h.quality_binning != BINNING_LOSSLESS:
        raise ArchiveError(f"quality binning {h.quality_binning} is not read by this build")
    if h.qualities_len and h.quality_compressor not in (COMPRESSOR_BSC, COMPRESSOR_QUALITY_CTX):
        raise ArchiveError(f"quality compressor {h.quality_compressor} is not read by this build")
    if h.sequence_compressor != COMPRESSOR_BSC or h.header_compressor != COMPRESSOR_BSC:
        raise ArchiveError("only BSC sequence and header streams are read by this build")
    if h.nmasks_len:
        raise ArchiveError("N-mask streams are not read by this build")


def decompress_streams(buf, batch: int = DECOMPRESS_BATCH_SIZE):
    """(header, raw streams h, s, q) of an archive whose three streams are
    BSC. The three streams advance together: each batch call takes up to
    `batch` blocks of every stream (qz runs three decompressor threads, each
    8 blocks at a time), so one device batch and one host coder pool serve
    all three."""
    h = parse_archive_header(buf)
    _check_supported(h)
    if h.qualities_len and h.quality_compressor == COMPRESSOR_QUALITY_CTX:
        raise ArchiveError("the quality stream is quality_ctx, not BSC: decode the archive with decompress_archive")
    per = [stream_blocks(buf, o, n) for o, n in stream_offsets(h)]
    out = [[], [], []]
    for i in range(0, max(len(b) for b in per), batch):
        chunk, owner = [], []
        for k in range(3):
            part = per[k][i:i + batch]
            chunk += part
            owner += [k] * len(part)
        for k, d in zip(owner, bsc.decompress_batch(chunk)):
            out[k].append(d)
    return h, b"".join(out[0]), b"".join(out[1]), b"".join(out[2])


def _scan_lengths(data: bytes, count: int, what: str):
    """Varint-prefixed records: (starts, lengths) (decompress_impl.rs:444-471)."""
    starts = np.empty(count, np.int64)
    lens = np.empty(count, np.int64)
    pos = 0
    for i in range(count):
        try:
            ln, pos = read_varint(data, pos)
        except ArchiveError:
            raise ArchiveError(f"truncated {what} stream at record {i}") from None
        starts[i] = pos
        lens[i] = ln
        pos += ln if what != "quality" else packed_len(ln)
    if pos > len(data):
        raise ArchiveError(f"truncated {what} stream")
    return starts, lens


def streams_to_records(h: ArchiveHeader, hs: bytes, ss: bytes, qs: bytes) -> list:
    """reconstruct_records_parallel (decompress_impl.rs:504-636) for
    encoding 0: records from the decoded streams."""
    n = h.num_reads
    hst, hln = _scan_lengths(hs, n, "header")
    if h.const_seq_len:
        L = h.const_seq_len
        if len(ss) < n * L:
            raise ArchiveError("truncated sequence stream")
        seqs = [ss[i * L:(i + 1) * L] for i in range(n)]
    else:
        sst, sln = _scan_lengths(ss, n, "sequence")
        seqs = [ss[int(a):int(a + b)] for a, b in zip(sst, sln)]
    if h.qualities_len == 0:
        quals = [b""] * n
    elif h.const_qual_len:
        L = h.const_qual_len
        P = packed_len(L)
        if len(qs) < n * P:
            raise ArchiveError("truncated quality stream")
        if n:
            codes = _packed_to_codes(np.frombuffer(qs, np.uint8)[:n * P].reshape(n, P), L) + 33
            flat = codes.tobytes()
            quals = [flat[i * L:(i + 1) * L] for i in range(n)]
        else:
            quals = []
    else:
        qst, qln = _scan_lengths(qs, n, "quality")
        quals = [unpack_qualities(qs[int(a):int(a) + packed_len(int(b))], int(b)) for a, b in zip(qst, qln)]
    return [FastqRecord(hs[int(a):int(a + b)], seqs[i], quals[i]) for i, (a, b) in enumerate(zip(hst, hln))]


def _new_bytes(size: int):
    """An uninitialised bytes object of `size` bytes and a pointer to its
    storage, for native code to fill before the object is used anywhere
    (CPython's PyBytes_FromStringAndSize(NULL, size) idiom)."""
    import ctypes as C
    api = C.pythonapi
    api.PyBytes_FromStringAndSize.restype = C.py_object
    api.PyBytes_FromStringAndSize.argtypes = (C.c_void_p, C.c_ssize_t)
    api.PyBytes_AsString.restype = C.c_void_p
    api.PyBytes_AsString.argtypes = (C.py_object,)
    b = api.PyBytes_FromStringAndSize(None, size)
    return b, C.cast(api.PyBytes_AsString(b), u8p)


def streams_to_fastq(h: ArchiveHeader, hs, ss, qs) -> bytes:
    """The decoded streams (bytes or uint8 arrays) -> FASTQ text by the
    native assembly (qzbsc_streams_to_fastq); the same text as
    format_fastq(streams_to_records(...))."""
    a = [np.frombuffer(x, dtype=np.uint8) if len(x) else np.zeros(1, np.uint8)[:0] for x in (hs, ss, qs)]
    args = (_ptr(a[0]), a[0].size, _ptr(a[1]), a[1].size, _ptr(a[2]), a[2].size, h.num_reads, h.const_seq_len,
            h.const_qual_len)
    size = lib().qzbsc_streams_to_fastq(*args, None, 0)
    if size < 0:
        raise ArchiveError("truncated stream: the streams hold fewer records than the header states")
    out, ptr = _new_bytes(size)
    rc = lib().qzbsc_streams_to_fastq(*args, ptr, size)
    if rc != size:
        raise ArchiveError("record assembly failed")
    return out


def _decompress_stream_buffer(buf):
    """(header, array, (hl, sl, ql)): every block of the three streams in
    one batch call, decoded end to end into one array, which then holds the
    header, sequence and quality streams in that order."""
    h = parse_archive_header(buf)
    _check_supported(h)
    per = [stream_blocks(buf, o, n) for o, n in stream_offsets(h)]
    out, dsz = bsc.decompress_batch_buffer(per[0] + per[1] + per[2])
    n0, n1 = len(per[0]), len(per[0]) + len(per[1])
    return h, out, (sum(dsz[:n0]), sum(dsz[n0:n1]), sum(dsz[n1:]))


def _sequence_list(h: ArchiveHeader, ss) -> list:
    """The records' sequences from a decoded sequence stream."""
    n = h.num_reads
    ss = bytes(ss)
    if h.const_seq_len:
        L = h.const_seq_len
        if len(ss) < n * L:
            raise ArchiveError("truncated sequence stream")
        return [ss[i * L:(i + 1) * L] for i in range(n)]
    st, ln = _scan_lengths(ss, n, "sequence")
    return [ss[a:a + b] for a, b in zip(st.tolist(), ln.tolist())]


def _decompress_quality_ctx_archive(buf, h: ArchiveHeader) -> bytes:
    """decompress_impl.rs:1103-1200 with quality_ctx qualities: the header
    and sequence streams through one BSC batch call, then the quality_ctx
    blobs with the sequences as their context, then the records."""
    from . import quality_ctx as QC
    spans = stream_offsets(h)
    per = [stream_blocks(buf, o, n) for o, n in spans[:2]]
    out, dsz = bsc.decompress_batch_buffer(per[0] + per[1])
    hl = sum(dsz[:len(per[0])])
    hs, ss = out[:hl].tobytes(), out[hl:].tobytes()
    seqs = _sequence_list(h, ss)
    o, n = spans[2]
    try:
        quals = QC.decompress_quality_ctx_multiblock(bytes(memoryview(buf)[o:o + n]), seqs)
    except QC.QualityCtxError as e:
        raise ArchiveError(f"quality_ctx stream: {e}") from None
    hst, hln = _scan_lengths(hs, h.num_reads, "header")
    parts = []
    for i, (a, b) in enumerate(zip(hst.tolist(), hln.tolist())):
        parts += [hs[a:a + b], b"\n", seqs[i], b"\n+\n", quals[i], b"\n"]
    return b"".join(parts)


def decompress_archive(buf, batch: int | None = None) -> bytes:
    """.qz v2 archive bytes -> FASTQ text. By default all blocks go through
    one batch call; batch=k decodes k blocks of every stream per call, as
    qz's decompressor threads do (decompress_streams). Quality streams
    written by quality_ctx decode with the sequences as context."""
    h0 = parse_archive_header(buf)
    _check_supported(h0)
    if h0.qualities_len and h0.quality_compressor == COMPRESSOR_QUALITY_CTX:
        return _decompress_quality_ctx_archive(buf, h0)
    if batch is not None:
        h, hs, ss, qs = decompress_streams(buf, batch)
        return streams_to_fastq(h, hs, ss, qs)
    h, out, (hl, sl, ql) = _decompress_stream_buffer(buf)
    return streams_to_fastq(h, out[:hl], out[hl:hl + sl], out[hl + sl:hl + sl + ql])


def compress_file(fastq_path: str, archive_path: str) -> int:
    with open(fastq_path, "rb") as f:
        data = compress_fastq(f.read())
    with open(archive_path, "wb") as f:
        f.write(data)
    return len(data)


def decompress_file(archive_path: str, fastq_path: str) -> int:
    """Returns the FASTQ size."""
    with open(archive_path, "rb") as f:
        text = decompress_archive(f.read())
    with open(fastq_path, "wb") as f:
        f.write(text)
    return len(text)
