"""Python mirror of qz-lib's quality_ctx codec
(crates/qz-lib/src/compression/quality_ctx.rs) over this library's C-ABI.

Same names and error behaviour: empty input gives an empty blob, a quality
string count that differs from the sequence count, a quality byte below 33
or a quality string longer than its sequence raise, and decoding raises on a
blob that does not hold the requested reads."""
from __future__ import annotations

import struct

import numpy as np

from ._native import lib, u8p, i32p


class QualityCtxError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"quality_ctx {what} failed with error code: {code}")
        self.code = code


def _concat(parts):
    arrs = [np.frombuffer(bytes(p), np.uint8) for p in parts]
    lens = np.array([a.size for a in arrs], dtype=np.int32)
    buf = np.concatenate(arrs) if arrs else np.zeros(0, np.uint8)
    if buf.size == 0:
        buf = np.zeros(1, np.uint8)
    return np.ascontiguousarray(buf), lens


def compress_qualities_ctx(qualities, sequences) -> bytes:
    """quality_ctx.rs:296-391."""
    if len(qualities) == 0:
        return b""
    if len(qualities) != len(sequences):
        raise QualityCtxError(-1, f"compress: {len(qualities)} quality strings but {len(sequences)} sequences")
    q, ql = _concat(qualities)
    s, sl = _concat(sequences)
    cap = 1 + 256 + 10 + int(ql.sum()) * 2 + 64
    out = np.zeros(cap, np.uint8)
    r = lib().qzbsc_quality_ctx_compress(q.ctypes.data_as(u8p), ql.ctypes.data_as(i32p), s.ctypes.data_as(u8p),
                                         sl.ctypes.data_as(i32p), len(qualities), out.ctypes.data_as(u8p), cap)
    if r < 0:
        raise QualityCtxError(int(r), "compress")
    return out[:r].tobytes()


def decompress_qualities_ctx(data, sequences, num_reads: int):
    """quality_ctx.rs:393-505: returns num_reads quality strings."""
    if num_reads == 0:
        return []
    if len(data) == 0:
        raise QualityCtxError(-6, f"decode: stream is empty but num_reads={num_reads}")
    if num_reads > len(sequences):
        raise QualityCtxError(-6, f"decode: num_reads={num_reads} but only {len(sequences)} sequences")
    s, sl = _concat(sequences[:num_reads])
    d = np.frombuffer(bytes(data), np.uint8).copy()
    out = np.zeros(max(int(sl.sum()), 1), np.uint8)
    ql = np.zeros(num_reads, np.int32)
    rc = lib().qzbsc_quality_ctx_decompress(d.ctypes.data_as(u8p), d.size, s.ctypes.data_as(u8p),
                                            sl.ctypes.data_as(i32p), num_reads, out.ctypes.data_as(u8p),
                                            ql.ctypes.data_as(i32p))
    if rc != 0:
        raise QualityCtxError(rc, "decode")
    res, o = [], 0
    for n in ql.tolist():
        res.append(out[o:o + n].tobytes())
        o += n
    return res


def wrap_as_multiblock(blob: bytes) -> bytes:
    """quality_ctx.rs:605-611."""
    return struct.pack("<II", 1, len(blob)) + bytes(blob)


def decompress_quality_ctx_multiblock(data, sequences):
    """quality_ctx.rs:512-603: [num_blocks u32][len u32, blob]..., each blob
    taking its read count from its own header, sequences consumed in order."""
    if len(sequences) == 0:
        return []
    if len(data) == 0:
        raise QualityCtxError(-6, "multiblock: stream is empty but sequences are present")
    data = bytes(data)
    if len(data) < 4:
        raise QualityCtxError(-6, "multiblock: truncated block count")
    nb = struct.unpack_from("<I", data, 0)[0]
    if nb == 0:
        raise QualityCtxError(-6, f"multiblock: 0 blocks but {len(sequences)} sequences")
    off, cur, parts = 4, 0, []
    for b in range(nb):
        if off + 4 > len(data):
            raise QualityCtxError(-6, f"multiblock: truncated block {b} length")
        ln = struct.unpack_from("<I", data, off)[0]
        off += 4
        if off + ln > len(data):
            raise QualityCtxError(-6, f"multiblock: truncated block {b} data")
        blob = data[off:off + ln]
        off += ln
        if not blob:
            raise QualityCtxError(-6, f"multiblock: block {b} is empty")
        nro = 1 + blob[0] + 2
        if nro + 4 > len(blob):
            raise QualityCtxError(-6, f"multiblock: block {b} header truncated")
        nr = struct.unpack_from("<I", blob, nro)[0]
        if cur + nr > len(sequences):
            raise QualityCtxError(-6, f"multiblock: block {b} wants {nr} reads starting at {cur}")
        parts.append((blob, cur, nr))
        cur += nr
    if cur != len(sequences):
        raise QualityCtxError(-6, f"multiblock: blocks cover {cur} reads but {len(sequences)} sequences provided")
    # the blobs decode independently, in parallel (as the reference's second
    # phase does); the native decoder runs without the interpreter lock
    from concurrent.futures import ThreadPoolExecutor
    import os
    threads = max(1, min(len(parts), os.cpu_count() or 1))
    with ThreadPoolExecutor(threads) as ex:
        decoded = list(ex.map(lambda p: decompress_qualities_ctx(p[0], sequences[p[1]:p[1] + p[2]], p[2]), parts))
    out = []
    for d in decoded:
        out.extend(d)
    return out
