"""Python mirror of qz-lib's BSC wrapper (crates/qz-lib/src/compression/bsc.rs)
over this library's libbsc-compatible C-ABI.

Same names, argument meaning and error behaviour: empty input gives empty
output (bsc.rs:162-164), inputs above i32::MAX are rejected (bsc.rs:165-167),
the output buffer is n + 28 + 1024 bytes (bsc.rs:170), decompression asks
bsc_block_info for the data size first (bsc.rs:401-419), and any negative
libbsc code raises.
"""
from __future__ import annotations

import struct

import numpy as np

from ._native import lib, u8p, i32p
from .bwt import BscError

LIBBSC_NO_ERROR = 0
LIBBSC_BAD_PARAMETER = -1
LIBBSC_NOT_ENOUGH_MEMORY = -2
LIBBSC_NOT_COMPRESSIBLE = -3
LIBBSC_NOT_SUPPORTED = -4
LIBBSC_UNEXPECTED_EOB = -5
LIBBSC_DATA_CORRUPT = -6
QZBSC_GPU_ERROR = -7

LIBBSC_BLOCKSORTER_BWT = 1
LIBBSC_CODER_QLFC_STATIC = 1
LIBBSC_CODER_QLFC_ADAPTIVE = 2
LIBBSC_CODER_QLFC_FAST = 3
# the coder id this library writes into a block's mode word (its own model)
QZBSC_CODER_QZ_ADAPTIVE = 4
LIBBSC_FEATURE_FASTMODE = 1
LIBBSC_FEATURE_MULTITHREADING = 2
LIBBSC_FEATURE_CUDA = 8
LIBBSC_HEADER_SIZE = 28
BSC_BLOCK_SIZE = 25 * 1024 * 1024  # bsc.rs:89

_I32_MAX = 2**31 - 1
_inited = False


def _ensure_initialized():
    """bsc.rs:128-141: one bsc_init, a nonzero result only warns."""
    global _inited
    if not _inited:
        rc = lib().bsc_init(LIBBSC_FEATURE_FASTMODE | LIBBSC_FEATURE_MULTITHREADING | LIBBSC_FEATURE_CUDA)
        if rc != LIBBSC_NO_ERROR:
            import sys
            print(f"Warning: BSC initialization failed with code {rc}", file=sys.stderr)
        _inited = True


def _arr(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8)
    return np.frombuffer(bytes(data), dtype=np.uint8)


def compress_with_params(data, lzp_hash_size: int, lzp_min_len: int, block_sorter: int, coder: int,
                         features: int = LIBBSC_FEATURE_FASTMODE | LIBBSC_FEATURE_CUDA) -> bytes:
    """bsc.rs:153-191."""
    _ensure_initialized()
    a = _arr(data)
    if a.size == 0:
        return b""
    if a.size > _I32_MAX:
        raise BscError(LIBBSC_BAD_PARAMETER, f"BSC input too large: {a.size} bytes (max {_I32_MAX})")
    out = np.zeros(a.size + LIBBSC_HEADER_SIZE + 1024, dtype=np.uint8)
    r = lib().bsc_compress(a.ctypes.data_as(u8p), out.ctypes.data_as(u8p), a.size, lzp_hash_size, lzp_min_len,
                           block_sorter, coder, features)
    if r < 0:
        raise BscError(r, "BSC compression")
    return out[:r].tobytes()


def compress(data) -> bytes:
    """bsc.rs:146-150: the static-QLFC request, served by this library's
    adaptive model (the block header records QZBSC_CODER_QZ_ADAPTIVE = 4)."""
    return compress_with_params(data, 0, 0, LIBBSC_BLOCKSORTER_BWT, LIBBSC_CODER_QLFC_STATIC)


def compress_adaptive(data) -> bytes:
    """bsc.rs:194-196: LZP 16/128 + adaptive QLFC."""
    return compress_with_params(data, 16, 128, LIBBSC_BLOCKSORTER_BWT, LIBBSC_CODER_QLFC_ADAPTIVE)


def compress_adaptive_mt(data) -> bytes:
    """bsc.rs:212-215,224-255: LZP 16/128, segment coding on several threads."""
    return compress_with_params(data, 16, 128, LIBBSC_BLOCKSORTER_BWT, LIBBSC_CODER_QLFC_ADAPTIVE,
                                LIBBSC_FEATURE_FASTMODE | LIBBSC_FEATURE_MULTITHREADING | LIBBSC_FEATURE_CUDA)


def compress_adaptive_no_lzp(data) -> bytes:
    """bsc.rs:201-203: the default sequence/quality/header block coder."""
    return compress_with_params(data, 0, 0, LIBBSC_BLOCKSORTER_BWT, LIBBSC_CODER_QLFC_ADAPTIVE)


def compress_adaptive_mt_no_lzp(data) -> bytes:
    """bsc.rs:218-220,224-255 without LZP: segment coding on several threads."""
    return compress_with_params(data, 0, 0, LIBBSC_BLOCKSORTER_BWT, LIBBSC_CODER_QLFC_ADAPTIVE,
                                LIBBSC_FEATURE_FASTMODE | LIBBSC_FEATURE_MULTITHREADING | LIBBSC_FEATURE_CUDA)


def compress_batch(blocks, lzp_hash_size: int = 0, lzp_min_len: int = 0,
                   block_sorter: int = LIBBSC_BLOCKSORTER_BWT, coder: int = LIBBSC_CODER_QLFC_ADAPTIVE,
                   features: int = LIBBSC_FEATURE_FASTMODE | LIBBSC_FEATURE_CUDA) -> list:
    """Many blocks through one qzbsc_compress_batch call (device
    sub-batches, host coder on all segments at once); the same bytes per
    block as compress_with_params. Replaces the per-block par_iter of
    compress_stream_to_bsc_blocks (mod.rs:388-413)."""
    arrs = [_arr(b) for b in blocks]
    if not arrs:
        return []
    inp = np.concatenate(arrs) if len(arrs) > 1 else arrs[0]
    return compress_batch_buffer(inp, [a.size for a in arrs], lzp_hash_size, lzp_min_len, block_sorter, coder,
                                 features)


def compress_batch_buffer(buf: np.ndarray, sizes, lzp_hash_size: int = 0, lzp_min_len: int = 0,
                          block_sorter: int = LIBBSC_BLOCKSORTER_BWT, coder: int = LIBBSC_CODER_QLFC_ADAPTIVE,
                          features: int = LIBBSC_FEATURE_FASTMODE | LIBBSC_FEATURE_CUDA) -> list:
    """compress_batch over consecutive blocks of one uint8 buffer (block k
    is the sizes[k] bytes after the first k), without gathering them."""
    _ensure_initialized()
    if not len(sizes):
        return []
    for n in sizes:
        if n > _I32_MAX:
            raise BscError(LIBBSC_BAD_PARAMETER, f"BSC: input too large ({n} bytes, max {_I32_MAX})")
    sizes = np.asarray(sizes, dtype=np.int32)
    buf = np.ascontiguousarray(buf, dtype=np.uint8)
    if int(sizes.astype(np.int64).sum()) > buf.size:
        raise BscError(LIBBSC_BAD_PARAMETER, "BSC: block sizes exceed the buffer")
    caps = sizes.astype(np.int64) + LIBBSC_HEADER_SIZE + 1024
    out = np.empty(int(caps.sum()), dtype=np.uint8)
    osz = np.zeros(len(sizes), dtype=np.int32)
    rc = lib().qzbsc_compress_batch(buf.ctypes.data_as(u8p), sizes.ctypes.data_as(i32p), len(sizes),
                                    out.ctypes.data_as(u8p), osz.ctypes.data_as(i32p), lzp_hash_size, lzp_min_len,
                                    block_sorter, coder, features)
    if rc < 0:
        raise BscError(rc, "BSC batch compression")
    res, off = [], 0
    for k in range(len(sizes)):
        res.append(out[off:off + int(osz[k])].tobytes())
        off += int(caps[k])
    return res


def decompress_batch(blocks, features: int = LIBBSC_FEATURE_FASTMODE | LIBBSC_FEATURE_CUDA) -> list:
    """Many blocks through one qzbsc_decompress_batch call: host coder
    decode of all segments in parallel, device sub-batches of inverse BWTs.
    Each block is checked as bsc.rs:384-436 checks one (size, header, all
    three checksums). Replaces the batches of 8 per-block decompress calls
    of decompress_impl.rs:97-121."""
    out, dsz = decompress_batch_buffer(blocks, features)
    res, off = [], 0
    for n in dsz:
        res.append(out[off:off + n].tobytes())
        off += n
    return res


def decompress_batch_buffer(blocks, features: int = LIBBSC_FEATURE_FASTMODE | LIBBSC_FEATURE_CUDA):
    """decompress_batch with the decoded blocks laid end to end in one
    uint8 array: (array, list of block sizes)."""
    _ensure_initialized()
    arrs = [_arr(b) for b in blocks]
    if not arrs:
        return np.zeros(0, np.uint8), []
    dsz = np.zeros(len(arrs), dtype=np.int32)
    for k, a in enumerate(arrs):
        if a.size < LIBBSC_HEADER_SIZE:
            raise BscError(LIBBSC_UNEXPECTED_EOB,
                           f"BSC: compressed data too small (need at least {LIBBSC_HEADER_SIZE} bytes for header)")
        bs = np.zeros(1, np.int32)
        ds = np.zeros(1, np.int32)
        rc = lib().bsc_block_info(a.ctypes.data_as(u8p), a.size, bs.ctypes.data_as(i32p), ds.ctypes.data_as(i32p),
                                  features)
        if rc != LIBBSC_NO_ERROR:
            raise BscError(rc, "BSC block_info")
        if ds[0] <= 0:
            raise BscError(LIBBSC_DATA_CORRUPT, f"BSC: invalid decompressed size: {int(ds[0])}")
        dsz[k] = ds[0]
    isz = np.array([a.size for a in arrs], dtype=np.int32)
    inp = np.concatenate(arrs) if len(arrs) > 1 else arrs[0]
    out = np.empty(int(dsz.astype(np.int64).sum()), dtype=np.uint8)
    rcs = np.zeros(len(arrs), dtype=np.int32)
    rc = lib().qzbsc_decompress_batch(inp.ctypes.data_as(u8p), isz.ctypes.data_as(i32p), len(arrs),
                                      out.ctypes.data_as(u8p), dsz.ctypes.data_as(i32p), features,
                                      rcs.ctypes.data_as(i32p))
    if rc != LIBBSC_NO_ERROR:
        raise BscError(rc, "BSC batch decompression")
    return out, [int(n) for n in dsz]


def _compress_parallel_with(data, block_compressor) -> bytes:
    """bsc.rs:257-301: [num_blocks u32][len u32, block]... even for one block."""
    a = _arr(data)
    if a.size == 0:
        return b""
    blocks = [a[i:i + BSC_BLOCK_SIZE] for i in range(0, a.size, BSC_BLOCK_SIZE)]
    parts = [struct.pack("<I", len(blocks))]
    for b in blocks:
        c = block_compressor(b)
        parts.append(struct.pack("<I", len(c)))
        parts.append(c)
    return b"".join(parts)


def compress_parallel_adaptive_no_lzp(data) -> bytes:
    return _compress_parallel_with(data, compress_adaptive_no_lzp)


def compress_parallel_adaptive(data) -> bytes:
    """bsc.rs:206-210: the columnar header path (LZP 16/128)."""
    return _compress_parallel_with(data, compress_adaptive)


def lzp_compress(data, hash_size: int = 16, min_len: int = 128) -> bytes:
    """The LZP stage alone (libbsc lzp.h bsc_lzp_compress); raises -3 when the
    result would not be smaller than the input."""
    a = _arr(data)
    out = np.zeros(max(a.size, 1), dtype=np.uint8)
    r = lib().qzbsc_lzp_compress(a.ctypes.data_as(u8p), out.ctypes.data_as(u8p), a.size, hash_size, min_len)
    if r < 0:
        raise BscError(r, "LZP compression")
    return out[:r].tobytes()


def lzp_decompress(data, n: int, hash_size: int = 16, min_len: int = 128) -> bytes:
    a = _arr(data)
    out = np.zeros(max(n, 1), dtype=np.uint8)
    r = lib().qzbsc_lzp_decompress(a.ctypes.data_as(u8p), a.size, out.ctypes.data_as(u8p), n, hash_size, min_len)
    if r < 0:
        raise BscError(r, "LZP decompression")
    return out[:n].tobytes()


def decompress_with_features(data, features: int = LIBBSC_FEATURE_FASTMODE | LIBBSC_FEATURE_CUDA) -> bytes:
    """bsc.rs:384-436."""
    _ensure_initialized()
    a = _arr(data)
    if a.size == 0:
        return b""
    if a.size < LIBBSC_HEADER_SIZE:
        raise BscError(LIBBSC_UNEXPECTED_EOB,
                       f"BSC: compressed data too small (need at least {LIBBSC_HEADER_SIZE} bytes for header)")
    bs = np.zeros(1, np.int32)
    ds = np.zeros(1, np.int32)
    rc = lib().bsc_block_info(a.ctypes.data_as(u8p), a.size, bs.ctypes.data_as(i32p), ds.ctypes.data_as(i32p),
                              features)
    if rc != LIBBSC_NO_ERROR:
        raise BscError(rc, "BSC block_info")
    if ds[0] <= 0:
        raise BscError(LIBBSC_DATA_CORRUPT, f"BSC: invalid decompressed size: {int(ds[0])}")
    out = np.zeros(int(ds[0]), dtype=np.uint8)
    rc = lib().bsc_decompress(a.ctypes.data_as(u8p), a.size, out.ctypes.data_as(u8p), int(ds[0]), features)
    if rc != LIBBSC_NO_ERROR:
        raise BscError(rc, "BSC decompression")
    return out.tobytes()


def decompress(data) -> bytes:
    return decompress_with_features(data)


def decompress_mt(data) -> bytes:
    return decompress_with_features(data, LIBBSC_FEATURE_FASTMODE | LIBBSC_FEATURE_MULTITHREADING
                                    | LIBBSC_FEATURE_CUDA)


def decompress_parallel(data) -> bytes:
    """bsc.rs:324-371."""
    a = bytes(data)
    if not a:
        return b""
    if len(a) < 4:
        raise BscError(LIBBSC_UNEXPECTED_EOB, "BSC parallel: data too small for header")
    (nb,) = struct.unpack_from("<I", a, 0)
    off = 4
    blocks = []
    for _ in range(nb):
        if off + 4 > len(a):
            raise BscError(LIBBSC_UNEXPECTED_EOB, "BSC parallel: truncated block length")
        (ln,) = struct.unpack_from("<I", a, off)
        off += 4
        if off + ln > len(a):
            raise BscError(LIBBSC_UNEXPECTED_EOB, "BSC parallel: truncated block data")
        blocks.append(a[off:off + ln])
        off += ln
    return b"".join(decompress(b) for b in blocks)
