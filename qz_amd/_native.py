"""Loads qz_amd/libqzbsc.so (the HIP build) and declares its C-ABI.

There is no fallback: if the library is missing or cannot be loaded, every
entry point raises. Build it with `python -m qz_amd.build` (or
__graft_entry__.build()).
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libqzbsc.so"
HEADER = PKG.parent / "include" / "qzbsc.h"

u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)
f32p = C.POINTER(C.c_float)

_SIGS = {
    "bsc_init": (C.c_int, [C.c_int]),
    "bsc_compress": (C.c_int, [u8p, u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "bsc_block_info": (C.c_int, [u8p, C.c_int, i32p, i32p, C.c_int]),
    "bsc_decompress": (C.c_int, [u8p, C.c_int, u8p, C.c_int, C.c_int]),
    "qzbsc_compress_batch": (C.c_int, [u8p, i32p, C.c_int, u8p, i32p, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_int]),
    "qzbsc_decompress_batch": (C.c_int, [u8p, i32p, C.c_int, u8p, i32p, C.c_int, i32p]),
    "qzbsc_fastq_to_streams": (C.c_longlong, [u8p, C.c_longlong, u8p, u8p, u8p, C.POINTER(C.c_longlong)]),
    "qzbsc_streams_to_fastq": (C.c_longlong, [u8p, C.c_longlong, u8p, C.c_longlong, u8p, C.c_longlong,
                                              C.c_longlong, C.c_longlong, C.c_longlong, u8p, C.c_longlong]),
    "qzbsc_bwt_encode": (C.c_int, [u8p, C.c_int, u8p, i32p, C.c_int]),
    "qzbsc_bwt_decode": (C.c_int, [u8p, C.c_int, C.c_int, C.c_uint8, i32p, C.c_int]),
    "qzbsc_bwt_encode_batch": (C.c_int, [u8p, u8p, i32p, C.c_int, i32p, i32p]),
    "qzbsc_bwt_decode_batch": (C.c_int, [u8p, u8p, i32p, C.c_int, i32p]),
    "qzbsc_dev_bwt_encode_batch": (C.c_int, [C.c_void_p, C.c_void_p, i32p, C.c_int, i32p, i32p, f32p]),
    "qzbsc_dev_bwt_decode_batch": (C.c_int, [C.c_void_p, C.c_void_p, i32p, C.c_int, i32p, f32p]),
    "qzbsc_dev_bwt_decode_batch_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, i32p, C.c_int, i32p, C.c_void_p]),
    "qzbsc_set_max_engines": (C.c_int, [C.c_int]),
    "qzbsc_qlfc_front": (C.c_int, [u8p, C.c_int, i32p, i32p, i32p, i32p, u8p, i32p, u8p, u8p, i32p]),
    "qzbsc_segment_header": (C.c_int, [u8p, C.c_int, i32p, u8p, i32p]),
    "qzbsc_coder_compress": (C.c_int, [u8p, C.c_int, u8p, C.c_int, C.c_int]),
    "qzbsc_coder_decompress": (C.c_int, [u8p, C.c_int, u8p, C.c_int, C.c_int]),
    "qzbsc_gpu_available": (C.c_int, []),
    "qzbsc_server_create": (C.c_void_p, [C.c_longlong, C.c_longlong, C.c_int]),
    "qzbsc_server_create_lanes": (C.c_void_p, [C.c_longlong, C.c_longlong, C.c_int, C.c_longlong]),
    "qzbsc_server_create_opts": (C.c_void_p, [C.c_longlong, C.c_longlong, C.c_int, C.c_longlong, C.c_longlong,
                                              C.c_int]),
    "qzbsc_server_destroy": (None, [C.c_void_p]),
    "qzbsc_server_batch_cap": (C.c_longlong, [C.c_void_p]),
    "qzbsc_host_alloc": (C.c_void_p, [C.c_longlong]),
    "qzbsc_host_free": (None, [C.c_void_p]),
    "qzbsc_device_memory": (C.c_int, [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "qzbsc_server_submit": (C.c_int, [C.c_void_p, C.c_int, u8p, u8p, i32p, C.c_int, i32p, i32p, i32p]),
    "qzbsc_server_wait": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int)]),
    "qzbsc_server_peak_bytes": (C.c_longlong, [C.c_void_p]),
    "qzbsc_footprint": (C.c_longlong, [C.c_int, C.c_longlong, C.c_int]),
    "qzbsc_tuning": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "qzbsc_tune_knob": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_double)]),
    "qzbsc_graph_stats": (C.c_int, [C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)]),
    "qzbsc_trim": (C.c_longlong, []),
    "qzbsc_lzp_compress": (C.c_int, [u8p, u8p, C.c_int, C.c_int, C.c_int]),
    "qzbsc_lzp_decompress": (C.c_int, [u8p, C.c_int, u8p, C.c_int, C.c_int, C.c_int]),
    "qzbsc_quality_ctx_compress": (C.c_longlong, [u8p, i32p, u8p, i32p, C.c_int, u8p, C.c_longlong]),
    "qzbsc_quality_ctx_decompress": (C.c_int, [u8p, C.c_longlong, u8p, i32p, C.c_int, u8p, i32p]),
    "qzbsc_calib_access": (C.c_int, [C.c_int, C.c_longlong, C.c_longlong, C.c_int, C.POINTER(C.c_double)]),
}

_lib = None


class NativeMissing(RuntimeError):
    pass


def header_symbols() -> list[str]:
    """Every function declared in include/qzbsc.h."""
    txt = HEADER.read_text()
    return re.findall(r"QZBSC_API\s+[\w\s\*]+?\b(\w+)\s*\(", txt)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NativeMissing(f"{LIB_PATH} is not built; run `python -m qz_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _lib = L
    return _lib
