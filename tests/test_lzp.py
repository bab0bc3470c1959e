"""LZP stage (host, libqzbsc.so): round trips over the parameter range qz
uses and the edges of the token format, one hand-derived token vector,
parameter checks, and corrupt streams failing instead of decoding wrong.

The stage follows the LZP scheme libbsc applies before the BWT; libbsc's own
byte format is not in the reference tree and no fixture holds an LZP block,
so byte parity with libbsc is unpinned. What is pinned here is this
library's format, stated in qz_amd/csrc/lzp.cpp."""
import numpy as np
import pytest

from qz_amd import bsc, synth
from qz_amd.bwt import BscError


def _rt(data, hs, ml):
    try:
        z = bsc.lzp_compress(data, hs, ml)
    except BscError as e:
        assert e.code == -3
        return None
    assert len(z) < len(data)
    assert bsc.lzp_decompress(z, len(data), hs, ml) == bytes(data)
    return z


def test_roundtrip_parameter_range():
    rng = np.random.default_rng(5)
    hdr = synth.header_stream(20_000, 8).tobytes()
    cases = [hdr, hdr * 3, b"abcdabcdabcdabcd" * 50, b"\x07" * 10_000, bytes(range(256)) * 40,
             (bytes(range(256)) + b"Q" * 300) * 8,       # every byte value present: literal escapes occur
             rng.integers(0, 4, 50_000).astype(np.uint8).tobytes(),
             synth.quality_stream(2_000, 150, 9).tobytes()]
    shrunk = 0
    for data in cases:
        for hs, ml in ((16, 128), (10, 4), (28, 255), (12, 32), (20, 5)):
            shrunk += _rt(data, hs, ml) is not None
    assert shrunk >= 20


def test_incompressible_and_tiny_inputs():
    rng = np.random.default_rng(6)
    for data in (rng.integers(0, 256, 20_000).astype(np.uint8).tobytes(), b"xyz", b"", b"abcd"):
        with pytest.raises(BscError) as e:
            bsc.lzp_compress(data, 16, 128)
        assert e.value.code == -3


def test_hand_derived_tokens():
    # 16 bytes, four of each of a, b, c, d: byte 0 is the rarest (escape).
    # Positions 0..7 are literals (the table first sees "abcd" at 4); at 8
    # the context "abcd" predicts position 4 and the next 8 bytes match:
    # escape, varint(8 - 4 + 1).
    z = bsc.lzp_compress(b"abcdabcdabcdabcd", 10, 4)
    assert z == b"\x00abcdabcd\x00\x05"
    # a literal equal to the escape byte is sent as escape, 0
    data = bytes(range(1, 256)) + b"\x00" + b"\x01" * 300
    z = bsc.lzp_compress(data, 10, 4)
    esc = z[0]
    assert esc != 1 and bytes([esc, 0]) in z
    assert bsc.lzp_decompress(z, len(data), 10, 4) == data


def test_parameter_checks():
    for hs, ml in ((9, 128), (29, 128), (16, 3), (16, 256)):
        with pytest.raises(BscError) as e:
            bsc.lzp_compress(b"abcd" * 100, hs, ml)
        assert e.value.code == -1


def test_corrupt_streams_fail():
    data = (synth.header_stream(5_000, 10).tobytes() * 4)
    z = bytearray(bsc.lzp_compress(data, 12, 8))
    with pytest.raises(BscError):
        bsc.lzp_decompress(bytes(z[:len(z) // 2]), len(data), 12, 8)     # truncated
    with pytest.raises(BscError):
        bsc.lzp_decompress(bytes(z) + b"\x00", len(data), 12, 8)          # trailing bytes
    # a match token before any prediction exists
    with pytest.raises(BscError):
        bsc.lzp_decompress(b"\x00ab\x00\x05", 10, 10, 4)
    # an unterminated length
    with pytest.raises(BscError):
        bsc.lzp_decompress(b"\x00abcdabcd\x00\xff\xff\xff\xff\xff\xff", 16, 10, 4)
