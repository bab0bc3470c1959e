"""quality_ctx model pins derived by hand (crates/qz-lib/src/compression/
quality_ctx.rs): each test lists the (cum, freq, total) triple that every
quality symbol is coded with, worked out on paper from the reference's
context rule (context_id, quality_ctx.rs:241-253), Laplace start
(alloc, :187-198), +1 update (:219-228) and rescale (:230-240). The expected
blob is those triples through the range coder, which the reference's own
coder cases and the two one- and two-symbol blobs of test_quality_ctx.py pin.

Each triple list is built so that a plausible slip in the model changes at
least one triple, as noted beside the entry that catches it."""
import struct

import numpy as np

from oracle import quality_ctx as QO
from qz_amd import quality_ctx as Q


def _blob(symbols, read_len, num_reads, triples):
    enc = QO.RangeEncoder()
    for cum, freq, total in triples:
        enc.encode(cum, freq, total)
    data = enc.finish()
    return bytes([len(symbols)]) + bytes(symbols) + struct.pack("<HII", read_len, num_reads, len(data)) + data


# Two reads, alphabet {D = phred 35 -> symbol 0, F = phred 37 -> symbol 1}.
# A fresh context codes with cum [0, 1, 2] (total 2). Context = (pos // 5,
# prev_q, |prev_q - prev_q2| <= 2 ? 0 : 1, prev_base * 5 + cur_base), with
# prev_q = prev_q2 = 0 and prev_base = 4 at each read's start, and A=0 C=1
# G=2 T=3 (either case), anything else 4.
#
# read 1: seq AAAAAAA  qual F F D F F F F
#   p0 c1 = (0, 0, 0, 20)  fresh        F -> (1,1,2)   c1 = [0,1,3]
#   p1 c2 = (0, 37, 1, 0)  fresh        F -> (1,1,2)   c2 = [0,1,3]
#   p2 c3 = (0, 37, 0, 0)  fresh        D -> (0,1,2)   c3 = [0,2,3]
#   p3 c4 = (0, 35, 0, 0)  fresh        F -> (1,1,2)   c4 = [0,1,3]
#   p4 c3 (|37-35| = 2 is stable)       F -> (2,1,3)   c3 = [0,2,4]
#        (prev_q2 as the quality, or "< 2" as stable, would read c4 / c2:
#         (1,2,3))
#   p5 c5 = (1, 37, 0, 0)  fresh        F -> (1,1,2)   c5 = [0,1,3]
#        (a position bin other than 5 // 5 = 1 would read c3: (2,2,4))
#   p6 c5                               F -> (1,2,3)   c5 = [0,1,4]
# read 2: seq AAAAC  qual F F D F F
#   p0 c1 (start state reset)           F -> (1,2,3)   c1 = [0,1,4]
#   p1 c2                               F -> (1,2,3)   c2 = [0,1,4]
#   p2 c3                               D -> (0,2,4)   c3 = [0,3,5]
#   p3 c4                               F -> (1,2,3)   c4 = [0,1,4]
#   p4 c6 = (0, 37, 0, 0*5 + 1) fresh   F -> (1,1,2)
#        (dropping the base from the context would read c3: (3,2,5))
CONTEXT_TRIPLES = [(1, 1, 2), (1, 1, 2), (0, 1, 2), (1, 1, 2), (2, 1, 3), (1, 1, 2), (1, 2, 3),
                   (1, 2, 3), (1, 2, 3), (0, 2, 4), (1, 2, 3), (1, 1, 2)]


def test_hand_derived_context_switches():
    quals = [b"FFDFFFF", b"FFDFF"]
    seqs = [b"AAAAAAA", b"AAAAC"]
    want = _blob([35, 37], 0, 2, CONTEXT_TRIPLES)  # read_len 0: reads differ in length
    blob = Q.compress_qualities_ctx(quals, seqs)
    assert blob == want
    assert Q.decompress_qualities_ctx(blob, seqs, 2) == quals


def _one_context_triples(is_d, quirk=True):
    """Every symbol in ONE context (reads of one quality, sequence "A"):
    counts a (D) and b (F) start at 1; D codes (0, a, a+b), F codes
    (a, b, a+b); after each +1 the total is checked against 2^20. The
    rescale walks the cumulative array in place, so the second frequency is
    taken against the already halved first one:
        a' = max(a >> 1, 1);  b' = max(((a + b) - a') >> 1, 1).
    quirk=False gives the plain halving b' = max(b >> 1, 1) instead."""
    a, b = 1, 1
    out = []
    rescales = 0
    for d in is_d:
        if d:
            out.append((0, a, a + b))
            a += 1
        else:
            out.append((a, b, a + b))
            b += 1
        if a + b >= 1 << 20:
            a2 = max(a >> 1, 1)
            b = max(((a + b) - a2) >> 1, 1) if quirk else max(b >> 1, 1)
            a = a2
            rescales += 1
    return out, rescales


def test_hand_derived_rescale():
    n = 1_100_000
    rng = np.random.default_rng(21)
    is_d = (rng.random(n) < 0.3).tolist()
    triples, rescales = _one_context_triples(is_d)
    assert rescales == 1
    quals = [b"D" if d else b"F" for d in is_d]
    seqs = [b"A"] * n
    blob = Q.compress_qualities_ctx(quals, seqs)
    assert blob == _blob([35, 37], 1, n, triples)
    # a plain halving of the second count gives other bytes
    plain, _ = _one_context_triples(is_d, quirk=False)
    assert _blob([35, 37], 1, n, plain) != blob
    assert Q.decompress_qualities_ctx(blob, seqs, n) == quals
