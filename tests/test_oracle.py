"""Pins the CPU oracle against independent checks (CPU only).

* adler32 against Python's zlib;
* suffix array / BWT / sampled indexes against brute-force suffix sorting;
* inverse BWT against the forward transform (round trip);
* QLFC front end against its direct characterisation and its inverse.
"""
import random
import zlib

import numpy as np
import pytest

from oracle import oracle as O


def brute_sa(t: bytes):
    return sorted(range(len(t)), key=lambda i: t[i:])


def brute_bwt(t: bytes):
    """libsais conventions: n+1 rows incl. the empty suffix; $ removed."""
    n = len(t)
    sa = brute_sa(t)
    full = [t[n - 1]]
    prim = None
    for k, s in enumerate(sa):
        if s == 0:
            prim = k + 1
            full.append(None)
        else:
            full.append(t[s - 1])
    u = bytes(c for c in full if c is not None)
    r = O.sample_rate(n)
    rank = {s: k for k, s in enumerate(sa)}
    aux = [rank[k * r] + 1 for k in range((n - 1) // r + 1)]
    return u, prim, aux


def corpus(seed=0):
    rng = random.Random(seed)
    out = [b"a", b"ab", b"ba", b"aa", b"\x00", b"\x00\x00", b"\x00\x01\x00", b"banana", b"mississippi",
           b"abracadabra" * 7, b"A" * 100, b"AC" * 77, b"ACG" * 50 + b"A", bytes(range(256)),
           bytes(range(255, -1, -1)) * 2]
    for _ in range(60):
        n = rng.randint(1, 400)
        sigma = rng.choice([1, 2, 3, 4, 5, 20, 256])
        out.append(bytes(rng.randrange(sigma) for _ in range(n)))
    # periodic with a perturbation (deep groups)
    for p in (1, 2, 3, 7):
        base = bytes(rng.randrange(4) + 65 for _ in range(p)) * (300 // p)
        b = bytearray(base)
        b[rng.randrange(len(b))] = 90
        out.append(bytes(b))
    return out


def test_adler32_matches_zlib():
    rng = np.random.default_rng(1)
    for n in (0, 1, 5551, 5552, 5553, 100_000):
        d = rng.integers(0, 256, size=n, dtype=np.uint8).tobytes()
        assert O.adler32(d) == zlib.adler32(d)
    assert O.adler32(b"\xff" * 1_000_000) == zlib.adler32(b"\xff" * 1_000_000)


def test_suffix_array_bruteforce():
    for t in corpus():
        assert O.suffix_array(t).tolist() == brute_sa(t), t[:40]


def test_bwt_conventions_bruteforce():
    for t in corpus(3):
        u, p, aux = O.bwt_aux(t)
        bu, bp, baux = brute_bwt(t)
        assert bytes(u) == bu and p == bp and aux.tolist() == baux, t[:40]


def test_sample_rate_rule():
    # largest power of two <= n/8, and 1 below 16
    for n in list(range(1, 300)) + [2_904_234, 4_450_000, 5_050_000, 15_000_000, 26_214_400]:
        r = O.sample_rate(n)
        want = 1
        while want * 2 <= n // 8:
            want *= 2
        assert r == want, n


def test_bsc_trailer_from_aux():
    t = bytes(random.Random(5).randrange(4) + 65 for _ in range(5000))
    u, p, idx = O.bsc_bwt_encode(t)
    u2, p2, aux = O.bwt_aux(t)
    assert bytes(u) == bytes(u2) and p == p2 == aux[0]
    assert idx.tolist() == [a - 1 for a in aux[1:].tolist()]
    assert len(idx) == (len(t) - 1) // O.sample_rate(len(t))


def test_unbwt_roundtrip():
    for t in corpus(7):
        u, p, _ = O.bwt_aux(t)
        assert bytes(O.unbwt(u, p)) == t


def test_unbwt_roundtrip_large():
    rng = np.random.default_rng(9)
    t = rng.choice(np.frombuffer(b"ACGT", np.uint8), size=300_000)
    u, p, _ = O.bwt_aux(t)
    assert np.array_equal(O.unbwt(u, p), t)


def _qlfc_direct(seg: bytes):
    """Rank of each run = #distinct symbols strictly between the run's end and
    the next occurrence of its symbol; #distinct symbols after the run when
    there is none."""
    runs = []
    i = 0
    while i < len(seg):
        j = i
        while j < len(seg) and seg[j] == seg[i]:
            j += 1
        runs.append((seg[i], i, j))
        i = j
    ranks = []
    for c, i, j in runs:
        nxt = seg.find(bytes([c]), j)
        window = seg[j:nxt] if nxt >= 0 else seg[j:]
        ranks.append(len(set(window)))
    first = []
    for c, _, _ in runs:
        if c not in first:
            first.append(c)
    return [c for c, _, _ in runs], ranks, [j - i for _, i, j in runs], first


def test_qlfc_front_end_direct():
    for t in corpus(11):
        sym, rank, ln, table = O.qlfc_transform(t)
        ds, dr, dl, dt = _qlfc_direct(t)
        assert sym.tolist() == ds and rank.tolist() == dr and ln.tolist() == dl and table.tolist() == dt


def test_qlfc_front_end_roundtrip():
    for t in corpus(13):
        sym, rank, ln, table = O.qlfc_transform(t)
        assert bytes(O.qlfc_inverse(table, rank, ln, len(t))) == t
