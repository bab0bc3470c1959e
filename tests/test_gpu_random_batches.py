"""Seeded random batches against the oracle: blocks of random size from
generators that reach every path of the forward sort (alphabets of 1..256
symbols on both sides of the 4- and 6-bit char-mode limits, run-heavy and
periodic texts, slices of the read, name and quality streams), batched
large enough to be split by alphabet class, once more with 1 MiB chunks of
the sort workspace. Every block's BWT, primary and sampled indexes equal
the oracle's, and the inverse gives every block back."""
import numpy as np
import pytest

from oracle import oracle as O
from qz_amd import bwt, synth

pytestmark = pytest.mark.gpu


def _blocks(seed: int, count: int, hi: int = 900_000):
    rng = np.random.default_rng(seed)
    seq = synth.sequence_stream(1_000_000, 20_000, 150, seed, 0.002, 0.0005)
    names = synth.header_stream(40_000, seed)
    quals = synth.quality_stream(20_000, 150, seed)
    out, kinds = [], []
    for _ in range(count):
        n = int(np.exp(rng.uniform(np.log(1), np.log(hi))))
        kind = int(rng.integers(0, 7))
        if kind == 0:  # uniform over sigma symbols
            sigma = int(rng.choice([1, 2, 3, 4, 5, 15, 16, 31, 32, 63, 64, 200, 256]))
            alpha = rng.choice(256, size=sigma, replace=False).astype(np.uint8)
            b = alpha[rng.integers(0, sigma, size=n)]
        elif kind == 1:  # runs of geometric length
            sigma = int(rng.integers(2, 9))
            alpha = rng.choice(256, size=sigma, replace=False).astype(np.uint8)
            lens = rng.geometric(0.05, size=n // 10 + 2)
            b = np.repeat(alpha[rng.integers(0, sigma, size=lens.size)], lens)[:n]
        elif kind == 2:  # periodic
            per = int(rng.integers(1, 300))
            b = np.resize(rng.integers(0, 256, size=per).astype(np.uint8), n)
        else:
            src = (seq, names, quals, seq)[kind - 3]
            o = int(rng.integers(0, max(1, src.size - n)))
            b = src[o:o + n]
        b = np.ascontiguousarray(b, dtype=np.uint8)
        if b.size == 0:
            b = np.array([65], np.uint8)
        out.append(b)
        kinds.append(kind)
    return out, kinds


def _check(blocks, kinds, label):
    out, prim, aux = bwt.encode_batch(blocks)
    for i, blk in enumerate(blocks):
        u, p, a = O.bwt_aux(blk)
        assert np.array_equal(out[i], u) and prim[i] == p and np.array_equal(aux[i], a), \
            f"{label}: block {i} kind {kinds[i]} size {blk.size}"
    back = bwt.decode_batch(out, prim)
    for i, blk in enumerate(blocks):
        assert np.array_equal(back[i], blk), f"{label}: round trip {i}"


@pytest.mark.parametrize("seed", [101, 102])
def test_random_batches_match_oracle(gpu, seed):
    blocks, kinds = _blocks(seed, 60)
    print(f"seed {seed}: {len(blocks)} blocks, {sum(b.size for b in blocks)} B", flush=True)
    _check(blocks, kinds, f"seed {seed}")


def test_random_large_batch_splits_by_class(gpu, capfd, monkeypatch):
    blocks, kinds = _blocks(104, 160, hi=2_500_000)
    print(f"seed 104: {len(blocks)} blocks, {sum(b.size for b in blocks)} B", flush=True)
    monkeypatch.setenv("QZBSC_TRACE", "1")
    out, prim, aux = bwt.encode_batch(blocks)
    err = capfd.readouterr().err
    monkeypatch.delenv("QZBSC_TRACE")
    assert err.count("qzbsc: sub-batch of") >= 2, err[-2000:]
    _check(blocks, kinds, "seed 104")


def test_random_batch_with_tiny_chunks(gpu, monkeypatch):
    blocks, kinds = _blocks(103, 60)
    print(f"seed 103: {len(blocks)} blocks, {sum(b.size for b in blocks)} B, 1 MiB chunks", flush=True)
    monkeypatch.setenv("QZBSC_CHUNK_MIB", "1")
    gpu.lib().qzbsc_trim()
    try:
        _check(blocks, kinds, "seed 103, 1 MiB chunks")
    finally:
        monkeypatch.delenv("QZBSC_CHUNK_MIB")
        gpu.lib().qzbsc_trim()
