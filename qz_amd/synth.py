"""Seeded synthetic streams shaped like qz's BSC inputs.

qz cuts three FASTQ streams into 25 MiB BSC blocks
(crates/qz-lib/src/compression/mod.rs:388-413):
  * sequences: concatenated reads, no separators, for constant read length
    (mod.rs:297-299);
  * headers: varint(len) + name per read (codecs.rs:97-105);
  * qualities: Phred+33 bytes (mod.rs:312-320; fed unpacked here).

The generators are vectorised with numpy so a full 1.5 GB stream can be made
in seconds; each takes an explicit seed.
"""
from __future__ import annotations

import numpy as np

BLOCK = 25 * 1024 * 1024  # qz's bsc_block_size_mb default (cli.rs:101-102,133)

_ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)
_COMP = np.zeros(256, dtype=np.uint8)
for a, b in zip(b"ACGTN", b"TGCAN"):
    _COMP[a] = b


def genome(length: int, seed: int, gc: float = 0.41) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p_at, p_cg = (1 - gc) / 2, gc / 2
    idx = rng.choice(4, size=length, p=[p_at, p_cg, p_cg, p_at])
    return _ACGT[idx]


def reads(g: np.ndarray, n_reads: int, read_len: int, seed: int, sub_rate: float,
          n_rate: float = 0.0, chunk: int = 1 << 20) -> np.ndarray:
    """Concatenated reads from uniform positions on both strands."""
    rng = np.random.default_rng(seed)
    out = np.empty(n_reads * read_len, dtype=np.uint8)
    win = np.lib.stride_tricks.sliding_window_view(g, read_len)
    code_of = np.zeros(256, dtype=np.uint8)
    code_of[_ACGT] = np.arange(4, dtype=np.uint8)
    for s in range(0, n_reads, chunk):
        k = min(chunk, n_reads - s)
        starts = rng.integers(0, g.size - read_len + 1, size=k)
        r = win[starts]
        rc = np.flatnonzero(rng.random(k) < 0.5)
        r[rc] = _COMP[r[rc, ::-1]]
        flat = r.reshape(-1)
        if sub_rate > 0:
            pos = rng.integers(0, flat.size, size=rng.binomial(flat.size, sub_rate))
            code = (code_of[flat[pos]] + rng.integers(1, 4, size=pos.size)) % 4
            flat[pos] = _ACGT[code]
        if n_rate > 0:
            flat[rng.integers(0, flat.size, size=rng.binomial(flat.size, n_rate))] = ord("N")
        out[s * read_len:(s + k) * read_len] = flat
    return out


def sequence_stream(genome_len: int, n_reads: int, read_len: int, seed: int, sub_rate: float,
                    n_rate: float = 0.0) -> np.ndarray:
    return reads(genome(genome_len, seed), n_reads, read_len, seed + 1, sub_rate, n_rate)


def header_stream(n_reads: int, seed: int, reads_per_tile: int = 4000) -> np.ndarray:
    """varint(len) + "@INSTR:RUN:FLOWCELL:LANE:TILE:X:Y" per read."""
    rng = np.random.default_rng(seed)
    parts = []
    chunk = 200_000
    for s in range(0, n_reads, chunk):
        k = min(chunk, n_reads - s)
        i = np.arange(s, s + k)
        tile = 1101 + (i // reads_per_tile) % 2000
        lane = 1 + (i // (reads_per_tile * 2000)) % 4
        x = rng.integers(0, 30000, size=k)
        y = rng.integers(0, 30000, size=k)
        names = [f"@A00123:456:HGK7WDSXY:{l}:{t}:{a}:{b}".encode()
                 for l, t, a, b in zip(lane.tolist(), tile.tolist(), x.tolist(), y.tolist())]
        buf = bytearray()
        for nm in names:
            buf.append(len(nm))  # names are < 128 bytes: one varint byte
            buf += nm
        parts.append(np.frombuffer(bytes(buf), dtype=np.uint8))
    return np.concatenate(parts) if parts else np.zeros(0, dtype=np.uint8)


_QV = np.array([2, 12, 23, 37], dtype=np.uint8) + 33


def quality_stream(n_reads: int, read_len: int, seed: int, p_repeat: float = 0.85,
                   chunk: int = 1 << 20) -> np.ndarray:
    """NovaSeq-binned Phred+33 first-order Markov qualities."""
    rng = np.random.default_rng(seed)
    head = np.array([0.01, 0.04, 0.15, 0.80])
    tail = np.array([0.05, 0.45, 0.30, 0.20])
    prof = np.empty((read_len, 4))
    for pos in range(read_len):
        t = max(0, pos - (read_len - 30)) / 30.0
        prof[pos] = (1 - t) * head + t * tail
    cdf = np.cumsum(prof, axis=1)
    out = np.empty(n_reads * read_len, dtype=np.uint8)
    for s in range(0, n_reads, chunk):
        k = min(chunk, n_reads - s)
        q = np.empty((k, read_len), dtype=np.uint8)
        cur = (rng.random(k)[:, None] > cdf[0][None, :]).sum(1)
        q[:, 0] = cur
        for pos in range(1, read_len):
            fresh = (rng.random(k)[:, None] > cdf[pos][None, :]).sum(1)
            rep = rng.random(k) < p_repeat
            cur = np.where(rep, cur, fresh)
            q[:, pos] = cur
        out[s * read_len:(s + k) * read_len] = _QV[np.minimum(q, 3)].reshape(-1)
    return out


def config1_block(seed: int = 1) -> np.ndarray:
    """1 Mbp genome, 100K x 150 bp reads, 0.1% substitutions: one 15 MB block."""
    return sequence_stream(1_000_000, 100_000, 150, seed, 0.001)


def config2_sequences(n_reads: int, seed: int = 2) -> np.ndarray:
    """50 Mbp genome, 150 bp reads, 0.2% substitutions, 0.01% N."""
    return sequence_stream(50_000_000, n_reads, 150, seed, 0.002, 0.0001)


def cut_blocks(stream: np.ndarray, block: int = BLOCK):
    return [stream[i:i + block] for i in range(0, stream.size, block)]


def mixed_blocks(count: int, seed: int, lo: int = 100_000, hi: int = BLOCK, total_cap: int | None = None):
    """Blocks of log-uniform size in [lo, hi] from a seeded mix of the three
    generators (sequence / header / quality)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)
    kinds = rng.integers(0, 3, size=count)
    if total_cap is not None:
        keep = np.cumsum(sizes) <= total_cap
        sizes, kinds = sizes[keep], kinds[keep]
    need = [int(sizes[kinds == k].sum()) for k in range(3)]
    pools = [
        config2_sequences(need[0] // 150 + 2, seed + 11) if need[0] else np.zeros(0, np.uint8),
        header_stream(need[1] // 38 + 2, seed + 12) if need[1] else np.zeros(0, np.uint8),
        quality_stream(need[2] // 150 + 2, 150, seed + 13) if need[2] else np.zeros(0, np.uint8),
    ]
    pos = [0, 0, 0]
    blocks = []
    for sz, k in zip(sizes.tolist(), kinds.tolist()):
        blocks.append(pools[k][pos[k]:pos[k] + sz])
        pos[k] += sz
    return blocks, kinds
