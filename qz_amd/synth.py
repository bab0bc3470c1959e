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


def _ascii_uint(v: np.ndarray, width: int):
    """Decimal digits of v (< 10**width) as a (len, width) uint8 array and the
    mask of digits to keep (no leading zeros)."""
    d = np.empty((v.size, width), dtype=np.uint8)
    x = v.astype(np.int64).copy()
    for j in range(width - 1, -1, -1):
        d[:, j] = 48 + (x % 10)
        x //= 10
    ndig = np.maximum(1, np.floor(np.log10(np.maximum(v, 1))).astype(np.int64) + 1)
    keep = np.arange(width)[None, :] >= (width - ndig)[:, None]
    return d, keep


_COMMENT = np.frombuffer(b" 1:N:0:", dtype=np.uint8)
_INDEX = np.frombuffer(b"GATCAGAT+AGATCTCG", dtype=np.uint8)
_INDEX_BASES = np.array([i for i in range(17) if i != 8])


def header_stream(n_reads: int, seed: int, reads_per_tile: int = 4000) -> np.ndarray:
    """varint(len) + "@A00123:456:HGK7WDSXY:LANE:TILE:X:Y 1:N:0:I7+I5" per
    read (about 64 bytes); tiles increment every `reads_per_tile` reads, x and
    y are uniform below 30000, and the dual sample index reads GATCAGAT +
    AGATCTCG except for one substituted base in 1% of reads."""
    rng = np.random.default_rng(seed)
    parts = []
    chunk = 1 << 20
    prefix = np.frombuffer(b"@A00123:456:HGK7WDSXY:", dtype=np.uint8)
    for s in range(0, n_reads, chunk):
        k = min(chunk, n_reads - s)
        i = np.arange(s, s + k)
        tile = 1101 + (i // reads_per_tile) % 2000
        lane = 1 + (i // (reads_per_tile * 2000)) % 4
        x = rng.integers(0, 30000, size=k)
        y = rng.integers(0, 30000, size=k)
        cols, keeps = [], []
        colon = np.full((k, 1), ord(":"), np.uint8)
        on = np.ones((k, 1), bool)
        cols.append(np.broadcast_to(prefix, (k, prefix.size)))
        keeps.append(np.ones((k, prefix.size), bool))
        for v, w in ((lane, 1), (tile, 4), (x, 5), (y, 5)):
            d, kp = _ascii_uint(v, w)
            cols += [d, colon]
            keeps += [kp, on]
        cols, keeps = cols[:-1], keeps[:-1]
        # Illumina read comment: read 1, not filtered, control 0, dual index
        idx = np.broadcast_to(_INDEX, (k, _INDEX.size)).copy()
        hit = np.flatnonzero(rng.random(k) < 0.01)
        pos = _INDEX_BASES[rng.integers(0, _INDEX_BASES.size, size=hit.size)]
        idx[hit, pos] = np.frombuffer(b"ACGT", dtype=np.uint8)[rng.integers(0, 4, size=hit.size)]
        cols += [np.broadcast_to(_COMMENT, (k, _COMMENT.size)), idx]
        keeps += [np.ones((k, _COMMENT.size), bool), np.ones((k, _INDEX.size), bool)]
        body = np.concatenate(cols, axis=1)
        keep = np.concatenate(keeps, axis=1)
        length = keep.sum(axis=1).astype(np.uint8)  # < 128: one varint byte
        rows = np.concatenate([length[:, None], body], axis=1)
        mask = np.concatenate([on, keep], axis=1)
        parts.append(rows[mask])
    return np.concatenate(parts) if parts else np.zeros(0, dtype=np.uint8)


_QV = np.array([2, 12, 23, 37], dtype=np.uint8) + 33


def quality_stream(n_reads: int, read_len: int, seed: int, p_repeat: float = 0.85,
                   chunk: int = 1 << 17) -> np.ndarray:
    """NovaSeq-binned Phred+33 qualities, one first-order Markov chain per
    read: with probability p_repeat the previous value repeats, otherwise a
    fresh value is drawn from a position profile that decays from Q37
    toward Q12 over the last 30 cycles."""
    rng = np.random.default_rng(seed)
    head = np.array([0.01, 0.04, 0.15, 0.80])
    tail = np.array([0.05, 0.45, 0.30, 0.20])
    t = np.clip(np.arange(read_len) - (read_len - 30), 0, None) / 30.0
    prof = (1 - t)[:, None] * head[None, :] + t[:, None] * tail[None, :]
    cdf = np.round(np.cumsum(prof, axis=1)[:, :3] * 65536).astype(np.uint32)  # (L, 3), 16-bit fixed point
    rep_t = int(round(p_repeat * 65536))
    out = np.empty((n_reads, read_len), dtype=np.uint8)
    cdf_t = np.ascontiguousarray(cdf.T[:, :, None])  # (3, L, 1)
    sub = 4096  # reads per cache-sized piece
    for s in range(0, n_reads, chunk):
        k = min(chunk, n_reads - s)
        u = rng.integers(0, 65536, size=(k, read_len), dtype=np.uint32)
        rep = rng.integers(0, 65536, size=(k, read_len), dtype=np.uint32) < rep_t
        for r0 in range(0, k, sub):
            ut = u[r0:r0 + sub].T  # cycle-major: the chain runs one cycle at a time over all reads
            rt = np.ascontiguousarray(rep[r0:r0 + sub].T)
            ft = (ut > cdf_t[0]).astype(np.uint8)
            ft += ut > cdf_t[1]
            ft += ut > cdf_t[2]
            for j in range(1, read_len):
                np.copyto(ft[j], ft[j - 1], where=rt[j])
            out[s + r0:s + r0 + ft.shape[1]] = _QV[ft.T]
    return out.reshape(-1)


def reads_15mb_block(seed: int = 1) -> np.ndarray:
    """1 Mbp genome, 100K x 150 bp reads, 0.1% substitutions: one 15 MB block."""
    return sequence_stream(1_000_000, 100_000, 150, seed, 0.001)


def reads_50mbp_stream(n_reads: int, seed: int = 2) -> np.ndarray:
    """50 Mbp genome, 150 bp reads, 0.2% substitutions, 0.01% N."""
    return sequence_stream(50_000_000, n_reads, 150, seed, 0.002, 0.0001)


def fastq(n_reads: int, read_len: int, seed: int, genome_len: int = 1_000_000, sub_rate: float = 0.001,
          variable: bool = False) -> bytes:
    """FASTQ text from the same generators: Illumina-style names, reads of
    read_len bases (0.01% N), binned qualities. variable=True trims each read
    and its qualities to a random length in [read_len // 2, read_len]."""
    seq = sequence_stream(genome_len, n_reads, read_len, seed, sub_rate, 0.0001).reshape(n_reads, read_len)
    qual = quality_stream(n_reads, read_len, seed + 7).reshape(n_reads, read_len)
    hs = header_stream(n_reads, seed + 3).tobytes()
    rng = np.random.default_rng(seed + 11)
    lens = rng.integers(read_len // 2, read_len + 1, n_reads) if variable else np.full(n_reads, read_len)
    parts, p = [], 0
    for i in range(n_reads):
        ln = hs[p]  # one varint byte: names are < 128 bytes
        L = int(lens[i])
        parts += [hs[p + 1:p + 1 + ln], b"\n", seq[i, :L].tobytes(), b"\n+\n", qual[i, :L].tobytes(), b"\n"]
        p += 1 + ln
    return b"".join(parts)


def cut_blocks(stream: np.ndarray, block: int = BLOCK):
    return [stream[i:i + block] for i in range(0, stream.size, block)]


def mixed_blocks(count: int, seed: int, lo: int = 100_000, hi: int = BLOCK, total_cap: int | None = None,
                 pool_cap: int = 128 << 20):
    """Blocks of log-uniform size in [lo, hi] from a seeded mix of the three
    generators (sequence / header / quality). Each kind draws its blocks at
    seeded offsets from one generated pool of at most pool_cap bytes (blocks
    are transformed independently, so sharing a pool costs no realism in the
    per-block work and keeps generation to seconds)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), size=count)).astype(np.int64)
    kinds = rng.integers(0, 3, size=count)
    if total_cap is not None:
        keep = np.cumsum(sizes) <= total_cap
        sizes, kinds = sizes[keep], kinds[keep]
    need = [min(pool_cap, max(int(sizes[kinds == k].sum()), int(sizes[kinds == k].max(initial=0)))) for k in range(3)]
    need = [max(n, int(sizes[kinds == k].max(initial=0))) for k, n in enumerate(need)]
    pools = [
        reads_50mbp_stream(need[0] // 150 + 2, seed + 11) if need[0] else np.zeros(0, np.uint8),
        header_stream(need[1] // 60 + 2, seed + 12) if need[1] else np.zeros(0, np.uint8),
        quality_stream(need[2] // 150 + 2, 150, seed + 13) if need[2] else np.zeros(0, np.uint8),
    ]
    blocks = []
    for sz, k in zip(sizes.tolist(), kinds.tolist()):
        o = int(rng.integers(0, pools[k].size - sz + 1))
        blocks.append(pools[k][o:o + sz])
    return blocks, kinds
