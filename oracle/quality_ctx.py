"""Oracle (test infrastructure only): a pure-Python restatement of qz's
quality_ctx codec, crates/qz-lib/src/compression/quality_ctx.rs, used to
check the host C++ coder byte for byte on small inputs.

Each piece follows the Rust line ranges cited on it."""
from __future__ import annotations

import struct

RC_TOP = 1 << 24                       # quality_ctx.rs:15
POS_BIN, MAX_POS_BINS, MAX_QUAL = 5, 64, 50
N_BASE_CTX, N_DELTA = 25, 2            # quality_ctx.rs:20-25
MAX_CONTEXTS = MAX_POS_BINS * MAX_QUAL * N_DELTA * N_BASE_CTX
RESCALE = 1 << 20                      # quality_ctx.rs:26
M32 = 0xFFFFFFFF


class RangeEncoder:                    # quality_ctx.rs:32-96
    def __init__(self):
        self.low, self.range, self.cache, self.cache_size, self.out = 0, M32, 0, 1, bytearray()

    def shift_low(self):
        hi = (self.low >> 32) & 0xFF
        if hi != 0 or (self.low & M32) < 0xFF000000:
            byte = self.cache
            while True:
                self.out.append((byte + hi) & 0xFF)
                byte = 0xFF
                self.cache_size -= 1
                if self.cache_size == 0:
                    break
            self.cache = (self.low >> 24) & 0xFF
        self.cache_size += 1
        self.low = ((self.low & M32) << 8) & M32

    def encode(self, cum, freq, total):
        r = self.range // total
        self.low += cum * r
        if cum + freq < total:
            self.range = (r * freq) & M32
        else:
            self.range = (self.range - r * cum) & M32
        while self.range < RC_TOP:
            self.range = (self.range << 8) & M32
            self.shift_low()

    def finish(self):
        for _ in range(5):
            self.shift_low()
        return bytes(self.out)


class RangeDecoder:                    # quality_ctx.rs:98-163
    def __init__(self, data: bytes):
        self.range, self.code, self.data, self.pos = M32, 0, data, 1 if data else 0
        for _ in range(4):
            self.code = ((self.code << 8) | self._next()) & M32

    def _next(self):
        if self.pos < len(self.data):
            b = self.data[self.pos]
            self.pos += 1
            return b
        return 0

    def decode(self, cum, n, total):
        r = self.range // total
        off = min(self.code // r, total - 1)
        s = 0
        while s + 1 < n and cum[s + 1] <= off:
            s += 1
        c, f = cum[s], cum[s + 1] - cum[s]
        self.code = (self.code - c * r) & M32
        if c + f < total:
            self.range = (r * f) & M32
        else:
            self.range = (self.range - r * c) & M32
        while self.range < RC_TOP:
            self.code = ((self.code << 8) | self._next()) & M32
            self.range = (self.range << 8) & M32
        return s


class Models:                          # quality_ctx.rs:171-247
    def __init__(self, n):
        self.n, self.cum, self.tot = n, [], []

    def alloc(self):
        self.cum.append(list(range(self.n + 1)))
        self.tot.append(self.n)
        return len(self.tot) - 1

    def update(self, slot, sym):
        c = self.cum[slot]
        for i in range(sym + 1, self.n + 1):
            c[i] += 1
        self.tot[slot] += 1
        if self.tot[slot] >= RESCALE:
            acc = 0
            c[0] = 0
            for i in range(self.n):
                f = (c[i + 1] - c[i]) & M32
                acc += max(f >> 1, 1)
                c[i + 1] = acc
            self.tot[slot] = acc


def context_id(pos, pq1, pq2, pb, cb):  # quality_ctx.rs:253-268
    pos_bin = min(pos // POS_BIN, MAX_POS_BINS - 1)
    pq = min(pq1, MAX_QUAL - 1)
    db = 0 if abs(pq1 - pq2) <= 2 else 1
    bc = min(min(pb, 4) * 5 + min(cb, 4), N_BASE_CTX - 1)
    return pos_bin * (MAX_QUAL * N_DELTA * N_BASE_CTX) + pq * (N_DELTA * N_BASE_CTX) + db * N_BASE_CTX + bc


_BASE = {ord(c): i for i, c in enumerate("ACGT")}
_BASE.update({ord(c): i for i, c in enumerate("acgt")})


def _base(b):                          # quality_ctx.rs:271-279
    return _BASE.get(b, 4)


def _slot(slots, models, ctx):         # quality_ctx.rs:347-356 (16-bit slot ids)
    if slots[ctx] == 0xFFFF:
        i = models.alloc()
        slots[ctx] = i & 0xFFFF
        return i
    return slots[ctx]


def compress(qualities, sequences) -> bytes:   # quality_ctx.rs:296-391
    if not qualities:
        return b""
    assert len(qualities) == len(sequences)
    seen = sorted({b - 33 for q in qualities for b in q})
    assert seen and seen[0] >= 0
    smap = {s: i for i, s in enumerate(seen)}
    first = len(qualities[0])
    read_len = 0 if any(len(q) != first for q in qualities) else first & 0xFFFF
    slots = [0xFFFF] * MAX_CONTEXTS
    models = Models(len(seen))
    enc = RangeEncoder()
    for q, s in zip(qualities, sequences):
        assert len(q) <= len(s)
        pq1 = pq2 = 0
        pb = 4
        for j in range(len(q)):
            ph = q[j] - 33
            sym = smap[ph]
            cb = _base(s[j])
            slot = _slot(slots, models, context_id(j, pq1, pq2, pb, cb))
            c = models.cum[slot]
            enc.encode(c[sym], c[sym + 1] - c[sym], models.tot[slot])
            models.update(slot, sym)
            pq2, pq1, pb = pq1, ph, cb
    comp = enc.finish()
    return (bytes([len(seen)]) + bytes(seen) + struct.pack("<HII", read_len, len(qualities), len(comp)) + comp)


def decompress(data: bytes, sequences, num_reads):   # quality_ctx.rs:393-505
    if num_reads == 0:
        return []
    n = data[0]
    symbols = data[1:1 + n]
    read_len, nr, clen = struct.unpack_from("<HII", data, 1 + n)
    assert nr == num_reads
    dec = RangeDecoder(data[11 + n:11 + n + clen])
    slots = [0xFFFF] * MAX_CONTEXTS
    models = Models(n)
    out = []
    for i in range(num_reads):
        s = sequences[i]
        ln = read_len if read_len else len(s)
        pq1 = pq2 = 0
        pb = 4
        q = bytearray()
        for j in range(ln):
            cb = _base(s[j])
            slot = _slot(slots, models, context_id(j, pq1, pq2, pb, cb))
            sym = dec.decode(models.cum[slot], n, models.tot[slot])
            models.update(slot, sym)
            ph = symbols[sym]
            q.append(ph + 33)
            pq2, pq1, pb = pq1, ph, cb
        out.append(bytes(q))
    return out
