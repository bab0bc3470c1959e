"""The batched uneven launch at its full shape: 512 blocks of log-uniform
size in 100 KB..25 MiB from a seeded mix of the sequence, header and quality
generators (2.7 GB at seed 4: the size law gives a mean of 4.7 MB per block, so the "about 4 GB" of the configuration is its rough total), forward-BWT'd in one call
and inverse-BWT'd in one call (a batch may hold up to 4 GB: the forward sorts
it by alphabet class in chunks of whole blocks, the inverse in sub-batches
of whole blocks below 2^31 bytes).

Every block: its BWT is a permutation of the block, its primary index is the
first sampled index, and the inverse gives the block back. The smallest and
the largest block and one block of each generator kind equal the oracle byte
for byte, and three blocks equal their own one-block call."""
import numpy as np
import pytest

from oracle import oracle as O
from qz_amd import bwt, synth

pytestmark = pytest.mark.gpu

CAP = 4_000_000_000  # one call for the whole workload


def _sub_batches(sizes):
    groups, cur, acc = [], [], 0
    for b, s in enumerate(sizes):
        if cur and acc + s > CAP:
            groups.append(cur)
            cur, acc = [], 0
        cur.append(b)
        acc += s
    groups.append(cur)
    return groups


def test_uneven_512_blocks():
    blocks, kinds = synth.mixed_blocks(512, seed=4)
    sizes = [int(b.size) for b in blocks]
    print(f"uneven 512: {len(blocks)} blocks, {sum(sizes)} B, {min(sizes)}..{max(sizes)} B", flush=True)
    assert len(blocks) == 512 and min(sizes) >= 100_000 and max(sizes) <= synth.BLOCK
    out, prim, aux = [None] * 512, [0] * 512, [None] * 512
    assert len(_sub_batches(sizes)) == 1
    for g in _sub_batches(sizes):
        o, p, a = bwt.encode_batch([blocks[b] for b in g])
        for i, b in enumerate(g):
            out[b], prim[b], aux[b] = o[i], p[i], a[i]
    print("uneven 512: forward done", flush=True)
    back = [None] * 512
    for g in _sub_batches(sizes):
        r = bwt.decode_batch([out[b] for b in g], [prim[b] for b in g])
        for i, b in enumerate(g):
            back[b] = r[i]
    print("uneven 512: inverse done", flush=True)
    for b, blk in enumerate(blocks):
        assert np.array_equal(np.bincount(out[b], minlength=256), np.bincount(blk, minlength=256)), b
        assert prim[b] == aux[b][0] and aux[b].size == (blk.size - 1) // O.sample_rate(blk.size) + 1, b
        assert np.array_equal(back[b], blk), f"round trip {b}"
    picks = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    for k in range(3):
        picks.add(int(np.flatnonzero(kinds == k)[0]))
    for b in sorted(picks):
        u, p, a = O.bwt_aux(blocks[b])
        print(f"uneven 512: oracle block {b} kind {int(kinds[b])} {sizes[b]} B", flush=True)
        assert np.array_equal(out[b], u) and prim[b] == p and np.array_equal(aux[b], a), f"oracle block {b}"
    for b in sorted(picks)[:3]:
        o1, p1, a1 = bwt.encode_batch([blocks[b]])
        assert np.array_equal(o1[0], out[b]) and p1[0] == prim[b] and np.array_equal(a1[0], aux[b])
