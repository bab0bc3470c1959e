"""Full-size workloads on the GPU, through the batch C-ABI.

* 1.5e9 B of sequence reads (50 Mbp genome, 10M x 150 bp) in 58 x 25 MiB
  blocks, one batch;
* 1.5e9 B of binned qualities in 25 MiB blocks, one batch;
* the names of the same 10M reads (Illumina style, about 0.64 GB).

Every block is checked by size-independent properties: the BWT is a
permutation of the block (byte histogram kept), the primary index lies in
[1, n] and is the first sampled index, there are (n-1)/r + 1 sampled
indexes, and the inverse BWT gives the block back. The first and the last
block of each batch are also compared with the oracle byte for byte
(BWT, primary and every sampled index).
"""
import numpy as np
import pytest

from oracle import oracle as O
from qz_amd import bwt, synth

pytestmark = pytest.mark.gpu


def _check(name, stream, oracle_blocks=(0, -1)):
    blocks = synth.cut_blocks(stream)
    print(f"{name}: {stream.size} B in {len(blocks)} blocks", flush=True)
    out, prim, aux = bwt.encode_batch(blocks)
    print(f"{name}: forward done", flush=True)
    for b, (blk, u) in enumerate(zip(blocks, out)):
        n = blk.size
        assert np.array_equal(np.bincount(u, minlength=256), np.bincount(blk, minlength=256)), f"{name} block {b}"
        assert 1 <= prim[b] <= n and aux[b][0] == prim[b], f"{name} block {b} primary {prim[b]}"
        assert aux[b].size == (n - 1) // bwt.sample_rate(n) + 1
    back = bwt.decode_batch(out, prim)
    print(f"{name}: inverse done", flush=True)
    for b, blk in enumerate(blocks):
        assert np.array_equal(back[b], blk), f"{name} block {b} round trip"
    for b in oracle_blocks:
        u, p, a = O.bwt_aux(blocks[b])
        assert np.array_equal(out[b], u) and prim[b] == p and np.array_equal(aux[b], a), f"{name} block {b} vs oracle"
        print(f"{name}: block {b} matches the oracle", flush=True)


def test_reads_1_5gb_full_size():
    _check("reads 1.5e9 B", synth.reads_50mbp_stream(10_000_000, seed=2))


def test_binned_qualities_1_5gb_full_size():
    _check("binned qualities 1.5e9 B", synth.quality_stream(10_000_000, 150, 3), oracle_blocks=(0,))


def test_illumina_names_full_size():
    _check("Illumina names", synth.header_stream(10_000_000, 4), oracle_blocks=(0,))
