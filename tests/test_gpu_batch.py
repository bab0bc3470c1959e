"""Batched uneven blocks (sequence / header / quality mix) in one device
pass: every block bit-exact against the oracle, and identical to the same
block transformed alone."""
import numpy as np
import pytest

from oracle import oracle as O
from qz_amd import bwt, synth

pytestmark = pytest.mark.gpu


def test_uneven_mixed_batch_matches_oracle_and_alone():
    blocks, kinds = synth.mixed_blocks(24, seed=9, lo=2_000, hi=1_500_000)
    blocks = blocks + [np.full(30_000, 65, np.uint8), np.frombuffer(b"ACGT" * 20, np.uint8)]
    print("sizes", [b.size for b in blocks], flush=True)
    out, prim, aux = bwt.encode_batch(blocks)
    for i, blk in enumerate(blocks):
        u, p, a = O.bwt_aux(blk)
        assert np.array_equal(out[i], u) and prim[i] == p and np.array_equal(aux[i], a), f"block {i}"
    # same result alone as in the mix
    for i in (0, 5, len(blocks) - 1):
        o1, p1, a1 = bwt.encode_batch([blocks[i]])
        assert np.array_equal(o1[0], out[i]) and p1[0] == prim[i] and np.array_equal(a1[0], aux[i])
    back = bwt.decode_batch(out, prim)
    for i, blk in enumerate(blocks):
        assert np.array_equal(back[i], blk), f"round trip {i}"
