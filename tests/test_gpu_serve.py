"""Serving loop: jobs submitted concurrently come back identical to the same
job run alone; admission under a tight HBM budget serialises batches."""
import threading

import numpy as np
import pytest

from qz_amd import bwt, serve, synth

pytestmark = pytest.mark.gpu


def _jobs(seed, count):
    rng = np.random.default_rng(seed)
    pool, _ = synth.mixed_blocks(60, seed, lo=5_000, hi=900_000)
    jobs = []
    for _ in range(count):
        k = int(rng.integers(1, 6))
        idx = rng.choice(len(pool), size=k, replace=False)
        jobs.append([pool[i] for i in idx])
    return jobs


def test_results_same_alone_and_in_the_mix():
    jobs = _jobs(3, 24)
    alone = [bwt.encode_batch(j) for j in jobs]
    with serve.Server(hbm_budget=64 << 30, max_batch_bytes=64 << 20, workers=2) as srv:
        fw = [None] * len(jobs)

        def client(lo, hi):
            for i in range(lo, hi):
                fw[i] = srv.submit_forward(jobs[i])

        ths = [threading.Thread(target=client, args=(a, a + 8)) for a in (0, 8, 16)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        inv = []
        for i, j in enumerate(fw):
            srv.wait(j)
            out, prim, aux = alone[i]
            for b, blk in enumerate(j.blocks()):
                assert np.array_equal(blk, out[b]), f"job {i} block {b}"
            assert j.prim_out.tolist() == list(prim)
            for b in range(len(prim)):
                assert np.array_equal(j.aux[b, :aux[b].size], aux[b])
            inv.append(srv.submit_inverse(out, prim))
        batched = max(j.batch_jobs for j in fw)
        for i, j in enumerate(inv):
            srv.wait(j)
            for b, blk in enumerate(j.blocks()):
                assert np.array_equal(blk, jobs[i][b]), f"inverse job {i} block {b}"
    assert batched >= 2, "no job was batched with another"


def test_tight_budget_admits_one_batch_at_a_time():
    jobs = _jobs(5, 6)
    need = max(serve.footprint(0, sum(b.size for b in j), len(j)) for j in jobs)
    with serve.Server(hbm_budget=need, max_batch_bytes=1 << 30, workers=2) as srv:
        sub = [srv.submit_forward(j) for j in jobs]
        for j in sub:
            srv.wait(j)
        assert srv.peak_bytes() <= need
    for j, jb in zip(sub, jobs):
        out, prim, _ = bwt.encode_batch(jb)
        for b, blk in enumerate(j.blocks()):
            assert np.array_equal(blk, out[b])
