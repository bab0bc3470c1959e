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


def test_small_job_lane_skips_the_queue():
    """One regular worker busy with a 100 MB job: a small job submitted right
    after goes to the small-job lane and finishes first, with the same result
    as alone."""
    big = synth.cut_blocks(synth.config2_sequences(700_000, seed=8))[:4]
    small = [synth.sequence_stream(100_000, 1_000, 150, 9, 0.002)]
    print("lane test:", [b.size for b in big], small[0].size, flush=True)
    with serve.Server(hbm_budget=64 << 30, max_batch_bytes=1 << 30, workers=1, small_lane_bytes=32 << 20) as srv:
        srv.wait(srv.submit_forward(big))  # workspaces warm
        jb = srv.submit_forward(big)
        js = srv.submit_forward(small)
        srv.wait(js)
        srv.wait(jb)
        assert js.latency_ms < jb.latency_ms
        assert js.batch_jobs == 1 and jb.batch_jobs == 1
    for j, blocks in ((jb, big), (js, small)):
        out, prim, _ = bwt.encode_batch(blocks)
        assert j.prim_out.tolist() == list(prim)
        for b, blk in enumerate(j.blocks()):
            assert np.array_equal(blk, out[b])
