"""Serving loop: jobs submitted concurrently come back identical to the
oracle and to the same job run alone; admission under a tight HBM budget
serialises batches without growing device memory; jobs are cut into work
units; the small-job lane skips the queue; a job at the serving
configuration's shape (Poisson arrivals, 1..40 blocks of 100 KB..25 MiB)
matches the oracle on sampled blocks and round-trips through the server."""
import threading
import time

import numpy as np
import pytest

from oracle import oracle as O
from qz_amd import bwt, serve, synth
from qz_amd.bwt import BscError

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


def test_results_same_as_oracle_alone_and_in_the_mix():
    jobs = _jobs(3, 24)
    with serve.Server(hbm_budget=64 << 30, max_batch_bytes=64 << 20, workers=2, unit_bytes=1 << 20) as srv:
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
            for b, blk in enumerate(j.blocks()):
                u, p, a = O.bwt_aux(jobs[i][b])
                assert np.array_equal(blk, u), f"job {i} block {b}"
                assert j.prim_out[b] == p and np.array_equal(j.aux[b, :a.size], a)
            inv.append(srv.submit_inverse(j.blocks(), j.prim_out))
        batched = max(j.batch_jobs for j in fw)
        for i, j in enumerate(inv):
            srv.wait(j)
            for b, blk in enumerate(j.blocks()):
                assert np.array_equal(blk, jobs[i][b]), f"inverse job {i} block {b}"
    assert batched >= 2, "no job was batched with another"
    for i in (0, 7, 23):  # alone
        out, prim, aux = bwt.encode_batch(jobs[i])
        assert all(np.array_equal(x, y) for x, y in zip(out, fw[i].blocks())) and list(prim) == fw[i].prim_out.tolist()


def test_budget_bounds_inflight_and_resident_memory_does_not_grow():
    """A 4 MiB batch cap and a budget of exactly the two workers' resident
    workspaces: in-flight footprints stay within it, a job larger than the
    cap runs in units, and device memory after the jobs is what server
    creation left (nothing grows)."""
    jobs = _jobs(5, 6)
    big = [b for j in jobs for b in j]  # one job of every block: larger than the batch cap
    cap = 4 << 20
    nb = cap // 65536 + 64
    budget = 2 * (serve.footprint(0, cap, nb) + serve.footprint(1, cap, nb))
    with serve.Server(hbm_budget=budget, max_batch_bytes=cap, workers=2, small_lane_bytes=0,
                      unit_bytes=2 << 20) as srv:
        assert srv.batch_cap() == cap
        free0, _ = serve.device_memory()
        sub = [srv.submit_forward(j) for j in jobs] + [srv.submit_forward(big)]
        for j in sub:
            srv.wait(j)
        free1, _ = serve.device_memory()
        assert 0 < srv.peak_bytes() <= budget
        assert free1 >= free0 - (64 << 20), (free0, free1)  # the resident workspaces did not grow
    for j, jb in zip(sub, jobs + [big]):
        out, prim, _ = bwt.encode_batch(jb)
        for b, blk in enumerate(j.blocks()):
            assert np.array_equal(blk, out[b])


def test_block_beyond_cap_refused_and_double_wait():
    with serve.Server(hbm_budget=8 << 30, max_batch_bytes=2 << 20, workers=1, small_lane_bytes=0) as srv:
        with pytest.raises(BscError) as e:
            srv.submit_forward([np.full(3 << 20, 65, np.uint8)])
        assert e.value.code == -2
        j = srv.submit_forward([synth.sequence_stream(100_000, 1_000, 150, 4, 0.002)])
        srv.wait(j)
        with pytest.raises(BscError):
            srv.wait(j)


def test_small_job_lane_skips_the_queue():
    """One regular worker busy with a 100 MB job (one unit): a small job
    submitted right after goes to the small-job lane and finishes first, with
    the same result as alone."""
    big = synth.cut_blocks(synth.reads_50mbp_stream(700_000, seed=8))[:4]
    small = [synth.sequence_stream(100_000, 1_000, 150, 9, 0.002)]
    print("lane test:", [b.size for b in big], small[0].size, flush=True)
    with serve.Server(hbm_budget=64 << 30, max_batch_bytes=1 << 30, workers=1, small_lane_bytes=32 << 20,
                      unit_bytes=256 << 20) as srv:
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


@pytest.mark.parametrize("policy", [0, 1])
def test_serving_shape_poisson_jobs(policy):
    """12 compress jobs of 1..40 blocks of 100 KB..25 MiB from the mixed
    generators, arriving as a Poisson stream at 50 jobs/s, through pinned
    request buffers; each job's inverse is submitted as it completes. Every
    block round-trips; the smallest block of each job and the largest
    block overall equal the oracle; every block equals its result alone."""
    rng = np.random.default_rng(40 + policy)
    pool, _ = synth.mixed_blocks(120, 41)
    jobs = []
    for _ in range(12):
        k = int(rng.integers(1, 41))
        jobs.append([pool[i] for i in rng.choice(len(pool), size=k, replace=False)])
    print("serving shape:", [sum(b.size for b in j) for j in jobs], flush=True)
    bufs = []
    with serve.Server(hbm_budget=288 * 10**9, max_batch_bytes=1 << 30, workers=3, small_lane_bytes=128 << 20,
                      unit_bytes=64 << 20, policy=policy) as srv:
        arrivals = np.cumsum(rng.exponential(1 / 50.0, size=len(jobs)))
        t0 = time.perf_counter()
        fw = []
        for a, blocks in zip(arrivals, jobs):
            dt = t0 + a - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            n = sum(b.size for b in blocks)
            pin_in, pin_out = serve.PinnedBuffer(n), serve.PinnedBuffer(n)
            bufs += [pin_in, pin_out]
            np.concatenate(blocks, out=pin_in.array)
            sizes = np.array([b.size for b in blocks], np.int32)
            fw.append(srv.submit_raw(0, pin_in.array, sizes, pin_out.array))
        inv = []
        for j in fw:
            srv.wait(j)
            inv.append(srv.submit_raw(1, j.out, j.sizes, np.empty_like(j.out), j.prim_out))
        for j in inv:
            srv.wait(j)
        lat = [j.latency_ms for j in fw]
        print(f"policy {policy}: latency p50 {np.percentile(lat, 50):.1f} ms max {max(lat):.1f} ms", flush=True)
    largest = max((b.size, ji, bi) for ji, blocks in enumerate(jobs) for bi, b in enumerate(blocks))
    for ji, (blocks, jf, jv) in enumerate(zip(jobs, fw, inv)):
        got = jf.blocks()
        for bi, blk in enumerate(jv.blocks()):
            assert np.array_equal(blk, blocks[bi]), f"job {ji} block {bi} round trip"
        checks = {int(np.argmin([b.size for b in blocks]))}
        if largest[1] == ji:
            checks.add(largest[2])
        for bi in checks:
            u, p, _ = O.bwt_aux(blocks[bi])
            assert np.array_equal(got[bi], u) and jf.prim_out[bi] == p, f"job {ji} block {bi} oracle"
        out, prim, _ = bwt.encode_batch(blocks)
        assert all(np.array_equal(x, y) for x, y in zip(out, got)) and list(prim) == jf.prim_out.tolist()
    for b in bufs:
        b.free()
