#!/usr/bin/env python3
"""Throughput of the BSC block-sorting hot path on one MI355X.

Workload (default): the full-size sequence stream — 50 Mbp synthetic genome,
10M x 150 bp reads, 0.2% substitutions, 0.01% N, 1.5e9 bytes — cut into
qz's 25 MiB BSC blocks (58 blocks, crates/qz-lib/src/compression/mod.rs:
388-413). One step = forward BWT + inverse BWT of every block, inputs
already resident in HBM.

Two forms are timed:
  * batched  — all blocks in one device pass per direction (the value);
  * per-block — one call per block, as qz's rayon map issues them; its
    per-block latency gives p50/p99.

The CPU baseline runs the oracle (SA-IS + LF-walk restatement of libbsc's
BWT stages, not upstream libsais) one block per thread on the host.

Further sections, each in the same JSON line (and each with a --no-* flag):
  * graph_replay — one-block jobs, eager launches versus HIP-graph replay;
  * uneven_batch — 512 blocks of 100 KB..25 MiB from mixed generators, and
    under small_blocks 2000 blocks of 100 KB, batched versus per-block;
  * serving — Poisson arrivals of compress / decompress jobs through the
    serving loop (continuous batching, footprint admission, small-job lane):
    p50 / p99 latency against each job alone;
  * streams — binned qualities (1.5e9 B) and Illumina names (10M reads) as
    their own batches;
  * archive — FASTQ -> .qz -> FASTQ end to end (BSC qualities, and qz's
    quality_ctx choice), and the batch block calls against one call per block;
  * random_access — the device's ceilings for random 4-byte gathers and
    stores over 100 MB and 6 GB arrays, beside a streaming copy: the shapes
    the suffix sort and the inverse are bound by.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
ALG_BYTES_FWD = 10  # per input byte: text 1 + SA write 4 + SA read 4 + BWT write 1
ALG_BYTES_INV = 10  # BWT 1 + LF write 4 + LF read 4 + text 1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(total: int):
    """HBM bytes per launch of the forward and inverse launch sequences, from
    the rocprofv3 request-count runs summarised under profiles/, when they
    were taken on this workload size."""
    p = ROOT / "profiles" / "traffic_seq58x25MiB.json"
    try:
        d = json.loads(p.read_text())
    except (OSError, ValueError):
        return None
    if d.get("input_bytes") != total:
        return None
    return d


def cpu_baseline(blocks, threads: int):
    """Oracle BWT + inverse BWT, one block per thread (qz's rayon model)."""
    from oracle import oracle as O
    O.lib()
    sample = blocks[:threads]
    done = []

    def work(b):
        u, p, _ = O.bwt_aux(b)
        back = O.unbwt(u, p)
        done.append(np.array_equal(back, b))

    t0 = time.perf_counter()
    ths = [threading.Thread(target=work, args=(b,)) for b in sample]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    nbytes = sum(int(b.size) for b in sample)
    assert all(done) and len(done) == len(sample)
    return {
        "value": round(nbytes / dt / 1e6, 3),
        "unit": "MB/s",
        "cores": len(sample),
        "kind": "port",
        "sample": f"{len(sample)} x {sample[0].size} B sequence blocks of the same stream, forward BWT "
                  f"(SA-IS) + inverse BWT (LF walk), one block per thread, {dt:.1f} s; restatement, "
                  "not upstream libsais",
        "per_core_mb_s": round(nbytes / dt / 1e6 / len(sample), 3),
        "nproc": os.cpu_count(),
    }


def random_access_section(lib, array_sizes=(100_000_000, 6_000_000_000), items=1 << 28, iters=5):
    """The device's own ceilings for the access shapes the suffix sort and the
    inverse are bound by: uniform random 4-byte gathers (index and output
    streamed) and random 4-byte stores over arrays of one block's rank array
    (100 MB) and of a whole batch's (6 GB), beside a streaming copy."""
    import ctypes as C
    rows = []
    for nbytes in array_sizes:
        for mode, shape in ((0, "gather"), (1, "scatter"), (2, "copy")):
            ms = C.c_double(0.0)
            rc = lib.qzbsc_calib_access(mode, nbytes, items, iters, C.byref(ms))
            if rc != 0:
                raise RuntimeError(f"qzbsc_calib_access({mode}, {nbytes}) returned {rc}")
            row = {"array_bytes": nbytes, "shape": shape, "ms": round(ms.value, 4)}
            if mode == 2:
                row["gb_s"] = round(2 * nbytes / ms.value / 1e6, 1)
            else:
                row["items"] = items
                row["g_items_per_s"] = round(items / ms.value / 1e6, 3)
            rows.append(row)
    return {"workload": f"{items} uniform random 4-byte accesses per launch (gather: dst[i] = src[idx[i]], "
                        "scatter: dst[idx[i]] = i) and a 16-byte copy of the array; mean of "
                        f"{iters} launches after one warm-up", "rows": rows}


def uneven_section(torch, bwt, synth, count: int, seed: int, cap: int, lo: int = 100_000, hi: int | None = None):
    """Batched vs per-block over blocks of log-uniform size in lo..hi bytes
    (default 100 KB..25 MiB) from the mixed generators (one device pass per
    sub-batch of <= cap bytes, versus one call per block)."""
    t = time.perf_counter()
    hi = synth.BLOCK if hi is None else hi
    blocks, kinds = synth.mixed_blocks(count, seed, lo, hi)
    sizes = np.array([b.size for b in blocks], dtype=np.int32)
    stream = np.concatenate(blocks)
    total = int(sizes.sum())
    log(f"uneven: {count} blocks, {total} B generated in {time.perf_counter() - t:.1f} s")
    dev = torch.device("cuda:0")
    d_in = torch.from_numpy(stream).to(dev)
    d_bwt = torch.empty_like(d_in)
    d_back = torch.empty_like(d_in)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    # batched: contiguous sub-batches of <= cap bytes
    groups, cur, acc = [], [], 0
    for b in range(count):
        if cur and acc + int(sizes[b]) > cap:
            groups.append(cur)
            cur, acc = [], 0
        cur.append(b)
        acc += int(sizes[b])
    groups.append(cur)
    # warm-up, untimed: one block alone, then the largest sub-batch (the
    # workspace grows to its size and block count once, the tuner settles)
    for g in ([0], max(groups, key=lambda g: int(sizes[g[0]:g[-1] + 1].sum()))):
        o = int(offs[g[0]])
        sz = sizes[g[0]:g[-1] + 1]
        prim, _, _ = bwt.dev_encode_batch(d_in.data_ptr() + o, d_bwt.data_ptr() + o, sz)
        bwt.dev_decode_batch(d_bwt.data_ptr() + o, d_back.data_ptr() + o, sz, prim)
    torch.cuda.synchronize()
    # per-block
    lat = []
    t0 = time.perf_counter()
    for b in range(count):
        ts = time.perf_counter()
        o = int(offs[b])
        sz = sizes[b:b + 1]
        prim, _, _ = bwt.dev_encode_batch(d_in.data_ptr() + o, d_bwt.data_ptr() + o, sz)
        bwt.dev_decode_batch(d_bwt.data_ptr() + o, d_back.data_ptr() + o, sz, prim)
        lat.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    t_pb = time.perf_counter() - t0
    ok_pb = torch.equal(d_back, d_in)
    d_back.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    blat = []
    for g in groups:
        ts = time.perf_counter()
        o = int(offs[g[0]])
        sz = sizes[g[0]:g[-1] + 1]
        prim, _, _ = bwt.dev_encode_batch(d_in.data_ptr() + o, d_bwt.data_ptr() + o, sz)
        bwt.dev_decode_batch(d_bwt.data_ptr() + o, d_back.data_ptr() + o, sz, prim)
        blat.append((time.perf_counter() - ts) * 1e3)
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    ok_b = torch.equal(d_back, d_in)
    if not (ok_pb and ok_b):
        print("bench.py: uneven round trip failed", file=sys.stderr)
        sys.exit(1)
    del d_in, d_bwt, d_back
    torch.cuda.empty_cache()
    return {
        "workload": f"{count} blocks, log-uniform {lo / 1e3:.0f} KB..{hi / 2**20:.2f} MiB, seeded mix of "
                    f"sequence/header/quality generators ({total} B); forward BWT + inverse BWT",
        "blocks": count,
        "bytes": total,
        "per_block": {"jobs_per_s": round(count / t_pb, 2), "mb_s": round(total / t_pb / 1e6, 2),
                      "p50_ms": round(float(np.percentile(lat, 50)), 3),
                      "p99_ms": round(float(np.percentile(lat, 99)), 3)},
        "batched": {"jobs_per_s": round(count / t_b, 2), "mb_s": round(total / t_b / 1e6, 2),
                    "sub_batches": len(groups), "batch_ms": [round(x, 2) for x in blat]},
        "batched_vs_per_block_jobs": round(t_pb / t_b, 3),
    }


def graph_section(torch, bwt, synth, sizes=(100_000, 1_000_000, 4_000_000), jobs: int = 40):
    """Per-job latency of one-block forward + inverse BWT jobs on the same
    device buffers, eager launches versus the HIP-graph replay of the fixed
    launch sequences (inverse BWT, forward round 0)."""
    dev = torch.device("cuda:0")
    stream = synth.sequence_stream(5_000_000, 40_000, 150, 9, 0.002, 0.0001)
    rows = []
    for n in sizes:
        blk = stream[:n]
        d_in = torch.from_numpy(np.ascontiguousarray(blk)).to(dev)
        d_bwt = torch.empty_like(d_in)
        d_back = torch.empty_like(d_in)
        sz = np.array([n], dtype=np.int32)
        torch.cuda.synchronize()
        res = {"bytes": n}
        for mode in ("0", "1"):
            os.environ["QZBSC_GRAPH"] = mode
            lat = []
            for j in range(jobs + 2):  # two warm calls: eager, then capture
                ts = time.perf_counter()
                prim, _, _ = bwt.dev_encode_batch(d_in.data_ptr(), d_bwt.data_ptr(), sz)
                bwt.dev_decode_batch(d_bwt.data_ptr(), d_back.data_ptr(), sz, prim)
                if j >= 2:
                    lat.append((time.perf_counter() - ts) * 1e3)
            res["graph_p50_ms" if mode == "1" else "eager_p50_ms"] = round(float(np.percentile(lat, 50)), 4)
        torch.cuda.synchronize()
        if not torch.equal(d_back, d_in):
            print("bench.py: graph section round trip failed", file=sys.stderr)
            sys.exit(1)
        res["speedup"] = round(res["eager_p50_ms"] / res["graph_p50_ms"], 3)
        rows.append(res)
        del d_in, d_bwt, d_back
    os.environ.pop("QZBSC_GRAPH", None)
    c, r = bwt.graph_stats()
    return {"workload": f"{jobs} one-block jobs per size on the same device buffers (sequence stream), "
                        "forward + inverse BWT, host wall time per job",
            "rows": rows, "captures": c, "replays": r}


def streams_section(torch, bwt, synth, reads: int, steps: int):
    """The other streams of the configurations, each as one batch of 25 MiB
    blocks from HBM, forward + inverse BWT like the headline: binned
    qualities (150 per read) and the names of the same reads."""
    dev = torch.device("cuda:0")
    rows = []
    for name, gen in (("binned qualities 1.5e9 B", lambda: synth.quality_stream(reads, 150, 3)),
                      ("Illumina names", lambda: synth.header_stream(reads, 4))):
        stream = gen()
        sizes = np.array([b.size for b in synth.cut_blocks(stream)], dtype=np.int32)
        d_in = torch.from_numpy(stream).to(dev)
        d_b, d_o = torch.empty_like(d_in), torch.empty_like(d_in)
        torch.cuda.synchronize()
        fw, iv = [], []
        for k in range(steps + 1):  # the first is warmup
            t = time.perf_counter()
            prim, _, f = bwt.dev_encode_batch(d_in.data_ptr(), d_b.data_ptr(), sizes)
            i = bwt.dev_decode_batch(d_b.data_ptr(), d_o.data_ptr(), sizes, prim)
            wall = time.perf_counter() - t
            if k:
                fw.append(f)
                iv.append(i)
                walls = wall if k == 1 else min(walls, wall)
        torch.cuda.synchronize()
        ok = torch.equal(d_o, d_in)
        rows.append({"stream": name, "bytes": int(stream.size), "blocks": int(sizes.size),
                     "mb_s": round(stream.size / walls / 1e6, 1), "forward_ms": round(min(fw), 2),
                     "inverse_ms": round(min(iv), 2), "round_trip_ok": bool(ok)})
        del d_in, d_b, d_o
        if not ok:
            print(f"bench.py: {name} round trip failed", file=sys.stderr)
            sys.exit(1)
    return rows


def archive_section(synth, reads: int, threads: int):
    """FASTQ -> .qz v2 archive -> FASTQ, end to end (host stream building,
    batched GPU BWTs, host coder, container), against the per-block pattern
    of qz (each 25 MiB block its own bsc_compress / bsc_decompress call from
    a pool of host threads)."""
    from concurrent.futures import ThreadPoolExecutor
    from qz_amd import archive as A, bsc
    text = synth.fastq(reads, 150, 5, genome_len=5_000_000)
    mb = len(text) / 1e6
    arc = A.compress_fastq(text, quality_compressor="bsc")  # warm up: workspaces and coder tables
    assert A.decompress_archive(arc) == text, "archive round trip mismatch"
    tc, td = [], []
    for _ in range(2):
        t = time.perf_counter()
        arc = A.compress_fastq(text, quality_compressor="bsc")
        tc.append(time.perf_counter() - t)
        t = time.perf_counter()
        back = A.decompress_archive(arc)
        td.append(time.perf_counter() - t)
    assert back == text
    # qz's own choice at this size: quality_ctx qualities (host coder, one
    # thread per 500,000-record blob), BSC for names and reads
    t = time.perf_counter()
    arc_q = A.compress_fastq(text)
    tqc = time.perf_counter() - t
    t = time.perf_counter()
    back = A.decompress_archive(arc_q)
    tqd = time.perf_counter() - t
    assert back == text and A.parse_archive_header(arc_q).quality_compressor == A.COMPRESSOR_QUALITY_CTX
    # the block calls alone: batched (every block of the three streams in one
    # call each way, straight from and into one stream buffer) ...
    _, cs, cq, buf, lens = A.fastq_to_stream_buffer(text)
    B = bsc.BSC_BLOCK_SIZE
    sizes = [min(B, n - i) for n in lens for i in range(0, n, B)]
    bc, bd = [], []
    for _ in range(2):
        t = time.perf_counter()
        zs = bsc.compress_batch_buffer(buf, sizes)
        bc.append(time.perf_counter() - t)
        t = time.perf_counter()
        out, _ = bsc.decompress_batch_buffer(zs)
        bd.append(time.perf_counter() - t)
    assert np.array_equal(out, buf)
    # ... and the per-block pattern over the same blocks
    offs = np.concatenate([[0], np.cumsum(sizes)])
    allp = [buf[offs[k]:offs[k + 1]] for k in range(len(sizes))]
    with ThreadPoolExecutor(threads) as pool:
        t = time.perf_counter()
        comp = list(pool.map(bsc.compress_adaptive_no_lzp, allp))
        pc = time.perf_counter() - t
        t = time.perf_counter()
        dec = list(pool.map(bsc.decompress, comp))
        pd = time.perf_counter() - t
    assert comp == zs and dec == [p.tobytes() for p in allp]
    return {"workload": f"{reads} FASTQ records (Illumina names, 150 bp reads from a 5 Mbp genome, binned "
                        f"qualities) = {len(text)} B; BSC for all three streams (encoding 0, 7-bit lossless "
                        f"qualities, quality_compressor=bsc)",
            "qz_default_quality_ctx": {"ratio": round(len(arc_q) / len(text), 4),
                                       "compress_mb_s": round(mb / tqc, 1), "decompress_mb_s": round(mb / tqd, 1),
                                       "note": "qz's choice from 100,000 records on: quality_ctx qualities "
                                               "(exact-integer host coder, 500,000-record blobs coded in "
                                               "parallel), BSC names and reads; one timed run each"},
            "fastq_mb": round(mb, 1), "archive_bytes": len(arc), "ratio": round(len(arc) / len(text), 4),
            "blocks": len(allp),
            "compress_mb_s": round(mb / min(tc), 1), "decompress_mb_s": round(mb / min(td), 1),
            "batched_blocks_compress_mb_s": round(mb / min(bc), 1),
            "batched_blocks_decompress_mb_s": round(mb / min(bd), 1),
            "per_block_compress_mb_s": round(mb / pc, 1), "per_block_decompress_mb_s": round(mb / pd, 1),
            "host_threads": threads,
            "note": "MB/s of FASTQ text. compress/decompress: end to end (FASTQ<->streams, container, blocks). "
                    "batched_blocks / per_block: the block calls alone on the same 25 MiB blocks, one "
                    "qzbsc_compress_batch / qzbsc_decompress_batch call over all of them vs one bsc_compress / "
                    "bsc_decompress call per block from a pool of host threads"}


def serving_section(bwt, serve, synth, rates, jobs_per_rate: int, seed: int, workers: int,
                    lane_bytes: int = 32 << 20, unit_bytes: int = 64 << 20, policy: int = 0):
    """Poisson arrivals of compress (forward BWT) and decompress (inverse
    BWT) jobs, 70/30, each 1..40 blocks of 100 KB..25 MiB from the mixed
    generators; jobs cut into work units, continuous batching with
    HBM-footprint admission within 288 GB. Request buffers are pinned host
    memory; every job's result is compared with the same job run alone."""
    rng = np.random.default_rng(seed)
    pool, _ = synth.mixed_blocks(400, seed)
    templates = []
    for _ in range(24):
        k = int(rng.integers(1, 41))
        idx = rng.choice(len(pool), size=k, replace=False)
        blks = [pool[i] for i in idx]
        sizes = np.array([b.size for b in blks], dtype=np.int32)
        out, prim, _ = bwt.encode_batch(blks)
        inp = serve.PinnedBuffer(int(sizes.sum()))
        np.concatenate(blks, out=inp.array)
        bw = serve.PinnedBuffer(int(sizes.sum()))
        np.concatenate(out, out=bw.array)
        templates.append({"inp": inp, "sizes": sizes, "bwt": bw, "prim": np.array(prim, dtype=np.int32)})
    mean_bytes = float(np.mean([t["sizes"].sum() for t in templates]))
    # pinned output buffers: one as large as the largest template for the
    # jobs run alone, and per rate one for every job, allocated before its
    # arrivals start
    big = int(max(t["sizes"].sum() for t in templates))
    solo_out = serve.PinnedBuffer(big)
    res = {"workload": f"Poisson arrivals, 70% compress (forward BWT) / 30% decompress (inverse BWT), "
                       f"jobs of 1..40 blocks of 100 KB..25 MiB (mean {mean_bytes / 1e6:.1f} MB), "
                       f"{jobs_per_rate} jobs per rate, {workers} workers + a lane for jobs <= "
                       f"{lane_bytes >> 20} MiB, work units of {unit_bytes >> 20} MiB, "
                       f"{'oldest-job-first' if policy == 0 else 'fewest-bytes-left-first'} order, "
                       "pinned request buffers",
           "rates": []}
    budget = 288 * 10**9
    mismatches = 0

    def submit(srv, t, kind, buf):
        n = int(t["sizes"].sum())
        o = buf.array[:n]
        if kind == 0:
            return srv.submit_raw(0, t["inp"].array, t["sizes"], o)
        return srv.submit_raw(1, t["bwt"].array, t["sizes"], o, t["prim"])

    def check(j, t, kind):
        want = t["bwt"].array if kind == 0 else t["inp"].array
        ok = np.array_equal(j.out, want) and (kind == 1 or j.prim_out.tolist() == t["prim"].tolist())
        return 0 if ok else 1

    with serve.Server(hbm_budget=budget, max_batch_bytes=1 << 30, workers=workers, small_lane_bytes=lane_bytes,
                      unit_bytes=unit_bytes, policy=policy) as srv:
        res["batch_cap_bytes"] = srv.batch_cap()
        # each job alone (both kinds), twice; the second run counts
        solo = {}
        for rep in range(2):
            for k, t in enumerate(templates):
                for kind in (0, 1):
                    j = srv.wait(submit(srv, t, kind, solo_out))
                    mismatches += check(j, t, kind)
                    solo[(k, kind)] = j.latency_ms
        res["single_job_ms"] = round(float(np.median([solo[(k, 0)] for k in range(len(templates))])), 2)
        res["single_job_p99_ms"] = round(float(np.percentile(list(solo.values()), 99)), 2)
        for rate in rates:
            arrivals = np.cumsum(rng.exponential(1.0 / rate, size=jobs_per_rate))
            picks = rng.integers(0, len(templates), size=jobs_per_rate)
            kinds = (rng.random(jobs_per_rate) < 0.3).astype(int)
            outs = [serve.PinnedBuffer(int(templates[p]["sizes"].sum())) for p in picks]
            sub = []
            t0 = time.perf_counter()
            for i, (a, p, k) in enumerate(zip(arrivals, picks, kinds)):
                dt = t0 + a - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                sub.append(submit(srv, templates[p], int(k), outs[i]))
            lat, nbat, slow, alone = [], [], [], []
            for j, p, k in zip(sub, picks, kinds):
                srv.wait(j)
                lat.append(j.latency_ms)
                nbat.append(j.batch_jobs)
                a = solo[(int(p), int(k))]
                alone.append(a)
                slow.append(j.latency_ms / a)
            wall = time.perf_counter() - t0
            # results are checked after the clock stops: comparing 100 MB
            # per job on the host is not serving work
            for j, p, k in zip(sub, picks, kinds):
                mismatches += check(j, templates[p], int(k))
            for o in outs:
                o.free()
            nbytes = sum(int(templates[p]["sizes"].sum()) for p in picks)
            res["rates"].append({"offered_jobs_per_s": rate, "jobs_per_s": round(len(sub) / wall, 2),
                                 "mb_s": round(nbytes / wall / 1e6, 1),
                                 "p50_ms": round(float(np.percentile(lat, 50)), 1),
                                 "p99_ms": round(float(np.percentile(lat, 99)), 1),
                                 "alone_p99_ms": round(float(np.percentile(alone, 99)), 1),
                                 "p99_vs_alone_p99": round(float(np.percentile(lat, 99) / np.percentile(alone, 99)), 2),
                                 "p99_vs_median_single_job": round(float(np.percentile(lat, 99)) / res["single_job_ms"], 2),
                                 "slowdown_p50": round(float(np.percentile(slow, 50)), 2),
                                 "slowdown_p99": round(float(np.percentile(slow, 99)), 2),
                                 "mean_jobs_per_batch": round(float(np.mean(nbat)), 2),
                                 # the three slowest jobs: what sets the tail
                                 "slowest": [{"mb": round(int(templates[picks[i]]["sizes"].sum()) / 1e6, 1),
                                              "kind": "compress" if kinds[i] == 0 else "decompress",
                                              "ms": round(lat[i], 1), "alone_ms": round(alone[i], 1),
                                              "queue_ms": round(sub[i].queue_ms, 1)}
                                             for i in np.argsort(lat)[::-1][:3]]})
            log(f"serving rate {rate}: {res['rates'][-1]}")
        res["peak_inflight_bytes"] = srv.peak_bytes()
    for t in templates:
        t["inp"].free()
        t["bwt"].free()
    solo_out.free()
    # every served job (alone and in the mix) equals the same job's result
    # from one direct batch call: mix invariance, checked on every job
    res["results_checked"] = len(templates) * 4 + len(rates) * jobs_per_rate
    res["results_mismatched"] = mismatches
    if mismatches:
        print(f"bench.py: {mismatches} served jobs differ from their results alone", file=sys.stderr)
        sys.exit(1)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--reads", type=int, default=10_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-per-block", action="store_true", help="skip the per-block form")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-block batch section")
    ap.add_argument("--uneven-blocks", type=int, default=512)
    ap.add_argument("--batch-cap", type=int, default=4_000_000_000)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop section")
    ap.add_argument("--no-graph", action="store_true", help="skip the eager-vs-graph section")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="first copy this many MiB device-to-device (a known-size kernel for counter scaling)")
    ap.add_argument("--serve-rates", type=str, default="20,50,100,200")
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--serve-lane-mib", type=int, default=128, help="small-job lane: jobs of at most this many MiB")
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-unit-mib", type=int, default=128, help="serving work units (MiB of whole blocks)")
    ap.add_argument("--serve-policy", type=int, default=0, help="0 oldest job first, 1 fewest bytes left first")
    ap.add_argument("--no-archive", action="store_true", help="skip the .qz archive section")
    ap.add_argument("--no-streams", action="store_true", help="skip the quality / header stream section")
    ap.add_argument("--archive-reads", type=int, default=1_000_000)
    ap.add_argument("--no-ceiling", action="store_true", help="skip the random-access ceiling section")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this benchmark measures one GPU only; blocks are independent and would be "
              "dealt to GPUs round-robin with no collective. Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import qz_amd
    from qz_amd import synth, bwt, serve

    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    lib = qz_amd.lib()

    t = time.perf_counter()
    stream = synth.reads_50mbp_stream(args.reads, seed=2)
    blocks = synth.cut_blocks(stream)
    sizes = np.array([b.size for b in blocks], dtype=np.int32)
    total = int(sizes.sum())
    log(f"generated {total} B in {len(blocks)} blocks ({time.perf_counter() - t:.1f} s)")

    dev = torch.device("cuda:0")
    if args.calib_copy:
        # a plain copy of known size, its source first pushed out of the
        # caches by writing 2 GiB elsewhere
        src = torch.ones(args.calib_copy << 20, dtype=torch.uint8, device=dev)
        evict = torch.zeros(2 << 30, dtype=torch.uint8, device=dev)
        evict.fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst, evict
    d_in = torch.from_numpy(stream).to(dev)
    d_bwt = torch.empty_like(d_in)
    d_back = torch.empty_like(d_in)
    torch.cuda.synchronize()

    def step_batched():
        prim, _, fms = bwt.dev_encode_batch(d_in.data_ptr(), d_bwt.data_ptr(), sizes)
        ims = bwt.dev_decode_batch(d_bwt.data_ptr(), d_back.data_ptr(), sizes, prim)
        return fms, ims

    for _ in range(args.warmup):
        step_batched()
    torch.cuda.synchronize()
    fwd_ms, inv_ms, lat = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        f, i = step_batched()
        lat.append((time.perf_counter() - ts) * 1e3)
        fwd_ms.append(f)
        inv_ms.append(i)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ok = torch.equal(d_back, d_in)
    log(f"batched: {args.steps} steps in {wall:.3f} s, round trip ok={ok}")
    if not ok:
        print("bench.py: inverse BWT did not restore the input", file=sys.stderr)
        sys.exit(1)
    value = total * args.steps / wall / 1e6

    per_block = None
    if not args.no_per_block:
        blat = []
        offs = np.concatenate([[0], np.cumsum(sizes)])
        t1 = time.perf_counter()
        for b in range(len(blocks)):
            ts = time.perf_counter()
            o = int(offs[b])
            sz = sizes[b:b + 1]
            prim, _, _ = bwt.dev_encode_batch(d_in.data_ptr() + o, d_bwt.data_ptr() + o, sz)
            bwt.dev_decode_batch(d_bwt.data_ptr() + o, d_back.data_ptr() + o, sz, prim)
            blat.append((time.perf_counter() - ts) * 1e3)
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t1
        per_block = {
            "value": round(total / pwall / 1e6, 2),
            "unit": "MB/s",
            "jobs_per_s": round(len(blocks) / pwall, 3),
            "p50_ms": round(float(np.percentile(blat, 50)), 3),
            "p99_ms": round(float(np.percentile(blat, 99)), 3),
        }
        log(f"per-block: {per_block}")

    fwd = float(np.mean(fwd_ms)) / 1e3
    inv = float(np.mean(inv_ms)) / 1e3
    ach_fwd = ALG_BYTES_FWD * total / fwd / 1e9
    ach_inv = ALG_BYTES_INV * total / inv / 1e9
    roof = {
        "bound": "hbm",
        "kernel": "forward BWT (prefix-doubling suffix sort + BWT gather), one batched launch sequence",
        "achieved": round(ach_fwd, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(ach_fwd / HBM_PEAK_GBS, 5),
        "traffic": None,
        "launch_ms": round(fwd * 1e3, 3),
        "inverse": {"achieved": round(ach_inv, 2), "frac": round(ach_inv / HBM_PEAK_GBS, 5),
                    "launch_ms": round(inv * 1e3, 3), "traffic": None},
        "bwt_ibwt_frac": round((ALG_BYTES_FWD + ALG_BYTES_INV) * total / (fwd + inv) / 1e9 / HBM_PEAK_GBS, 5),
    }

    ceiling = None
    if not args.no_ceiling:
        ceiling = random_access_section(lib)
        log(f"random-access ceiling: {ceiling}")

    graphs = None
    if not args.no_graph:
        graphs = graph_section(torch, bwt, synth)
        log(f"graphs: {graphs}")

    tr = measured_traffic(total)
    if tr:
        roof["traffic"] = tr["forward"]["hbm_bytes"]
        roof["inverse"]["traffic"] = tr["inverse"]["hbm_bytes"]
        roof["traffic_note"] = ("HBM bytes per launch from rocprofv3 --pmc runs of this workload "
                                "(profiles/traffic_seq58x25MiB.json): TCC read and write requests resolved by "
                                "request size, the rule checked on a 6 GB copy kernel: "
                                f"{tr['forward']['hbm_bytes_per_input_byte']} B per input byte forward, "
                                f"{tr['inverse']['hbm_bytes_per_input_byte']} inverse")

    uneven = None
    if not args.no_uneven:
        uneven = uneven_section(torch, bwt, synth, args.uneven_blocks, 4, args.batch_cap)
        log(f"uneven: {uneven}")
        # many small jobs: where one call per block leaves the GPU idle
        small = uneven_section(torch, bwt, synth, 2000, 6, args.batch_cap, 100_000, 100_000)
        log(f"small blocks: {small}")
        uneven["small_blocks"] = small

    serving = None
    if not args.no_serve:
        from qz_amd import serve
        from qz_amd._native import lib as _lib
        # the server's workers keep their own full-batch workspaces resident:
        # give back what the idle engines of the sections above still hold
        torch.cuda.empty_cache()
        log(f"trim before serving: {_lib().qzbsc_trim() / 1e9:.1f} GB released")
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        serving = serving_section(bwt, serve, synth, rates, args.serve_jobs, 5, args.serve_workers,
                                  args.serve_lane_mib << 20, args.serve_unit_mib << 20, args.serve_policy)

    streams = None
    if not args.no_streams:
        streams = streams_section(torch, bwt, synth, args.reads, 2)
        log(f"streams: {streams}")

    archive = None
    if not args.no_archive:
        archive = archive_section(synth, args.archive_reads, max(1, min(args.cpu_threads, os.cpu_count() or 1)))
        log(f"archive: {archive}")

    cpu = None
    if not args.no_cpu:
        thr = max(1, min(args.cpu_threads, os.cpu_count() or 1, len(blocks)))
        cpu = cpu_baseline(blocks, thr)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "BWT+iBWT MB/s per GPU (batched vs per-block) + p50/p99 job latency, 25MB blocks",
        "value": round(value, 2),
        "unit": "MB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"full-size sequence stream: 50 Mbp genome, {args.reads} x 150 bp reads, 0.2% subs, "
                        f"0.01% N = {total} B in {len(blocks)} x 25 MiB blocks; forward BWT + inverse BWT",
            "blocks": len(blocks),
            "block_bytes": int(sizes.max()),
            "seed": 2,
        },
        "batched": {
            "value": round(value, 2),
            "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
        },
        "per_block": per_block,
        "roofline": roof,
        "cpu_baseline": cpu,
        "random_access": ceiling,
        "graph_replay": graphs,
        "uneven_batch": uneven,
        "serving": serving,
        "streams": streams,
        "archive": archive,
        "tuning": bwt.tuning(),
        # device bytes per input byte that one call over the headline batch
        # holds (qzbsc_footprint): staging scales with the batch, the forward
        # sort workspace with one 512 MiB chunk of whole blocks
        "footprint_b_per_input_b": {
            "forward": round(serve.footprint(0, total, len(blocks)) / total, 2),
            "forward_one_25MiB_block": round(serve.footprint(0, synth.BLOCK, 1) / synth.BLOCK, 2),
            "inverse": round(serve.footprint(1, total, len(blocks)) / total, 2)},
    }
    if per_block:
        out["batched_vs_per_block"] = round(value / per_block["value"], 3)
    if cpu:
        out["gpu_vs_cpu"] = round(value / cpu["value"], 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
