"""Python face of the serving loop (qzbsc_server_*): continuous batching of
forward / inverse BWT jobs, cut into work units of whole blocks, with
HBM-footprint admission."""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from ._native import lib, u8p, i32p
from .bwt import BscError


class Job:
    def __init__(self, jid, kind, inp, out, sizes, prim_in, prim_out, aux):
        self.id, self.kind = jid, kind
        self.out, self.prim_out, self.aux = out, prim_out, aux
        self.sizes = sizes
        self.latency_ms = self.queue_ms = None
        self.batch_jobs = None
        self._buffers = (inp, out, sizes, prim_in, prim_out, aux)

    def blocks(self):
        res, o = [], 0
        for s in self.sizes.tolist():
            res.append(self.out[o:o + s])
            o += s
        return res


class PinnedBuffer:
    """Page-locked host memory (qzbsc_host_alloc) as a numpy uint8 array:
    request buffers the server copies by DMA instead of through the driver's
    staging of pageable memory. Freed by free() or when collected."""

    def __init__(self, nbytes: int):
        self._p = lib().qzbsc_host_alloc(max(1, int(nbytes)))
        if not self._p:
            raise BscError(-2, "qzbsc_host_alloc")
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(1, int(nbytes))).from_address(self._p))[:int(nbytes)]

    def free(self):
        if self._p:
            self.array = None
            lib().qzbsc_host_free(self._p)
            self._p = None

    def __del__(self):
        self.free()


class Server:
    """workers regular workers plus, when small_lane_bytes > 0, one worker
    that takes only jobs of at most that many bytes. Jobs are cut into work
    units of unit_bytes (whole blocks); policy 0 serves the oldest job first,
    1 the job with the fewest bytes not yet started."""

    def __init__(self, hbm_budget: int, max_batch_bytes: int = 1 << 30, workers: int = 2,
                 small_lane_bytes: int = 32 << 20, unit_bytes: int = 64 << 20, policy: int = 0):
        self._h = lib().qzbsc_server_create_opts(hbm_budget, max_batch_bytes, workers, small_lane_bytes,
                                                 unit_bytes, policy)
        if not self._h:
            raise BscError(-7, "qzbsc_server_create")
        # every submitted job's arrays, by id, until wait() or close(): the
        # workers read and write them, so they must outlive the job
        self._live: dict[int, Job] = {}
        self._mu = threading.Lock()

    def close(self):
        if self._h:
            lib().qzbsc_server_destroy(self._h)  # runs every pending job to its end first
            self._h = None
        with self._mu:
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def peak_bytes(self) -> int:
        return int(lib().qzbsc_server_peak_bytes(self._h))

    def batch_cap(self) -> int:
        return int(lib().qzbsc_server_batch_cap(self._h))

    def submit_raw(self, kind: int, inp: np.ndarray, sizes: np.ndarray, out: np.ndarray, prim_in=None) -> Job:
        """Submit with caller-owned contiguous buffers (no copies)."""
        return self._submit_arrays(kind, inp, np.ascontiguousarray(sizes, dtype=np.int32), out, prim_in)

    def _submit(self, kind, blocks, primaries=None):
        arrs = [np.ascontiguousarray(b, dtype=np.uint8) for b in blocks]
        sizes = np.array([a.size for a in arrs], dtype=np.int32)
        inp = np.concatenate(arrs)
        out = np.empty_like(inp)
        return self._submit_arrays(kind, inp, sizes, out, primaries)

    def _submit_arrays(self, kind, inp, sizes, out, primaries):
        if not self._h:
            raise BscError(-1, "qzbsc_server_submit on a closed server")
        if inp.size != out.size or int(sizes.sum()) != inp.size:
            raise BscError(-1, "qzbsc_server_submit: buffer sizes")
        prim_in = np.ascontiguousarray(primaries, dtype=np.int32) if primaries is not None else None
        prim_out = np.zeros(sizes.size, dtype=np.int32)
        aux = np.zeros((sizes.size, 256), dtype=np.int32)
        with self._mu:
            jid = lib().qzbsc_server_submit(self._h, kind, inp.ctypes.data_as(u8p), out.ctypes.data_as(u8p),
                                            sizes.ctypes.data_as(i32p), sizes.size,
                                            prim_in.ctypes.data_as(i32p) if prim_in is not None else None,
                                            prim_out.ctypes.data_as(i32p), aux.ctypes.data_as(i32p))
            if jid < 0:
                raise BscError(jid, "qzbsc_server_submit")
            job = Job(jid, kind, inp, out, sizes, prim_in, prim_out, aux)
            self._live[jid] = job
        return job

    def submit_forward(self, blocks) -> Job:
        """Compress side: forward BWT of the blocks."""
        return self._submit(0, blocks)

    def submit_inverse(self, bwts, primaries) -> Job:
        """Decompress side: inverse BWT of the blocks."""
        return self._submit(1, bwts, primaries)

    def wait(self, job: Job) -> Job:
        lat, q, nb = C.c_double(0), C.c_double(0), C.c_int(0)
        rc = lib().qzbsc_server_wait(self._h, job.id, C.byref(lat), C.byref(q), C.byref(nb))
        if rc != -1:
            with self._mu:
                self._live.pop(job.id, None)  # the native job is gone once wait returned its status
        if rc != 0:
            raise BscError(rc, "qzbsc_server_wait")
        job.latency_ms, job.queue_ms, job.batch_jobs = lat.value, q.value, nb.value
        return job


def footprint(kind: int, nbytes: int, blocks: int) -> int:
    return int(lib().qzbsc_footprint(kind, nbytes, blocks))


def device_memory():
    """(free, total) device bytes."""
    f, t = C.c_longlong(0), C.c_longlong(0)
    rc = lib().qzbsc_device_memory(C.byref(f), C.byref(t))
    if rc != 0:
        raise BscError(rc, "qzbsc_device_memory")
    return f.value, t.value
