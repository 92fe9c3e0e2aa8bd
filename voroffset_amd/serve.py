"""Served form: resident volumes, HBM-budget admission, batched execution.

A ``MorphoServer`` keeps volumes resident in HBM (``put`` uploads once and
returns a handle); requests name a handle, an operation and a radius.  Each
``step`` admits waiting requests first-come first-served while the estimated
device footprint of the batch stays within the HBM budget (and the batch
within ``max_batch``), runs them as one ``vo_morph_batch_device`` call, and
keeps every result resident as a new volume.  There is no reference
counterpart: the reference runs one CLI invocation per volume
(offset3d.cpp:116-136).
"""

from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from . import _native

OPS = ("dilation", "erosion", "closing", "opening", "noop")
_DILATIONS = {"noop": 0, "dilation": 1, "erosion": 1, "closing": 2, "opening": 2}


@dataclasses.dataclass
class Resident:
    nx: int
    ny: int
    off: object          # device buffer (DeviceArray or a pooled DeviceBuffer)
    z: object
    z_lo: float          # device z-range (shifted for FP32 storage)
    z_hi: float
    shift: float         # FP32 storage: device z = true z - shift
    spacing: float
    endpoints: int
    pooled: bool = False  # buffers taken from the server's pool (given back on drop)

    @property
    def ncols(self):
        return self.nx * self.ny


class DevicePool:
    """Device buffers kept for reuse.  A result takes a free buffer of the
    smallest power-of-two size class that holds it (allocating one only when
    none is free) and a dropped result gives its buffers back, so a steady
    stream of requests allocates and frees nothing once the pool is warm."""

    def __init__(self):
        self.free = {}      # size class (bytes) -> [DeviceBuffer]
        self.allocated = 0  # bytes held by the pool, in use or free

    @staticmethod
    def size_class(nbytes):
        return max(256, 1 << max(0, int(nbytes) - 1).bit_length())

    def take(self, nbytes):
        c = self.size_class(nbytes)
        lst = self.free.get(c)
        if lst:
            return lst.pop()
        self.allocated += c
        return _native.DeviceBuffer(c)

    def give(self, buf):
        self.free.setdefault(buf.nbytes, []).append(buf)

    def close(self):
        for lst in self.free.values():
            for b in lst:
                b.free()
        self.free.clear()


@dataclasses.dataclass
class Request:
    rid: int
    handle: int
    op: str
    radius: float
    arrival: float
    start: float = 0.0
    done: float = 0.0
    result: int = -1     # handle of the resident result


def footprint_bytes(ncols, endpoints, op, storage_bytes=4, caps=96, ocap=24):
    """Estimated device bytes a job occupies inside a batch, per column: the
    pass-1 slots (lo, hi and a 16-bit line offset per slot), their block
    summaries (lo and hi per 8 slots) and last hi, the triple count, the
    gather's interval lists (2 * ocap doubles whatever the storage), its
    output arena (2 * ocap endpoints), the start, length and offset words;
    plus the outputs of the job's stages at their bounds (the batch program
    sizes them before it runs).  ocap is the largest list capacity a batch uses (24 when
    it holds a radius below 6, 16 otherwise)."""
    if op not in _DILATIONS:
        raise ValueError(f"unknown operation {op!r}")
    if _DILATIONS[op] == 0:
        return 0
    per_col = (caps * (2 * storage_bytes + 2) + 2 * (caps // 8) * storage_bytes + storage_bytes + 4
               + 2 * ocap * 8 + 2 * ocap * storage_bytes + 8 + 8 + 8)
    dilations = _DILATIONS[op]
    complements = {"dilation": 0, "erosion": 2, "closing": 2, "opening": 2}[op]
    # stage outputs, each sized for its bound: a dilation's column holds at
    # most ocap intervals, a complement adds at most two endpoints a column
    volume = (ncols + 1) * 8 + (endpoints + 2 * ncols) * storage_bytes
    dilated = (ncols + 1) * 8 + 2 * ocap * ncols * storage_bytes
    return ncols * per_col + dilations * dilated + complements * volume


class MorphoServer:
    def __init__(self, precision="fp32", device=0, hbm_budget=64 << 30, max_batch=64, ctx=None):
        if precision not in _native.PRECISIONS:
            raise ValueError("precision must be 'fp32' or 'fp64'")
        self.precision = precision
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.ctx = ctx if ctx is not None else _native.Context(device)
        self.vols: dict[int, Resident] = {}
        self.queue: collections.deque[Request] = collections.deque()
        self.pool = DevicePool()
        self._next = 0
        self._rid = 0
        self.batches = []    # (size, seconds) of every executed batch

    # -- resident volumes ------------------------------------------------
    def put(self, vol):
        """Uploads a CompressedVolume; returns its handle."""
        zlo, zhi = vol.z_range()
        shift = _native.storage_shift(zlo, zhi) if self.precision == "fp32" else 0.0
        dtype = np.float32 if self.precision == "fp32" else np.float64
        r = Resident(vol.nx, vol.ny, _native.DeviceArray(vol.off), _native.DeviceArray((vol.z - shift).astype(dtype)),
                     zlo - shift, zhi - shift, shift, vol.spacing, int(vol.z.size))
        return self._register(r)

    def _register(self, r):
        h = self._next
        self._next += 1
        self.vols[h] = r
        return h

    def get(self, handle, like):
        """Downloads a resident volume (geometry from `like`)."""
        r = self.vols[handle]
        dtype = np.float32 if self.precision == "fp32" else np.float64
        off = _native.to_host(r.off.address, np.int64, r.ncols + 1)
        z = _native.to_host(r.z.address, dtype, r.endpoints).astype(np.float64) + r.shift
        return like.like(off, z)

    def drop(self, handle):
        r = self.vols.pop(handle)
        if r.pooled:
            self.pool.give(r.off)
            self.pool.give(r.z)
        else:
            r.off.free()
            r.z.free()

    # -- requests ----------------------------------------------------------
    def submit(self, handle, op, radius, arrival=None):
        if handle not in self.vols:
            raise KeyError(f"no resident volume {handle}")
        if op not in OPS:
            raise ValueError(f"unknown operation {op!r}")
        req = Request(self._rid, handle, op, float(radius), time.perf_counter() if arrival is None else arrival)
        self._rid += 1
        self.queue.append(req)
        return req

    def footprint(self, req):
        r = self.vols[req.handle]
        return footprint_bytes(r.ncols, r.endpoints, req.op, 4 if self.precision == "fp32" else 8)

    def admit(self, now=None):
        """Requests to run next: FIFO while the batch fits the budget (the
        head of the queue always runs, alone if it exceeds the budget)."""
        batch, used = [], 0
        while self.queue and len(batch) < self.max_batch:
            req = self.queue[0]
            if now is not None and req.arrival > now:
                break
            need = self.footprint(req)
            if batch and used + need > self.budget:
                break
            batch.append(self.queue.popleft())
            used += need
        return batch

    def run(self, batch):
        """Runs admitted requests as one batch; results stay resident."""
        if not batch:
            return []
        jobs = []
        for req in batch:
            r = self.vols[req.handle]
            jobs.append((req.op, req.radius, r.nx, r.ny, r.off.address, r.z.address, r.z_lo, r.z_hi))
        t0 = time.perf_counter()
        for req in batch:
            req.start = t0
        outs, _ = self.ctx.morph_batch_device(self.precision, jobs)
        esize = 4 if self.precision == "fp32" else 8
        # results leave the context's workspace (reused by the next batch) for
        # pooled buffers: no allocation once the pool is warm, no host data
        copies = []
        for req, (v, cnt) in zip(batch, outs):
            src = self.vols[req.handle]
            ncols = v.nx * v.ny
            off = self.pool.take((ncols + 1) * 8)
            z = self.pool.take(max(cnt, 1) * esize)
            copies += [(off.address, v.offsets, (ncols + 1) * 8), (z.address, v.z, cnt * esize)]
            req.result = self._register(Resident(v.nx, v.ny, off, z, src.z_lo, src.z_hi, src.shift, src.spacing,
                                                 int(cnt), pooled=True))
        self.ctx.copy_device(copies)  # one synchronisation for the batch
        t1 = time.perf_counter()
        for req in batch:
            req.done = t1
        self.batches.append((len(batch), t1 - t0))
        return batch

    def step(self, now=None):
        return self.run(self.admit(now))

    def close(self):
        for h in list(self.vols):
            self.drop(h)
        self.pool.close()
        self.ctx.close()


def simulate(server, handles, rate, n_requests, radius_range=(2.0, 48.0), seed=0, saturate=False):
    """Poisson arrivals at `rate` per second of n requests over the resident
    `handles` (op uniform over the four operations, radius uniform); runs the
    server in real time and returns the finished requests.  The arrival
    times are a Poisson process conditioned on n arrivals in n / rate
    seconds (n sorted uniform times), so the offered rate is `rate` itself
    rather than one draw's.  With `saturate` the same requests all arrive at
    once (the server's capacity)."""
    rng = np.random.default_rng(seed)
    times = np.sort(rng.uniform(0.0, n_requests / rate, n_requests))
    t0 = time.perf_counter() + 0.01
    arrivals = t0 + (np.zeros(n_requests) if saturate else times)
    plan = [(int(rng.integers(len(handles))), OPS[int(rng.integers(4))], float(rng.uniform(*radius_range)))
            for _ in range(n_requests)]
    done, k = [], 0
    while len(done) < n_requests:
        now = time.perf_counter()
        while k < n_requests and arrivals[k] <= now:
            h, op, r = plan[k]
            server.submit(handles[h], op, r, arrival=float(arrivals[k]))
            k += 1
        if not server.queue:
            time.sleep(max(0.0, min(arrivals[k] - now, 0.01)))
            continue
        finished = server.step(now)
        for req in finished:
            server.drop(req.result)  # results are not reused here
        done += finished
    return done


def latency_report(done):
    lat = np.array([r.done - r.arrival for r in done]) * 1e3
    return {"requests": len(done), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3), "mean_ms": round(float(lat.mean()), 3)}
