"""Serving: continuous batching of assembly requests on device-resident meshes.

The server keeps every mesh's topology on the GPU: its tiled plan and its base
coordinates and coefficients. A request names one mesh and brings either a fresh
per-element coefficient field or perturbed coordinates. Serving it means uploading that
input, running one fused iteration (assembly + block-Jacobi prec, src/FEM.cc:183-232)
and leaving values and prec in HBM for the consumer.

Admission is FIFO. A request starts once it has arrived, a slot (stream) is free, and the
HBM budget can hold its footprint: input + values + prec. The budget is managed as an arena
carved from one device allocation, so serving itself makes no hipMalloc/hipFree calls
(hipFree would synchronise the device).

Requests in flight on different slots overlap on the device. The loop polls completion
events and frees memory and slots as requests finish, so new arrivals are admitted
while others still run. Latency is measured from arrival to completion of the request's
last kernel.

The mixed-precision server takes a request's coefficient field as float32, the precision
its element products use (coordinates stay float64: their differences form the element
Jacobian), which halves those uploads.
"""
import bisect
import collections
import time

import numpy as np

from . import gpu
from ._lib import OP_ELA, PREC_FP64, PREC_MIXED, PREC_MIXED_F32COEF
from .setup import Mesh, generate_coef, generate_cube

_ALIGN = 256


class Arena:
    """First-fit allocator over [0, size) bytes, 256-byte aligned blocks."""

    def __init__(self, size):
        self.size = int(size)
        self._free = [(0, self.size)]  # sorted (offset, length)
        self.used = 0

    def alloc(self, nbytes):
        n = (int(nbytes) + _ALIGN - 1) // _ALIGN * _ALIGN
        for i, (off, ln) in enumerate(self._free):
            if ln >= n:
                if ln == n:
                    self._free.pop(i)
                else:
                    self._free[i] = (off + n, ln - n)
                self.used += n
                return off, n
        return None

    def free(self, block):
        off, n = block
        i = bisect.bisect_left(self._free, (off, 0))
        self._free.insert(i, (off, n))
        # merge with the neighbours
        if i + 1 < len(self._free) and off + n == self._free[i + 1][0]:
            self._free[i] = (off, n + self._free[i + 1][1])
            self._free.pop(i + 1)
        if i > 0 and self._free[i - 1][0] + self._free[i - 1][1] == off:
            po, pn = self._free[i - 1]
            self._free[i - 1] = (po, pn + self._free[i][1])
            self._free.pop(i)
        self.used -= n


class Request:
    __slots__ = ("rid", "t_arrival", "mesh", "kind", "payload", "host", "t_start", "t_done")

    def __init__(self, rid, t_arrival, mesh, kind, payload):
        self.rid, self.t_arrival, self.mesh, self.kind, self.payload = rid, t_arrival, mesh, kind, payload
        self.host = None
        self.t_start = self.t_done = None


def served_meshes(count=16, n_lo=1000, n_hi=1000000, seed=1000, jitter=0.2):
    """The served workload's resident meshes: node counts log-uniform in [n_lo, n_hi],
    lap and ela alternating."""
    rng = np.random.default_rng(seed)
    sizes = np.sort(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), count)))
    meshes = []
    for b in range(count):
        n = max(2, int(round(sizes[b] ** (1.0 / 3.0))))
        op = OP_ELA if b % 2 else 0
        coord, e2n, code, nnz = generate_cube(n, jitter, seed + b, 0)
        coef = generate_coef(op, e2n.size // 4, seed + b)
        meshes.append(Mesh(coord, e2n, code, nnz, op, coef=coef))
    return meshes


def make_requests(meshes, count, rate_hz, seed=5000):
    """Poisson arrivals at `rate_hz` (seed), request r (seed 6000 + r): a uniformly chosen
    mesh; even r a fresh lognormal coefficient field, odd r coordinates moved by
    U[-0.1h, 0.1h] (h: the mesh spacing)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, count))
    reqs = []
    for r in range(count):
        rr = np.random.default_rng(6000 + r)
        mi = int(rr.integers(len(meshes)))
        m = meshes[mi]
        if r % 2 == 0:
            coef = generate_coef(m.operatorID, m.nbElem, 6000 + r)
            payload = m.permute_elements(coef)
            kind = "coef"
        else:
            n = round(m.nbNodes ** (1.0 / 3.0))
            h = 1.0 / max(n - 1, 1)
            payload = m.coord + rr.uniform(-0.1 * h, 0.1 * h, m.coord.size)
            kind = "coord"
        reqs.append(Request(r, float(t[r]), mi, kind, np.ascontiguousarray(payload, np.float64)))
    return reqs


class Server:
    """Resident meshes + an HBM budget + `slots` concurrent request streams."""

    def __init__(self, meshes, hbm_budget, slots=4, precision=PREC_FP64, plan_params=None):
        if precision not in (PREC_FP64, PREC_MIXED):
            # MF_PREC_MIXED_F32COEF is chosen per request, for a float32 coefficient
            # payload staged by a PREC_MIXED server; a float64 field read as float is wrong
            raise ValueError(f"Server precision must be PREC_FP64 or PREC_MIXED, got {precision}")
        self.meshes = list(meshes)
        self.precision = precision
        self.dms = [gpu.DeviceMesh(m) for m in self.meshes]
        self.plans = [gpu.Plan(m, **(plan_params or {})) for m in self.meshes]
        self.arena_mem = gpu.DeviceArray(int(hbm_budget), np.uint8)
        self.arena = Arena(hbm_budget)
        self.slots = [gpu.Stream() for _ in range(max(1, int(slots)))]
        self.events = [gpu.Event() for _ in self.slots]
        # one warm launch per plan before any timing (code objects loaded, caches warm)
        for p, dm in zip(self.plans, self.dms):
            p.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=precision)
        gpu.synchronize()

    def _views(self, req, off):
        m = self.meshes[req.mesh]
        base = self.arena_mem.ptr + off
        inp = gpu.DeviceView(base, req.host.size, req.host.dtype, self.arena_mem)
        vo = base + (req.host.nbytes + _ALIGN - 1) // _ALIGN * _ALIGN
        values = gpu.DeviceView(vo, m.nnz * m.dim, np.float64, self.arena_mem)
        po = vo + (values.nbytes + _ALIGN - 1) // _ALIGN * _ALIGN
        prec = gpu.DeviceView(po, m.nbNodes * m.dim, np.float64, self.arena_mem)
        return inp, values, prec

    def _footprint_aligned(self, req):
        m = self.meshes[req.mesh]
        a = lambda n: (n + _ALIGN - 1) // _ALIGN * _ALIGN
        return a(req.host.nbytes) + a(8 * m.dim * m.nnz) + a(8 * m.dim * m.nbNodes)

    def _launch(self, req, slot, inp, values, prec):
        s = self.slots[slot]
        dm = self.dms[req.mesh]
        gpu.copy_h2d_async(inp, req.host, s)
        coord = inp if req.kind == "coord" else dm.coord
        coef = inp if req.kind == "coef" else dm.coef
        precision = self.precision
        if req.kind == "coef" and req.host.dtype == np.float32:
            precision = PREC_MIXED_F32COEF
        self.plans[req.mesh].assemble(coord, coef, values, prec, precision=precision, stream=s)
        self.events[slot].record(s)

    def stage(self, requests):
        """Copies each request's payload into pinned host memory: the receive buffer a
        network front end would hand over (a coefficient field as float32 for the
        mixed-precision server)."""
        for r in requests:
            f32 = self.precision == PREC_MIXED and r.kind == "coef"
            r.host = gpu.HostArray(r.payload.size, np.float32 if f32 else np.float64)
            r.host.array[:] = r.payload

    def run(self, requests, keep_results=False):
        """Serves `requests` (arrival times relative to the start) and returns the stats;
        with keep_results, each request's (values, prec) copied back after it finishes."""
        for r in requests:
            if r.host is None:
                raise ValueError("stage() the requests first")
            if self._footprint_aligned(r) > self.arena.size:
                raise ValueError(f"request {r.rid} needs more than the HBM budget")
        pending = collections.deque(sorted(requests, key=lambda r: r.t_arrival))
        busy = [None] * len(self.slots)  # (request, arena block, views)
        results = {}
        done = []
        memory_waits = set()  # requests that found a slot free but not their HBM
        t0 = time.perf_counter()
        while pending or any(b is not None for b in busy):
            now = time.perf_counter() - t0
            progressed = False
            for i, b in enumerate(busy):
                if b is not None and self.events[i].query():
                    req, blk, views = b
                    req.t_done = time.perf_counter() - t0
                    if keep_results:
                        results[req.rid] = (views[1].numpy(), views[2].numpy())
                    self.arena.free(blk)
                    busy[i] = None
                    done.append(req)
                    progressed = True
            while pending and pending[0].t_arrival <= now:
                free = [i for i, b in enumerate(busy) if b is None]
                if not free:
                    break
                req = pending[0]
                blk = self.arena.alloc(self._footprint_aligned(req))
                if blk is None:
                    memory_waits.add(req.rid)
                    break  # FIFO: wait for memory to free up
                pending.popleft()
                views = self._views(req, blk[0])
                req.t_start = time.perf_counter() - t0
                self._launch(req, free[0], *views)
                busy[free[0]] = (req, blk, views)
                progressed = True
            if not progressed:
                time.sleep(0.00005)
        makespan = max(r.t_done for r in done) - min(r.t_arrival for r in done)
        lat = np.array([r.t_done - r.t_arrival for r in done])
        blocks = sum(self.meshes[r.mesh].nnz for r in done)
        stats = {"requests": len(done), "p50_ms": float(np.percentile(lat, 50) * 1e3),
                 "p99_ms": float(np.percentile(lat, 99) * 1e3),
                 "mean_ms": float(lat.mean() * 1e3), "makespan_s": float(makespan),
                 "blocks_per_s": float(blocks / makespan),
                 "requests_per_s": float(len(done) / makespan),
                 "memory_waits": len(memory_waits)}
        return (stats, results) if keep_results else stats

    def run_alone(self, req):
        """One request with the device to itself: (values, prec)."""
        blk = self.arena.alloc(self._footprint_aligned(req))
        try:
            views = self._views(req, blk[0])
            self._launch(req, 0, *views)
            self.slots[0].sync()
            return views[1].numpy(), views[2].numpy()
        finally:
            self.arena.free(blk)
