"""Serving: continuous batching of assembly requests on device-resident meshes.

The server keeps every mesh's topology on the GPU: its tiled plan and its base
coordinates and coefficients. A request names one mesh and brings either a fresh
per-element coefficient field or perturbed coordinates. Serving it means uploading that
input, running one fused iteration (assembly + block-Jacobi prec, src/FEM.cc:183-232)
and leaving values and prec in HBM for the consumer.

Admission: a request that has arrived is admitted once its footprint (input + values +
prec) fits the HBM budget, in arrival order. One that does not fit waits; later requests
may pass it (first fit) until it has been passed max_bypass times, then admission holds
until it fits. The budget is managed as an arena carved from one device allocation, so
serving itself makes no hipMalloc/hipFree calls (hipFree would synchronise the device).

When a slot (stream) is free, the admitted requests are taken in arrival order, up to
max_group of them with one precision mode, and served by ONE launch over all their plans
(mf_plans_assemble: each tile runs its own mesh and operator), so requests that arrived
together share a launch. The slot uploads their inputs on its stream, then launches; with
several slots, one slot's uploads run under another's kernel. With one slot and
max_group=1 this is the plain one-at-a-time server.

The defaults (4 slots, pairs) come from a sweep on the served workload (16 meshes of
1k-885k nodes, FP64 and mixed, saturating arrivals): FP64 serving moves about 42 GB/s of
inputs over the host-to-device link (57 GB/s measured), and 3-4 slots keep it busy. Pairs
cut the median latency under saturation by a fifth at the same throughput; groups of 4-8
lower throughput (a merged launch runs every member at the largest member's LDS footprint
and lasts as long as its longest job). Uploading on separate copy streams, with the
compute stream waiting for them on the device or on the host, was slower in every form.

The loop polls completion events and frees memory and streams as launches finish, so new
arrivals are admitted while others still run. Latency is measured from arrival to
completion of the launch that served the request.

The mixed-precision server takes a request's coefficient field as float32, the precision
its element products use (coordinates stay float64: their differences form the element
Jacobian), which halves those uploads.
"""
import bisect
import time

import numpy as np

from . import gpu
from ._lib import MAX_PLAN_JOBS, OP_ELA, PREC_FP64, PREC_MIXED, PREC_MIXED_F32COEF
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


def retime(requests, rate_hz, seed=5000):
    """The same requests (meshes, payloads, staged buffers) with Poisson arrivals at
    another rate: one staging serves streams of several rates."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_hz, len(requests)))
    out = []
    for r, ta in zip(requests, t):
        q = Request(r.rid, float(ta), r.mesh, r.kind, r.payload)
        q.host = r.host
        out.append(q)
    return out


class Server:
    """Resident meshes + an HBM budget + `slots` streams; each launch serves up to
    `max_group` requests at once (slots=1, max_group=1: the plain one-at-a-time server)."""

    def __init__(self, meshes, hbm_budget, slots=4, precision=PREC_FP64, plan_params=None,
                 max_group=2, max_bypass=8):
        if precision not in (PREC_FP64, PREC_MIXED):
            # MF_PREC_MIXED_F32COEF is chosen per request, for a float32 coefficient
            # payload staged by a PREC_MIXED server; a float64 field read as float is wrong
            raise ValueError(f"Server precision must be PREC_FP64 or PREC_MIXED, got {precision}")
        self.meshes = list(meshes)
        self.precision = precision
        self.max_group = max(1, min(int(max_group), MAX_PLAN_JOBS))
        self.max_bypass = max(0, int(max_bypass))
        self.dms = [gpu.DeviceMesh(m) for m in self.meshes]
        self._plan_params = dict(plan_params or {})
        self.plans = [gpu.Plan(m, **self._plan_params) for m in self.meshes]
        self.arena_mem = gpu.DeviceArray(int(hbm_budget), np.uint8)
        self.arena = Arena(hbm_budget)
        self.slots = [gpu.Stream() for _ in range(max(1, int(slots)))]
        self.events = [gpu.Event() for _ in self.slots]
        self._fits = {}  # (precision, sorted plan indices) -> one launch fits one CU's LDS
        # one warm launch per plan before any timing (code objects loaded, caches warm)
        for p, dm in zip(self.plans, self.dms):
            p.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=precision)
        gpu.synchronize()

    def add_mesh(self, mesh):
        """Admits a new resident mesh while the server runs (its topology arrives with a
        request stream of its own): device arrays, tiled plan and one warm launch on a slot
        stream. The mesh's host setup (coloring, CSR, mask: minifem.Mesh) comes first.
        Returns the mesh's index for requests."""
        dm = gpu.DeviceMesh(mesh)
        plan = gpu.Plan(mesh, **self._plan_params)
        s = self.slots[0]
        plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=self.precision, stream=s)
        s.sync()
        self.meshes.append(mesh)
        self.dms.append(dm)
        self.plans.append(plan)
        return len(self.meshes) - 1

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

    def _mode(self, req):
        """The precision mode of a request's launch: a float32 coefficient field is read as
        float (MF_PREC_MIXED_F32COEF); everything else runs in the server's precision."""
        if req.kind == "coef" and req.host.dtype == np.float32:
            return PREC_MIXED_F32COEF
        return self.precision

    def _job(self, req, views):
        inp, values, prec = views
        dm = self.dms[req.mesh]
        coord = inp if req.kind == "coord" else dm.coord
        coef = inp if req.kind == "coef" else dm.coef
        return (self.plans[req.mesh], coord, coef, values, prec)

    def _group_fits(self, mode, mesh_ids):
        """Whether the plans of these meshes run as one launch (their tiles fit one CU's LDS
        together; the check reads only the plans)."""
        key = (mode, tuple(sorted(set(mesh_ids))))
        if key not in self._fits:
            jobs = [(self.plans[i], self.dms[i].coord, None, self.dms[i].values, self.dms[i].prec)
                    for i in key[1]]
            self._fits[key] = gpu.plans_lds_bytes(jobs, mode) is not None
        return self._fits[key]

    def _launch(self, group, slot):
        """One launch for a group of requests on one slot: their inputs uploaded on the slot's
        stream, then one multi-plan launch over all of them."""
        s = self.slots[slot]
        jobs = []
        for req, _, views in group:
            gpu.copy_h2d_async(views[0], req.host, s)
            jobs.append(self._job(req, views))
        gpu.plans_assemble(jobs, self._mode(group[0][0]), stream=s)
        self.events[slot].record(s)

    def _take(self, pending, now, memory_waits):
        """Admission for one free slot: requests that have arrived, in arrival order, that fit
        the HBM budget, one precision mode per launch, at most max_group of them. A request
        that does not fit waits; later ones may pass it (first fit) until it has been passed
        max_bypass times, then admission holds until it fits."""
        group, mode, i = [], None, 0
        head = pending[0] if pending else None
        while i < len(pending) and len(group) < self.max_group:
            req = pending[i]
            if req.t_arrival > now:
                break
            m = self._mode(req)
            if mode is not None and (m != mode or not self._group_fits(
                    m, [g[0].mesh for g in group] + [req.mesh])):
                i += 1
                continue
            blk = self.arena.alloc(self._footprint_aligned(req))
            if blk is None:
                memory_waits.add(req.rid)
                if req is head and self._bypassed >= self.max_bypass:
                    break  # the head has waited long enough: nothing passes it
                i += 1
                continue
            del pending[i]
            group.append((req, blk, self._views(req, blk[0])))
            mode = m
        if head is not None and group and pending and pending[0] is head:
            self._bypassed += 1  # the head stays behind the requests just admitted
        elif not pending or pending[0] is not head:
            self._bypassed = 0
        return group

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
        pending = sorted(requests, key=lambda r: r.t_arrival)
        busy = [None] * len(self.slots)  # per slot: [(request, arena block, views), ...]
        results = {}
        done = []
        memory_waits = set()  # requests that found a slot free but not their HBM
        launches = 0
        self._bypassed = 0
        t0 = time.perf_counter()
        while pending or any(b is not None for b in busy):
            for i, b in enumerate(busy):
                if b is not None and self.events[i].query():
                    t_done = time.perf_counter() - t0
                    for req, blk, views in b:
                        req.t_done = t_done
                        if keep_results:
                            results[req.rid] = (views[1].numpy(), views[2].numpy())
                        self.arena.free(blk)
                        done.append(req)
                    busy[i] = None
            now = time.perf_counter() - t0
            for i in range(len(self.slots)):
                if busy[i] is not None or not pending or pending[0].t_arrival > now:
                    continue
                group = self._take(pending, now, memory_waits)
                if not group:
                    break  # nothing admissible until memory frees up or more arrives
                for req, _, _ in group:
                    req.t_start = now
                self._launch(group, i)
                busy[i] = group
                launches += 1
            time.sleep(0)  # yield, then poll again
        makespan = max(r.t_done for r in done) - min(r.t_arrival for r in done)
        lat = np.array([r.t_done - r.t_arrival for r in done])
        blocks = sum(self.meshes[r.mesh].nnz for r in done)
        stats = {"requests": len(done), "p50_ms": float(np.percentile(lat, 50) * 1e3),
                 "p99_ms": float(np.percentile(lat, 99) * 1e3),
                 "mean_ms": float(lat.mean() * 1e3), "makespan_s": float(makespan),
                 "blocks_per_s": float(blocks / makespan),
                 "requests_per_s": float(len(done) / makespan),
                 "memory_waits": len(memory_waits), "launches": launches,
                 "mean_requests_per_launch": len(done) / max(launches, 1)}
        return (stats, results) if keep_results else stats

    def run_alone(self, req):
        """One request with the device to itself, through the single-plan launch
        (mf_plan_assemble): (values, prec)."""
        blk = self.arena.alloc(self._footprint_aligned(req))
        try:
            views = self._views(req, blk[0])
            s = self.slots[0]
            gpu.copy_h2d_async(views[0], req.host, s)
            plan, coord, coef, values, prec = self._job(req, views)
            plan.assemble(coord, coef, values, prec, precision=self._mode(req), stream=s)
            s.sync()
            return views[1].numpy(), views[2].numpy()
        finally:
            self.arena.free(blk)
