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

The request loop is native (mf_server_run, csrc/host_server.cpp): it polls the streams'
completion events, frees memory and streams as launches finish, and admits new arrivals
while others still run. Latency is measured from arrival to the completion of the launch
that served the request, as the loop saw it.

The mixed-precision server takes a request's coefficient field as float32, the precision
its element products use (coordinates stay float64: their differences form the element
Jacobian), which halves those uploads.
"""
import ctypes

import numpy as np

from . import gpu
from ._lib import (MAX_PLAN_JOBS, MF_ERR_BUDGET, OP_ELA, PREC_FP64, PREC_MIXED,
                   PREC_MIXED_F32COEF, REQ_COEF, REQ_COORD, ServerMesh, ServerStats, check, lib)
from ._lib import Request as RequestDesc
from .setup import Mesh, generate_coef, generate_cube



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
    """Resident meshes + an HBM budget + `slots` streams, served by the native request loop
    (mf_server_run); each launch serves up to `max_group` requests at once (slots=1,
    max_group=1: the plain one-at-a-time server). A PREC_MIXED server sizes its plans'
    tiles by the float stage unless plan_params names a stage_precision."""

    def __init__(self, meshes, hbm_budget, slots=4, precision=PREC_FP64, plan_params=None,
                 max_group=2, max_bypass=8):
        if precision not in (PREC_FP64, PREC_MIXED):
            # MF_PREC_MIXED_F32COEF is chosen per request, for a float32 coefficient
            # payload staged by a PREC_MIXED server; a float64 field read as float is wrong
            raise ValueError(f"Server precision must be PREC_FP64 or PREC_MIXED, got {precision}")
        self.meshes = list(meshes)
        self.precision = precision
        self.max_group = max(1, min(int(max_group), MAX_PLAN_JOBS))
        self._lib = lib()
        self._srv = None
        self.dms = [gpu.DeviceMesh(m) for m in self.meshes]
        self._plan_params = dict(plan_params or {})
        self._plan_params.setdefault("stage_precision", precision)
        self.plans = [gpu.Plan(m, **self._plan_params) for m in self.meshes]
        self.budget = int(hbm_budget)
        self.arena_mem = gpu.DeviceArray(self.budget, np.uint8)
        p = ctypes.c_void_p()
        check(self._lib.mf_server_create(ctypes.byref(p), self.arena_mem.ptr, self.budget,
                                         int(slots), self.max_group, max(0, int(max_bypass))),
              "mf_server_create")
        self._srv = p.value
        for m, dm, plan in zip(self.meshes, self.dms, self.plans):
            self._register(m, dm, plan)
        # one warm launch per plan before any timing (code objects loaded, caches warm)
        for p_, dm in zip(self.plans, self.dms):
            p_.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=precision)
        gpu.synchronize()

    def _register(self, mesh, dm, plan):
        ncoef = mesh.nbElem * (2 if mesh.operatorID == OP_ELA else 1)
        sm = ServerMesh(plan.ptr, dm.coord.ptr, dm.coef.ptr if dm.coef is not None else None,
                        mesh.nnz * mesh.dim, mesh.nbNodes * mesh.dim, 3 * mesh.nbNodes, ncoef)
        return check(self._lib.mf_server_add_mesh(self._srv, ctypes.byref(sm)),
                     "mf_server_add_mesh")

    def __del__(self):
        if getattr(self, "_srv", None):
            self._lib.mf_server_destroy(self._srv)
            self._srv = None

    def add_mesh(self, mesh):
        """Admits a new resident mesh while the server holds others (its topology arrives
        with a request stream of its own): device arrays, tiled plan and one warm launch.
        The mesh's setup (coloring, CSR, mask: minifem.Mesh, on the host or the GPU) comes
        first. Returns the mesh's index for requests."""
        dm = gpu.DeviceMesh(mesh)
        plan = gpu.Plan(mesh, **self._plan_params)
        s = gpu.Stream()
        plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=self.precision, stream=s)
        s.sync()
        idx = self._register(mesh, dm, plan)
        self.meshes.append(mesh)
        self.dms.append(dm)
        self.plans.append(plan)
        assert idx == len(self.meshes) - 1
        return idx

    def arena_used(self):
        """Arena bytes held by requests (0 between runs)."""
        return int(self._lib.mf_server_arena_used(self._srv))

    def _mode(self, req):
        """The precision mode of a request's launch: a float32 coefficient field is read as
        float (MF_PREC_MIXED_F32COEF); everything else runs in the server's precision."""
        if req.kind == "coef" and req.host.dtype == np.float32:
            return PREC_MIXED_F32COEF
        return self.precision

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
        n = len(requests)
        arr = (RequestDesc * max(n, 1))()
        results = {}
        for i, r in enumerate(requests):
            if r.host is None:
                raise ValueError("stage() the requests first")
            vo = po = None
            if keep_results:
                m = self.meshes[r.mesh]
                results[r.rid] = (np.empty(m.nnz * m.dim), np.empty(m.nbNodes * m.dim))
                vo, po = results[r.rid][0].ctypes.data, results[r.rid][1].ctypes.data
            arr[i] = RequestDesc(r.mesh, REQ_COEF if r.kind == "coef" else REQ_COORD,
                                 self._mode(r), 0, r.host.ptr, r.host.nbytes,
                                 float(r.t_arrival), vo, po)
        t_start, t_done = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        st = ServerStats()
        rc = self._lib.mf_server_run(self._srv, ctypes.cast(arr, ctypes.c_void_p), n,
                                     t_start.ctypes.data, t_done.ctypes.data, ctypes.byref(st))
        if rc == MF_ERR_BUDGET:
            raise ValueError("a request needs more than the HBM budget")
        check(rc, "mf_server_run")
        for i, r in enumerate(requests):
            r.t_start, r.t_done = float(t_start[i]), float(t_done[i])
        lat = np.array([r.t_done - r.t_arrival for r in requests])
        blocks = sum(self.meshes[r.mesh].nnz for r in requests)
        makespan = float(st.makespan_s)
        stats = {"requests": n, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                 "p99_ms": float(np.percentile(lat, 99) * 1e3),
                 "mean_ms": float(lat.mean() * 1e3), "makespan_s": makespan,
                 "blocks_per_s": float(blocks / makespan),
                 "requests_per_s": float(n / makespan),
                 "memory_waits": int(st.memory_waits), "launches": int(st.launches),
                 "mean_requests_per_launch": n / max(int(st.launches), 1)}
        return (stats, results) if keep_results else stats

    def run_alone(self, req):
        """One request with the device to itself, through the single-plan launch
        (mf_plan_assemble), in buffers of its own: (values, prec)."""
        m = self.meshes[req.mesh]
        s = gpu.Stream()
        inp = gpu.DeviceArray(req.host.size, req.host.dtype)
        values = gpu.DeviceArray(m.nnz * m.dim, np.float64)
        prec = gpu.DeviceArray(m.nbNodes * m.dim, np.float64)
        gpu.copy_h2d_async(inp, req.host, s)
        dm = self.dms[req.mesh]
        coord = inp if req.kind == "coord" else dm.coord
        coef = inp if req.kind == "coef" else dm.coef
        self.plans[req.mesh].assemble(coord, coef, values, prec, precision=self._mode(req),
                                      stream=s)
        s.sync()
        return values.numpy(), prec.numpy()
