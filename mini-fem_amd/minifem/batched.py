"""Batched assembly: many independent meshes of different sizes and operators, each
mesh's fused iteration (mf_plan_assemble: assembly + block-Jacobi prec,
src/FEM.cc:183-232) captured into ONE HIP graph and replayed as one submission.

Inside the graph the meshes are dealt round-robin to `branches` parallel streams
(forked from and joined back to the capturing stream through events), so small
meshes that cannot fill the GPU on their own run side by side. The schedule parameters
of the tiled plans are picked for the whole batch by `tune_batch`.
"""
import numpy as np

from . import gpu
from ._lib import OP_ELA, OP_LAP, PREC_FP64
from .setup import Mesh, generate_coef, generate_cube


def batch_meshes(count=64, n_lo=5000, n_hi=200000, seed=1000, jitter=0.2):
    """The batched workload: `count` jittered cubes with node counts log-uniform in
    [n_lo, n_hi] (n = round(N^(1/3)) per side), lap and ela alternating, per-mesh
    lognormal coefficients; mesh b uses seed + b for geometry and coefficients."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), count))
    meshes = []
    for b in range(count):
        n = max(2, int(round(sizes[b] ** (1.0 / 3.0))))
        op = OP_LAP if b % 2 == 0 else OP_ELA
        coord, e2n, code, nnz = generate_cube(n, jitter, seed + b, 0)
        coef = generate_coef(op, e2n.size // 4, seed + b)
        meshes.append(Mesh(coord, e2n, code, nnz, op, coef=coef))
    return meshes


class Batch:
    """Device-resident meshes with their plans, and the captured batch graph."""

    def __init__(self, meshes, precision=PREC_FP64, branches=8, **plan_params):
        self.meshes = list(meshes)
        self.precision = precision
        self.dms = [gpu.DeviceMesh(m) for m in self.meshes]
        self.plans = [gpu.Plan(m, **plan_params) for m in self.meshes]
        self.stream = gpu.Stream()
        self.branches = max(1, min(int(branches), len(self.meshes)))
        self._side = [gpu.Stream() for _ in range(self.branches)]
        self._fork = gpu.Event()
        self._join = [gpu.Event() for _ in range(self.branches)]
        self.graph = None

    def enqueue(self, stream):
        """Every mesh's iteration, forked over the side streams and joined on `stream`."""
        self._fork.record(stream)
        for s in self._side:
            s.wait(self._fork)
        for i, (plan, dm) in enumerate(zip(self.plans, self.dms)):
            plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=self.precision,
                          stream=self._side[i % self.branches])
        for s, e in zip(self._side, self._join):
            e.record(s)
            stream.wait(e)

    def run_eager(self):
        self.enqueue(self.stream)
        self.stream.sync()

    def capture(self):
        """Builds the batch graph (after one eager run, which also fixes each plan's
        launch grid, so the captured launches make no runtime queries)."""
        self.run_eager()
        self.graph = gpu.Graph(self.stream, self.enqueue)
        return self

    def launch(self):
        if self.graph is None:
            self.capture()
        self.graph.launch(self.stream)

    def time_ms(self, iters=10, warmup=2, graph=True):
        """Mean time of one whole-batch iteration (HIP events on the batch stream)."""
        run = self.launch if graph else (lambda: self.enqueue(self.stream))
        for _ in range(warmup):
            run()
        self.stream.sync()
        e0, e1 = gpu.Event(), gpu.Event()
        e0.record(self.stream)
        for _ in range(iters):
            run()
        e1.record(self.stream)
        return e0.elapsed_ms(e1) / iters

    def blocks(self):
        """CSR blocks assembled per batch iteration."""
        return sum(m.nnz for m in self.meshes)

    def results(self):
        return [(dm.values.numpy(), dm.prec.numpy()) for dm in self.dms]


BATCH_CANDIDATES = [
    dict(tile_lds_bytes=lds, block_threads=bt, max_tile_rows=254)
    for lds, bt in ((40000, 512), (64000, 512), (80000, 512), (150000, 1024))
]


def tune_batch(meshes, precision=PREC_FP64, candidates=None, branches=8, iters=5, log=None):
    """Times the captured batch under each candidate schedule (one parameter set for all
    meshes) and returns (best_params, [(params, ms), ...])."""
    results = []
    for c in candidates or BATCH_CANDIDATES:
        b = Batch(meshes, precision=precision, branches=branches, **c).capture()
        ms = b.time_ms(iters=iters)
        results.append((c, ms))
        if log:
            log(f"batch tune: {c} -> {ms:.3f} ms")
        del b
    best = min(results, key=lambda r: r[1])[0]
    return best, results
