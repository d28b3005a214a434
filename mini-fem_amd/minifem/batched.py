"""Batched assembly: many independent meshes of different sizes and operators, each
mesh's fused iteration (mf_plan_assemble: assembly + block-Jacobi prec,
src/FEM.cc:183-232) captured into ONE HIP graph and replayed as one submission.

Two forms:
  - per mesh (union=False): one plan and one launch per mesh, the meshes dealt
    round-robin to `branches` parallel streams inside the graph (forked from and joined
    back to the capturing stream through events);
  - union (union=True): the meshes of one operator become ONE disjoint-union mesh with
    one plan, and the plans of both operators run as ONE persistent launch
    (mf_plans_assemble: the tiles of the lap union, then those of the ela union, each tile
    running its own operator). The workgroups of that launch take tiles of every mesh, so
    small meshes never leave the GPU idle and no launch waits for another.
    Results are bit-identical to the per-mesh form: the greedy coloring of a disjoint
    union colors every element as its own mesh does, and its color permutation keeps
    each mesh's elements in their own relative order, so every CSR block sums the same
    products in the same order.
The schedule parameters of the tiled plans are picked for the whole batch by `tune_batch`.
"""
import numpy as np

from . import gpu
from ._lib import OP_ELA, OP_LAP, PREC_FP64, PREC_MIXED
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


class UnionMesh:
    """The disjoint union of meshes of one operator as one Mesh: mesh k's nodes are
    node_off[k] .. node_off[k+1]-1 of the union, its elements keep their input order
    after those of meshes 0..k-1, its coefficients and boundary codes come along.

    partition_coord: the union's coordinates with the meshes set side by side along x
    (a gap of `gap` times the widest mesh between neighbours), used only to group nodes
    into tiles, so a tile holds nodes of one mesh except where a cut falls between two.
    """

    def __init__(self, meshes, gap=0.5):
        self.parts = list(meshes)
        if not self.parts:
            raise ValueError("UnionMesh needs at least one mesh")
        op = self.parts[0].operatorID
        if any(m.operatorID != op for m in self.parts):
            raise ValueError("UnionMesh: all meshes must have the same operator")
        have_coef = [m.coef_original is not None for m in self.parts]
        if any(have_coef) and not all(have_coef):
            raise ValueError("UnionMesh: either every mesh or none has coefficients")
        n = np.array([m.nbNodes for m in self.parts], np.int64)
        self.node_off = np.concatenate([[0], np.cumsum(n)])
        if self.node_off[-1] >= 2**31 - 1:
            raise ValueError("UnionMesh: too many nodes for 32-bit node ids")
        e2n = np.concatenate([m.elemToNode_original.astype(np.int64) + self.node_off[k]
                              for k, m in enumerate(self.parts)]).astype(np.int32)
        coord = np.concatenate([m.coord for m in self.parts])
        code = np.concatenate([m.boundNodesCode for m in self.parts])
        coef = (np.concatenate([m.coef_original for m in self.parts])
                if all(have_coef) else None)
        cap = sum(m.nnz for m in self.parts)
        if cap > 2**31 - 1:
            raise ValueError("UnionMesh: too many CSR blocks for 32-bit row pointers")
        self.mesh = Mesh(coord, e2n, code, cap, op, coef=coef)
        if self.mesh.nnz != cap:
            raise AssertionError("UnionMesh: CSR size differs from the sum of the parts")
        pc = coord.reshape(-1, 3).copy()
        lo = np.array([m.coord.reshape(-1, 3)[:, 0].min() for m in self.parts])
        hi = np.array([m.coord.reshape(-1, 3)[:, 0].max() for m in self.parts])
        step = (hi - lo).max() * (1.0 + gap) + 1e-12
        for k in range(len(self.parts)):
            a, b = self.node_off[k], self.node_off[k + 1]
            pc[a:b, 0] += k * step - lo[k]
        self.partition_coord = pc.ravel()

    def value_range(self, k):
        """Mesh k's CSR values in the union's value array (doubles)."""
        r, d = self.mesh.row, self.mesh.dim
        return int(r[self.node_off[k]]) * d, int(r[self.node_off[k + 1]]) * d

    def prec_range(self, k):
        """Mesh k's prec blocks in the union's prec array (doubles)."""
        d = self.mesh.dim
        return int(self.node_off[k]) * d, int(self.node_off[k + 1]) * d


class Batch:
    """Device-resident meshes with their plans, and the captured batch graph.

    A batch built for a mixed precision mode sizes its tiles by the float stage (larger
    tiles, mf_plan_params.stage_precision) unless plan_params names a stage_precision;
    such a batch runs the mixed modes only."""

    def __init__(self, meshes, precision=PREC_FP64, branches=8, union=False, **plan_params):
        self.meshes = list(meshes)
        self.precision = precision
        plan_params = dict(plan_params)
        plan_params.setdefault("stage_precision", precision)
        self.stage_precision = plan_params["stage_precision"]
        self.union = bool(union)
        self._where = None
        self._jobs = None  # union form: the job table of its one launch
        if self.union:
            # One union per operator, in order of first appearance; _where[i] = (union,
            # part) of input mesh i.
            groups = {}
            for i, m in enumerate(self.meshes):
                groups.setdefault(m.operatorID, []).append(i)
            self.unions = [UnionMesh([self.meshes[i] for i in idx]) for idx in groups.values()]
            self._where = [None] * len(self.meshes)
            for u, idx in enumerate(groups.values()):
                for k, i in enumerate(idx):
                    self._where[i] = (u, k)
            self.dms = [gpu.DeviceMesh(u.mesh) for u in self.unions]
            self.plans = self._union_plans(plan_params)
        else:
            self.dms = [gpu.DeviceMesh(m) for m in self.meshes]
            self.plans = [gpu.Plan(m, **plan_params) for m in self.meshes]
        self.stream = gpu.Stream()
        self.branches = max(1, min(int(branches), len(self.plans)))
        self._side = [gpu.Stream() for _ in range(self.branches)]
        self._fork = gpu.Event()
        self._join = [gpu.Event() for _ in range(self.branches)]
        self.graph = None

    def _union_plans(self, plan_params):
        """The union plans, sized so that they run as one launch: when their largest tiles
        do not fit one CU's LDS together, they are rebuilt with the tile budget lowered by
        the overflow."""
        params = dict(plan_params)
        for _ in range(3):
            plans = [gpu.Plan(u.mesh, partition_coord=u.partition_coord, **params)
                     for u in self.unions]
            jobs = [(p, dm.coord, dm.coef, dm.values, dm.prec) for p, dm in zip(plans, self.dms)]
            modes = (PREC_MIXED,) if params["stage_precision"] != PREC_FP64 else (PREC_FP64, PREC_MIXED)
            fits = [gpu.plans_lds_bytes(jobs, prec) for prec in modes]
            if all(f is not None for f in fits):
                self._jobs = jobs
                return plans
            budget = params.get("tile_lds_bytes") or 155000
            params["tile_lds_bytes"] = budget - 6144  # the front of the other operator's tiles
        raise RuntimeError("union plans do not fit one launch")

    def set_precision(self, precision):
        """Switches the precision mode of every launch; the captured graph, if any, is
        dropped and captured again on the next launch(). Plans sized for the FP64 stage
        serve every mode, plans sized for the float stage the mixed modes only."""
        if precision == PREC_FP64 and self.stage_precision != PREC_FP64:
            raise ValueError("this batch's tiles are sized for the float stage (mixed modes "
                             "only); build it with stage_precision=0 to run FP64")
        self.precision = precision
        self.graph = None

    def enqueue(self, stream):
        """Every mesh's iteration: the union form's one launch on `stream`, or the per-mesh
        launches forked over the side streams and joined on `stream`."""
        if self._jobs is not None:
            gpu.plans_assemble(self._jobs, self.precision, stream)
            return
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
        """Builds the batch graph after one eager run (every output then holds a result
        before the first replay). The plans fixed their launch grids at creation, so the
        captured launches make no runtime queries."""
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

    def launches(self):
        """Assembly launches per batch iteration."""
        return 1 if self._jobs is not None else len(self.plans)

    def results(self):
        """(values, prec) of every input mesh, in input order."""
        if not self.union:
            return [(dm.values.numpy(), dm.prec.numpy()) for dm in self.dms]
        full = [(dm.values.numpy(), dm.prec.numpy()) for dm in self.dms]
        out = []
        for u, k in self._where:
            (va, vb), (pa, pb) = self.unions[u].value_range(k), self.unions[u].prec_range(k)
            out.append((full[u][0][va:vb].copy(), full[u][1][pa:pb].copy()))
        return out


BATCH_CANDIDATES = [
    dict(tile_lds_bytes=lds, block_threads=bt, max_tile_rows=rows)
    for lds, bt, rows in ((40000, 512, 254), (56000, 512, 254), (64000, 512, 254),
                          (72000, 512, 254), (80000, 512, 254), (100000, 1024, 254),
                          (120000, 1024, 254), (135000, 1024, 254), (150000, 1024, 254),
                          (150000, 1024, 48))
]


def tune_batch(meshes, precision=PREC_FP64, candidates=None, branches=8, iters=5, log=None,
               union=False):
    """Times the captured batch under each candidate schedule (one parameter set for all
    meshes) and returns (best_params, [(params, ms), ...])."""
    results = []
    for c in candidates or BATCH_CANDIDATES:
        b = Batch(meshes, precision=precision, branches=branches, union=union, **c).capture()
        ms = b.time_ms(iters=iters)
        results.append((c, ms))
        if log:
            log(f"batch tune: {c} -> {ms:.3f} ms")
        del b
    best = min(results, key=lambda r: r[1])[0]
    return best, results
