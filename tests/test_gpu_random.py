"""Seeded randomized parity of the fused plan against the oracle, bit for bit. Cubes of
random size and jitter, node ids shuffled or not, the element order shuffled or not (so
coloring, permutation and tiles all change), with or without per-element coefficients,
under random tile schedules, in FP64 and mixed precision (against the oracle's mixed
restatement); the FP64 plans that fit one launch together then run as one multi-plan
launch, each with the bits of its own launch. Unstructured meshes: Delaunay
tetrahedralizations of random points (node valences up to ~90)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCHEDULES = [dict(), dict(block_threads=512), dict(tile_lds_bytes=8192),
             dict(tile_lds_bytes=155000, block_threads=1024, max_tile_rows=254),
             dict(tile_lds_bytes=78 * 1024, block_threads=512), dict(max_tile_rows=7)]


def random_case(mf, orc, rng):
    n = int(rng.integers(3, 21))
    op = int(rng.integers(0, 2))
    jitter = float(rng.choice([0.0, 0.1, 0.25, 0.3]))
    seed, shuffle = int(rng.integers(1, 1000)), int(rng.integers(0, 1000))
    coef_seed = int(rng.integers(1, 1000)) if rng.integers(0, 2) else None
    elem_shuffle = bool(rng.integers(0, 2))
    sched = SCHEDULES[int(rng.integers(0, len(SCHEDULES)))]
    precision = int(rng.choice([mf.PREC_FP64, mf.PREC_FP64, mf.PREC_MIXED]))
    desc = (f"n={n} op={op} jitter={jitter} seed={seed} shuffle={shuffle} coef={coef_seed} "
            f"elem_shuffle={elem_shuffle} schedule={sched} precision={precision}")
    coord, e2n, code, nnz = mf.generate_cube(n, jitter, seed, shuffle)
    nE = e2n.size // 4
    coef = None if coef_seed is None else mf.generate_coef(op, nE, coef_seed)
    if elem_shuffle:
        p = rng.permutation(nE)
        e2n = e2n.reshape(-1, 4)[p].ravel().copy()
        if coef is not None:
            coef = coef.reshape(nE, -1)[p].ravel().copy()
    mesh = mf.Mesh(coord, e2n, code, nnz, op, coef=coef)
    om = orc.OracleMesh(coord, e2n, code, nnz, op, coef=coef)
    return desc, mesh, om, sched, precision


def same_prec(a, b):
    """Equal bits, or equal where finite with infinities in the same places (a singular
    block: the reference's inversion divides by zero the same way)."""
    return np.array_equal(a, b) or (np.array_equal(np.isinf(a), np.isinf(b)) and
                                    np.array_equal(a[~np.isinf(a)], b[~np.isinf(b)]))


@pytest.mark.parametrize("seed", [11, 12])
def test_random_meshes_bit_exact(mf, orc, gpu, seed):
    from minifem._lib import PlanJob
    rng = np.random.default_rng(seed)
    pool = []
    for case in range(8):
        desc, mesh, om, sched, precision = random_case(mf, orc, rng)
        print(f"[random {seed}.{case}] {desc}", flush=True)
        plan = gpu.Plan(mesh, **sched)
        dm = gpu.DeviceMesh(mesh)
        dm.values.fill(np.nan)
        dm.prec.fill(np.nan)
        plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=precision)
        ov = om.assemble(mixed=precision == mf.PREC_MIXED)
        v, p = dm.values.numpy(), dm.prec.numpy()
        assert np.array_equal(v, ov), desc
        assert same_prec(p, om.prec(ov)[0]), desc
        if precision == mf.PREC_FP64 and sched.get("block_threads", 1024) == 1024:
            pool.append((plan, dm, v, p))
    lib = mf.lib()

    def table(items):
        return (PlanJob * len(items))(*[
            PlanJob(pl.ptr, d.coord.ptr, d.coef.ptr if d.coef is not None else None,
                    d.values.ptr, d.prec.ptr, None) for pl, d, _, _ in items])
    together = []
    for item in pool:
        if lib.mf_plans_lds_bytes(ctypes.cast(table(together + [item]), ctypes.c_void_p),
                                  len(together) + 1, mf.PREC_FP64) >= 0:
            together.append(item)
    print(f"[random {seed}] {len(together)} of {len(pool)} FP64 plans in one launch", flush=True)
    if len(together) < 2:
        return
    for _, d, _, _ in together:
        d.values.fill(np.nan)
        d.prec.fill(np.nan)
    mf.check(lib.mf_plans_assemble(ctypes.cast(table(together), ctypes.c_void_p), len(together),
                                   mf.PREC_FP64, None), "mf_plans_assemble")
    gpu.synchronize()
    for pl, d, v, p in together:
        assert np.array_equal(d.values.numpy(), v) and np.array_equal(d.prec.numpy(), p)


def delaunay_mesh(mf, orc, npts, seed, op):
    """An unstructured tet mesh: the Delaunay tetrahedralization of random points in the
    unit cube (nearly flat tets dropped), hull nodes with random boundary codes, random
    coefficients. Valences and colors vary from node to node, unlike the Kuhn cubes."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    pts = rng.random((npts, 3))
    tri = Delaunay(pts)
    tets = tri.simplices.astype(np.int64)
    a, b, c = (pts[tets[:, k]] - pts[tets[:, 3]] for k in range(3))
    tets = tets[np.abs(np.einsum('ij,ij->i', a, np.cross(b, c))) > 1e-9]
    i, j = tets[:, [0, 0, 0, 1, 1, 2]].ravel(), tets[:, [1, 2, 3, 2, 3, 3]].ravel()
    edges = np.unique(np.minimum(i, j) * npts + np.maximum(i, j))
    nnz = npts + 2 * edges.size
    code = np.zeros(npts, np.int32)
    hull = np.unique(tri.convex_hull)
    code[hull] = rng.choice([10, 52, 53, 54], hull.size)
    coord = pts.ravel().copy()
    e2n = (tets + 1).astype(np.int32).ravel()
    coef = mf.generate_coef(op, tets.shape[0], seed + 1)
    valence = int(np.bincount(tets.ravel()).max())
    return (mf.Mesh(coord, e2n, code, nnz, op, coef=coef),
            orc.OracleMesh(coord, e2n, code, nnz, op, coef=coef), valence)


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("npts,seed", [(3000, 5), (20000, 6)])
def test_delaunay_meshes_bit_exact(mf, orc, gpu, op, npts, seed):
    """Unstructured meshes, bit for bit against the oracle: the fused plan in FP64 and
    mixed precision, and the drop-in colored path."""
    mesh, om, valence = delaunay_mesh(mf, orc, npts, seed, op)
    print(f"[delaunay] op={op} points={npts} seed={seed} tets={mesh.nbElem} nnz={mesh.nnz} "
          f"max valence {valence}", flush=True)
    ov = om.assemble()
    op_ = om.prec(ov)[0]
    plan = gpu.Plan(mesh)
    dm = gpu.DeviceMesh(mesh)
    dm.values.fill(np.nan)
    dm.prec.fill(np.nan)
    plan.assemble(dm.coord, dm.coef, dm.values, dm.prec)
    assert np.array_equal(dm.values.numpy(), ov)
    assert same_prec(dm.prec.numpy(), op_)
    dm.values.fill(np.nan)
    dm.prec.fill(np.nan)
    plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=mf.PREC_MIXED)
    ovm = om.assemble(mixed=True)
    assert np.array_equal(dm.values.numpy(), ovm)
    assert same_prec(dm.prec.numpy(), om.prec(ovm)[0])
    dm.iterate_colored()
    assert np.array_equal(dm.values.numpy(), ov)
    assert same_prec(dm.prec.numpy(), op_)
