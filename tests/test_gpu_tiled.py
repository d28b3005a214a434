"""GPU parity of the fused tiled plan (mf_plan_create / mf_plan_assemble) against the
CPU oracle, through the C ABI.

Bar: values and prec bit-exact against the oracle's COLORING-order restatement for
FP64; mixed precision bit-exact against the oracle's mixed restatement and within
1e-6 relative Frobenius error of FP64. The benchmark workload itself (ela 100^3) is
checked bit for bit against the oracle; at 100^3 size-independent properties are checked
as well (exact block symmetry, vanishing row sums, prec times the masked diagonal block
= I, agreement with the drop-in colored path bit for bit).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def build(mf, orc, n, op, jitter=0.2, seed=1, shuffle=0, coef_seed=None, oracle=True):
    coord, e2n, code, nnz = mf.generate_cube(n, jitter, seed, shuffle)
    coef = None if coef_seed is None else mf.generate_coef(op, e2n.size // 4, coef_seed)
    mesh = mf.Mesh(coord, e2n, code, nnz, op, coef=coef)
    om = orc.OracleMesh(coord, e2n, code, nnz, op, coef=coef) if oracle else None
    return mesh, om


def run_plan(gpu, mesh, precision=0, **params):
    """One fused iteration into outputs poisoned with NaN, so a block the plan leaves
    unwritten cannot pass for a match."""
    plan = gpu.Plan(mesh, **params)
    dm = gpu.DeviceMesh(mesh)
    dm.values.fill(np.nan)
    dm.prec.fill(np.nan)
    plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=precision)
    return plan, dm


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("coef_seed", [None, 101])
def test_tiled_bit_exact_30_cube(mf, orc, gpu, op, coef_seed):
    print(f"[tiled] op={op} coef={coef_seed}", flush=True)
    mesh, om = build(mf, orc, 30, op, seed=1 + op, coef_seed=coef_seed)
    plan, dm = run_plan(gpu, mesh)
    ov = om.assemble()
    assert np.array_equal(dm.values.numpy(), ov)
    assert np.array_equal(dm.prec.numpy(), om.prec(ov)[0])


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("params", [
    dict(block_threads=512), dict(block_threads=1024),
    dict(tile_lds_bytes=8192),
    dict(tile_lds_bytes=150 * 1024, max_tile_rows=254), dict(use_coord=False),
    dict(max_tile_rows=1), dict(tile_lds_bytes=8192, block_threads=1024),
    dict(tile_lds_bytes=150 * 1024, block_threads=1024, max_tile_rows=254),
    dict(id_chunks=8), dict(id_chunks=5, block_threads=512),
])
def test_tiled_schedule_variants(mf, orc, gpu, op, params):
    print(f"[tiled variant] op={op} {params}", flush=True)
    mesh, om = build(mf, orc, 12, op, jitter=0.25, seed=3, shuffle=203, coef_seed=9)
    plan, dm = run_plan(gpu, mesh, **params)
    ov = om.assemble()
    assert np.array_equal(dm.values.numpy(), ov)
    assert np.array_equal(dm.prec.numpy(), om.prec(ov)[0])


def fan_mesh(mf, orc, n_fan, op, coef_seed=5):
    """n_fan tets around the edge between two hub nodes: the hubs have valence n_fan."""
    ang = np.linspace(0, 2 * np.pi, n_fan + 1)[:-1]
    ring = np.stack([np.cos(ang), np.sin(ang), 0.5 + 0.1 * np.sin(3 * ang)], 1)
    coord = np.concatenate([[[0, 0, 0], [0.05, -0.02, 1]], ring]).astype(np.float64).ravel()
    e2n = []
    for i in range(n_fan):
        e2n += [1, 2, 3 + i, 3 + (i + 1) % n_fan]
    e2n = np.array(e2n, np.int32)
    code = np.zeros(n_fan + 2, np.int32)
    code[5] = 52
    nnz = 2 * (n_fan + 2) + 5 * n_fan
    coef = mf.generate_coef(op, n_fan, coef_seed)
    return (mf.Mesh(coord, e2n, code, nnz, op, coef=coef),
            orc.OracleMesh(coord, e2n, code, nnz, op, coef=coef))


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("n_fan", [24, 40, 64, 65, 96, 127])
def test_tiled_high_valence_masks(mf, orc, gpu, op, n_fan):
    """Rows of valence 24 to 127: tasks with more than six contributions take the
    record's 64-bit mask over the row's incidence list up to valence 64, and walk the
    row's whole list beyond; bit-exact. 127 tets at one edge need 127 colors, the most a
    node can have in a mesh that colors within 128."""
    print(f"[tiled fan] op={op} n_fan={n_fan}", flush=True)
    mesh, om = fan_mesh(mf, orc, n_fan, op)
    plan, dm = run_plan(gpu, mesh)
    ov = om.assemble()
    assert np.array_equal(dm.values.numpy(), ov)
    assert np.array_equal(dm.prec.numpy(), om.prec(ov)[0])


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("stage", [0, 1])
def test_tiled_mixed(mf, orc, gpu, op, stage):
    """Mixed precision on a plan sized for the FP64 stage and on one sized for the float
    stage (larger tiles): both bit-exact against the oracle's mixed restatement."""
    print(f"[tiled mixed] op={op} stage={stage}", flush=True)
    mesh, om = build(mf, orc, 20, op, coef_seed=55)
    plan, dm = run_plan(gpu, mesh, precision=mf.PREC_MIXED, stage_precision=stage)
    vals = dm.values.numpy()
    assert np.array_equal(vals, om.assemble(mixed=True))
    ov = om.assemble()
    rel = np.linalg.norm(vals - ov) / np.linalg.norm(ov)
    print(f"tiled mixed vs fp64 relative Frobenius error: {rel:.3e}")
    assert rel <= 1e-6
    assert np.array_equal(dm.prec.numpy(), om.prec(vals)[0])


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("stage", [0, 2])
def test_tiled_mixed_f32_coefficients(mf, orc, gpu, op, stage):
    """MF_PREC_MIXED_F32COEF takes the coefficient field as float: bit-exact against the
    oracle's mixed restatement on the coefficients rounded to float, and within 1e-6 of
    FP64 on the original ones; on plans sized for the FP64 and for the float stage."""
    print(f"[tiled mixed f32 coef] op={op} stage={stage}", flush=True)
    coord, e2n, code, nnz = mf.generate_cube(20, 0.2, 1, 0)
    coef = mf.generate_coef(op, e2n.size // 4, 55)
    mesh = mf.Mesh(coord, e2n, code, nnz, op, coef=coef)
    plan = gpu.Plan(mesh, stage_precision=stage)
    dm = gpu.DeviceMesh(mesh)
    c32 = gpu.DeviceArray.from_numpy(mesh.coef.astype(np.float32))
    for _ in range(2):  # the second run must repeat the first bit for bit
        dm.values.fill(np.nan)
        dm.prec.fill(np.nan)
        plan.assemble(dm.coord, c32, dm.values, dm.prec, precision=mf.PREC_MIXED_F32COEF)
        vals, prec = dm.values.numpy(), dm.prec.numpy()
        om32 = orc.OracleMesh(coord, e2n, code, nnz, op,
                              coef=coef.astype(np.float32).astype(np.float64))
        assert np.array_equal(vals, om32.assemble(mixed=True))
        assert np.array_equal(prec, om32.prec(vals)[0])
    ov = orc.OracleMesh(coord, e2n, code, nnz, op, coef=coef).assemble()
    rel = np.linalg.norm(vals - ov) / np.linalg.norm(ov)
    print(f"mixed, float coefficients vs fp64 relative Frobenius error: {rel:.3e}")
    assert rel <= 1e-6


@pytest.mark.parametrize("op", [0, 1])
def test_tiled_tiny_and_isolated(mf, orc, gpu, op):
    # single tet, two tets, and a mesh with an unused node (empty CSR row)
    cases = [
        (np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0]], np.float64).ravel(),
         np.array([1, 2, 3, 4], np.int32), np.zeros(4, np.int32), 16),
        (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], np.float64).ravel(),
         np.array([1, 2, 3, 4, 2, 3, 4, 5], np.int32), np.array([52, 53, 54, 10, 0], np.int32), 23),
        (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [5, 5, 5]], np.float64).ravel(),
         np.array([1, 2, 3, 4], np.int32), np.array([0, 0, 0, 0, 53], np.int32), 16),
    ]
    for coord, e2n, code, nnz in cases:
        mesh = mf.Mesh(coord, e2n, code, nnz, op)
        om = orc.OracleMesh(coord, e2n, code, nnz, op)
        plan, dm = run_plan(gpu, mesh)
        ov = om.assemble()
        assert np.array_equal(dm.values.numpy(), ov)
        p_gpu, p_ref = dm.prec.numpy(), om.prec(ov)[0]
        assert np.array_equal(p_gpu, p_ref) or (
            np.array_equal(np.isinf(p_gpu), np.isinf(p_ref)) and
            np.array_equal(p_gpu[~np.isinf(p_gpu)], p_ref[~np.isinf(p_ref)]))


def test_tiled_equals_colored_and_deterministic(mf, orc, gpu):
    mesh, _ = build(mf, orc, 24, 1, jitter=0.25, seed=3, shuffle=203, coef_seed=103, oracle=False)
    plan, dm = run_plan(gpu, mesh)
    v1, p1 = dm.values.numpy(), dm.prec.numpy()
    dm.values.fill(np.nan)
    dm.prec.fill(np.nan)
    plan.assemble(dm.coord, dm.coef, dm.values, dm.prec)
    assert np.array_equal(v1, dm.values.numpy()) and np.array_equal(p1, dm.prec.numpy())
    dm.iterate_colored()
    assert np.array_equal(v1, dm.values.numpy()) and np.array_equal(p1, dm.prec.numpy())


@pytest.mark.parametrize("op", [0, 1])
def test_tiled_large_properties(mf, orc, gpu, op):
    """100^3 shuffled jittered cube (the benchmark workload): properties that do not need
    the oracle at full size."""
    print(f"[tiled 100^3] op={op}", flush=True)
    mesh, _ = build(mf, orc, 100, op, jitter=0.25, seed=3, shuffle=203, coef_seed=103,
                    oracle=False)
    plan, dm = run_plan(gpu, mesh)
    vals, prec = dm.values.numpy(), dm.prec.numpy()
    assert np.all(np.isfinite(vals)) and np.all(np.isfinite(prec))
    d = mesh.dim
    b = 3 if op else 1
    n = mesh.nbNodes
    rows = np.repeat(np.arange(n), np.diff(mesh.row))
    cols = mesh.col - 1
    V = vals.reshape(-1, b, b)
    # exact symmetry: block(i,j) == block(j,i)^T bit for bit
    key = rows.astype(np.int64) * n + cols
    tkey = cols.astype(np.int64) * n + rows
    order = np.argsort(key)
    pos_t = order[np.searchsorted(key[order], tkey)]
    assert np.array_equal(V, V[pos_t].transpose(0, 2, 1))
    # row sums of every block row vanish (constants are in the kernel of the operator)
    rs = np.zeros((n, b, b))
    np.add.at(rs, rows, V)
    scale = np.abs(V).max()
    assert np.abs(rs).max() < 1e-9 * scale
    # prec inverts the (masked) diagonal block
    diag = np.nonzero(rows == cols)[0]
    assert diag.size == n
    Dg = V[diag]
    P = prec.reshape(n, b, b)
    if op == 0:
        assert np.array_equal(P[:, 0, 0], 1.0 / Dg[:, 0, 0])
    else:
        cb = mesh.checkBounds.reshape(3, n).T.astype(bool)
        A = Dg.transpose(0, 2, 1).copy()  # Fortran view of each block
        for k in range(3):
            A[cb[:, k], k, :] = 0
            A[cb[:, k], :, k] = 0
            A[cb[:, k], k, k] = 1
        Pf = P.transpose(0, 2, 1)
        err = np.abs(np.einsum("nij,njk->nik", Pf, A) - np.eye(3)).max()
        assert err < 1e-10, err
    # the drop-in colored path gives the same bits
    dm.iterate_colored()
    assert np.array_equal(vals, dm.values.numpy())
    assert np.array_equal(prec, dm.prec.numpy())


def test_tiled_200_cube_equals_colored(mf, orc, gpu):
    """The largest mesh tested: an 8M-node ela cube (200^3, node-shuffled, 46.6M tets,
    118M CSR blocks, 8.5 GB of values: byte offsets past 2^32). The fused plan and the
    drop-in colored path, two independent kernels, give the same bits for values and
    prec, all finite."""
    print("[tiled 200^3] ela", flush=True)
    mesh, _ = build(mf, orc, 200, 1, jitter=0.25, seed=3, shuffle=209, coef_seed=109,
                    oracle=False)
    assert mesh.nnz * mesh.dim * 8 > 2 ** 32
    plan, dm = run_plan(gpu, mesh)
    vals, prec = dm.values.numpy(), dm.prec.numpy()
    assert np.all(np.isfinite(vals)) and np.all(np.isfinite(prec))
    del plan
    dm.iterate_colored()
    assert np.array_equal(vals, dm.values.numpy())
    assert np.array_equal(prec, dm.prec.numpy())


@pytest.mark.parametrize("op", [0, 1])
def test_write_probe_covers_every_value(mf, orc, gpu, op):
    """mf_plan_write_probe (measurement only): mode 0 (tile rows), mode 1 (front to back)
    and modes 2 + 3 (whole lines + partial end lines) each overwrite every value exactly
    where the assembly writes, and nothing past the array."""
    print(f"[write probe] op={op}", flush=True)
    mesh, _ = build(mf, orc, 14, op, seed=5, shuffle=207, oracle=False)
    plan = gpu.Plan(mesh)
    n = mesh.nnz * mesh.dim
    buf = gpu.DeviceArray(n + 64, np.float64)  # a guard band past the values
    for modes in ((0,), (1,), (2, 3)):
        buf.fill(np.nan)
        for m in modes:
            plan.write_probe(buf, m)
        out = buf.numpy()
        assert np.all(out[:n] == 0.0), modes
        assert np.all(np.isnan(out[n:])), modes
    if op == 1:  # ela rows (~1 KB) cover whole lines; lap rows (~120 B) rarely do
        buf.fill(np.nan)
        plan.write_probe(buf, 2)
        part = buf.numpy()[:n]
        assert np.any(np.isnan(part)) and np.any(part == 0.0)


def test_tiled_bit_exact_bench_workload(mf, orc, gpu):
    """The benchmark workload itself (ela 100^3, jitter 0.25h, seeds 3/103/203, node ids
    shuffled) under the benchmark's schedule (155000 B tiles, 1024 threads, 254 rows):
    FP64 values and prec bit-exact against the oracle, and the mixed form bit-exact against
    the oracle's mixed restatement."""
    print("[tiled bench workload] ela 100^3", flush=True)
    coord, e2n, code, nnz = mf.generate_cube(100, 0.25, 3, 203)
    coef = mf.generate_coef(1, e2n.size // 4, 103)
    mesh = mf.Mesh(coord, e2n, code, nnz, 1, coef=coef)
    plan, dm = run_plan(gpu, mesh, tile_lds_bytes=155000, block_threads=1024, max_tile_rows=254)
    vals, prec = dm.values.numpy(), dm.prec.numpy()
    om = orc.OracleMesh(coord, e2n, code, nnz, 1, coef=coef)
    ov = om.assemble()
    assert np.array_equal(vals, ov)
    assert np.array_equal(prec, om.prec(ov)[0])
    del ov, prec
    dm.values.fill(np.nan)
    plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=mf.PREC_MIXED)
    mixed = om.assemble(mixed=True)
    assert np.array_equal(dm.values.numpy(), mixed)
    # the mixed benchmark's plan: tiles sized by the float stage, fewer and larger
    p32 = gpu.Plan(mesh, tile_lds_bytes=155000, block_threads=1024, max_tile_rows=254,
                   stage_precision=mf.PREC_MIXED)
    assert p32.stats()["tiles"] < plan.stats()["tiles"]
    dm.values.fill(np.nan)
    dm.prec.fill(np.nan)
    p32.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=mf.PREC_MIXED)
    vals = dm.values.numpy()
    assert np.array_equal(vals, mixed)
    assert np.array_equal(dm.prec.numpy(), om.prec(vals)[0])


@pytest.mark.parametrize("op", [0, 1])
def test_float_stage_plan_runs_mixed_modes_only(mf, orc, gpu, op):
    """A plan sized for the float stage: no more tiles than the FP64-sized plan of the
    same mesh (fewer for ela; lap tiles at the default budget are bound by the row cap),
    both mixed modes bit-identical to that plan's, and FP64 launches of it rejected with
    MF_ERR_ARG (alone and in a multi-plan launch), leaving the outputs untouched."""
    from minifem._lib import MiniFemError
    print(f"[float stage] op={op}", flush=True)
    mesh, _ = build(mf, orc, 40, op, jitter=0.25, seed=4, shuffle=11, coef_seed=8, oracle=False)
    p64 = gpu.Plan(mesh)
    p32 = gpu.Plan(mesh, stage_precision=mf.PREC_MIXED)
    assert p32.stage_precision == mf.PREC_MIXED
    t32, t64 = p32.stats()["tiles"], p64.stats()["tiles"]
    assert t32 < t64 if op == 1 else t32 <= t64
    dm = gpu.DeviceMesh(mesh)
    c32 = gpu.DeviceArray.from_numpy(mesh.coef.astype(np.float32))
    want = {}
    for prec, coef in ((mf.PREC_MIXED, dm.coef), (mf.PREC_MIXED_F32COEF, c32)):
        got = []
        for plan in (p64, p32):
            dm.values.fill(np.nan)
            dm.prec.fill(np.nan)
            plan.assemble(dm.coord, coef, dm.values, dm.prec, precision=prec)
            got.append((dm.values.numpy(), dm.prec.numpy()))
        assert np.array_equal(got[0][0], got[1][0]) and np.array_equal(got[0][1], got[1][1])
        want[prec] = got[0]
    # both plans in one mixed multi-plan launch, each into outputs of its own
    dm2 = gpu.DeviceMesh(mesh)
    for d in (dm, dm2):
        d.values.fill(np.nan)
        d.prec.fill(np.nan)
    gpu.plans_assemble([(p64, dm.coord, dm.coef, dm.values, dm.prec),
                        (p32, dm2.coord, dm2.coef, dm2.values, dm2.prec)], mf.PREC_MIXED)
    gpu.synchronize()
    for d in (dm, dm2):
        assert np.array_equal(d.values.numpy(), want[mf.PREC_MIXED][0])
        assert np.array_equal(d.prec.numpy(), want[mf.PREC_MIXED][1])
    dm.values.fill(7.0)
    with pytest.raises(MiniFemError) as exc:
        p32.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=mf.PREC_FP64)
    assert exc.value.status == -1
    jobs = [(p64, dm.coord, dm.coef, dm.values, dm.prec), (p32, dm.coord, dm.coef, dm.values, dm.prec)]
    with pytest.raises(MiniFemError):
        gpu.plans_assemble(jobs, mf.PREC_FP64)
    assert gpu.plans_lds_bytes(jobs, mf.PREC_FP64) is None
    gpu.synchronize()
    assert np.all(dm.values.numpy() == 7.0)


def test_plan_launched_from_two_host_threads(mf, orc, gpu):
    """One plan launched from two host threads at once, each on its own stream with its
    own outputs: every result equals a single launch bit for bit."""
    import threading
    print("[tiled] two host threads", flush=True)
    mesh, _ = build(mf, orc, 20, 1, jitter=0.25, seed=3, shuffle=7, coef_seed=5, oracle=False)
    plan, dm = run_plan(gpu, mesh)
    want_v, want_p = dm.values.numpy(), dm.prec.numpy()
    outs = [(gpu.DeviceArray(mesh.nnz * 9, np.float64), gpu.DeviceArray(mesh.nbNodes * 9, np.float64))
            for _ in range(2)]
    for v, p in outs:  # a block a launch leaves unwritten cannot pass for a match
        v.fill(np.nan)
        p.fill(np.nan)
    errors = []

    def work(i):
        try:
            gpu.init(0)
            s = gpu.Stream()
            v, p = outs[i]
            for _ in range(5):
                plan.assemble(dm.coord, dm.coef, v, p, stream=s)
            s.sync()
        except Exception as exc:  # reported by the main thread
            errors.append(exc)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for v, p in outs:
        assert np.array_equal(v.numpy(), want_v) and np.array_equal(p.numpy(), want_p)
