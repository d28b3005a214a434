"""The batched graph and the serving loop on the GPU.

Batched: the captured batch graph, replayed several times with the program's own eager
work and synchronisations in between, reproduces the eager per-mesh results bit for bit
on every replay; the union form (one launch per operator) equals the per-mesh form bit
for bit; one member against the oracle; mixed precision within 1e-6 of FP64.

Serving: every request's (values, prec) served in the mix -- overlapping requests on
several streams under an HBM budget that forces admission to wait -- equals the same
request served alone, bit for bit; one served request against the oracle with that
request's input; the mixed-precision server within 1e-6 of FP64.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def small_batch(mf, count=8):
    from minifem.batched import batch_meshes
    return batch_meshes(count=count, n_lo=3000, n_hi=30000, seed=1000)


def test_batched_graph_replays_equal_eager(mf, orc, gpu):
    from minifem.batched import Batch
    print("[batched] graph vs eager", flush=True)
    meshes = small_batch(mf)
    b = Batch(meshes, branches=4)
    b.run_eager()
    eager = b.results()
    b.capture()
    for rep in range(3):
        for dm in b.dms:  # poison: every replay must write every block again
            dm.values.fill(np.nan)
            dm.prec.fill(np.nan)
        b.launch()
        b.stream.sync()
        got = b.results()
        for (v0, p0), (v1, p1) in zip(eager, got):
            assert np.array_equal(v0, v1) and np.array_equal(p0, p1), f"replay {rep}"
        b.run_eager()  # eager work between replays
        gpu.synchronize()
    # one member against the oracle
    m = meshes[1]
    om = orc.OracleMesh(m.coord, m.elemToNode_original, m.boundNodesCode, m.nnz, m.operatorID,
                        coef=m.coef_original)
    ov = om.assemble()
    assert np.array_equal(eager[1][0], ov)
    assert np.array_equal(eager[1][1], om.prec(ov)[0])


@pytest.mark.parametrize("precision", [0, 1])
def test_batched_union_equals_per_mesh(mf, gpu, precision):
    """The union form (one persistent launch per operator) against the per-mesh form:
    every mesh's values and prec bit for bit, on every graph replay."""
    from minifem.batched import Batch
    print(f"[batched] union vs per mesh, precision {precision}", flush=True)
    meshes = small_batch(mf)
    ref = Batch(meshes, precision=precision, branches=4)
    ref.run_eager()
    want = ref.results()
    del ref
    for params in (dict(), dict(tile_lds_bytes=150000, block_threads=1024, max_tile_rows=254)):
        u = Batch(meshes, precision=precision, union=True, **params)
        assert u.launches() == 1
        u.capture()
        for rep in range(2):
            for dm in u.dms:
                dm.values.fill(np.nan)
                dm.prec.fill(np.nan)
            u.launch()
            u.stream.sync()
            for i, ((v0, p0), (v1, p1)) in enumerate(zip(want, u.results())):
                assert np.array_equal(v0, v1) and np.array_equal(p0, p1), (params, rep, i)
            u.run_eager()
            gpu.synchronize()
        del u


@pytest.mark.parametrize("op", [0, 1])
def test_batched_union_with_isolated_node(mf, gpu, op):
    """A union whose part has an unused node (an empty CSR row, prec of a node without
    elements) equals the per-mesh form bit for bit."""
    from minifem.batched import Batch
    print(f"[batched] union with an isolated node, op {op}", flush=True)
    iso = mf.Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [5, 5, 5]],
                           np.float64).ravel(),
                  np.array([1, 2, 3, 4], np.int32), np.array([0, 0, 0, 0, 53], np.int32), 16, op)
    meshes = [mf.Mesh.cube(5, 0.2, 11, op), iso, mf.Mesh.cube(7, 0.2, 12, op, shuffle_seed=5)]
    ref = Batch(meshes, branches=2)
    ref.run_eager()
    want = ref.results()
    u = Batch(meshes, union=True).capture()
    u.launch()
    u.stream.sync()
    for i, ((v0, p0), (v1, p1)) in enumerate(zip(want, u.results())):
        assert np.array_equal(v0, v1) and np.array_equal(p0, p1), i


def test_batched_mixed_precision_error(mf, gpu):
    from minifem.batched import Batch
    print("[batched] mixed", flush=True)
    meshes = small_batch(mf, count=4)
    b64 = Batch(meshes, branches=2)
    b64.run_eager()
    bmx = Batch(meshes, precision=mf.PREC_MIXED, branches=2).capture()
    bmx.launch()
    bmx.stream.sync()
    for (v0, p0), (v1, _) in zip(b64.results(), bmx.results()):
        rel = np.linalg.norm(v1 - v0) / np.linalg.norm(v0)
        assert rel <= 1e-6, rel


def test_serving_mixed_mix_equals_alone_and_error(mf, orc, gpu):
    """The mixed-precision server (coefficient fields uploaded as float32): every result
    the same alone and in the mix, and within 1e-6 of FP64."""
    from minifem.serve import Server, make_requests, served_meshes
    print("[serve mixed] mix vs alone", flush=True)
    meshes = served_meshes(count=4, n_lo=1000, n_hi=30000, seed=1000)
    reqs = make_requests(meshes, count=24, rate_hz=20000.0, seed=5000)
    fp = max(8 * m.dim * (m.nnz + m.nbNodes) + 8 * max(m.coord.size, m.coef.size)
             for m in meshes)
    ref = Server(meshes, hbm_budget=int(2.2 * fp), slots=1)  # plain FP64
    ref.stage(reqs)
    v64 = {r.rid: ref.run_alone(r)[0] for r in reqs}
    del ref
    srv = Server(meshes, hbm_budget=int(2.2 * fp), slots=4, precision=mf.PREC_MIXED)
    srv.stage(reqs)
    assert any(r.host.dtype == np.float32 for r in reqs)
    _, mixed = srv.run(reqs, keep_results=True)
    worst = 0.0
    for r in reqs:
        v, p = srv.run_alone(r)
        assert np.array_equal(v, mixed[r.rid][0]), r.rid
        assert np.array_equal(p, mixed[r.rid][1]), r.rid
        worst = max(worst, np.linalg.norm(v - v64[r.rid]) / np.linalg.norm(v64[r.rid]))
    print(f"served mixed vs fp64, worst relative Frobenius error: {worst:.3e}")
    assert worst <= 1e-6


@pytest.mark.parametrize("max_group", [1, 2, 8])
def test_serving_mix_equals_alone(mf, orc, gpu, max_group):
    """Every result in the mix equals the request served alone, whether requests are served
    one per launch or merged, and two equal the oracle."""
    from minifem.serve import Server, make_requests, served_meshes
    print(f"[serve] mix vs alone, max_group {max_group}", flush=True)
    meshes = served_meshes(count=6, n_lo=1000, n_hi=60000, seed=1000)
    reqs = make_requests(meshes, count=40, rate_hz=20000.0, seed=5000)
    fp = max(8 * m.dim * (m.nnz + m.nbNodes) + 8 * max(m.coord.size, m.coef.size)
             for m in meshes)
    srv = Server(meshes, hbm_budget=int(2.2 * fp), slots=4, max_group=max_group)
    srv.stage(reqs)
    stats, mixed = srv.run(reqs, keep_results=True)
    assert stats["requests"] == len(reqs)
    assert 0 < stats["p50_ms"] <= stats["p99_ms"]
    assert srv.arena_used() == 0
    for r in reqs:
        v, p = srv.run_alone(r)
        assert np.array_equal(v, mixed[r.rid][0]), r.rid
        assert np.array_equal(p, mixed[r.rid][1]), r.rid
    # a coefficient request and a coordinate request against the oracle
    for r in reqs[:2]:
        m = meshes[r.mesh]
        coord = r.payload if r.kind == "coord" else m.coord
        coef = m.coef_original if r.kind == "coord" else m.unpermute_elements(r.payload)
        om = orc.OracleMesh(coord, m.elemToNode_original, m.boundNodesCode, m.nnz, m.operatorID,
                            coef=coef)
        ov = om.assemble()
        assert np.array_equal(mixed[r.rid][0], ov)
        assert np.array_equal(mixed[r.rid][1], om.prec(ov)[0])


def test_serving_rejects_request_above_budget(mf, gpu):
    from minifem.serve import Server, make_requests, served_meshes
    meshes = served_meshes(count=2, n_lo=2000, n_hi=3000, seed=1000)
    reqs = make_requests(meshes, count=2, rate_hz=1000.0)
    srv = Server(meshes, hbm_budget=4096, slots=1)
    srv.stage(reqs)
    with pytest.raises(ValueError):
        srv.run(reqs)


def test_serving_rejects_bad_requests(mf, gpu):
    """The native loop checks every request before it serves any: a payload of the wrong
    size (the kernel would read past it), an unknown mesh or kind, the float coefficient
    mode on coordinates; nothing is admitted and the arena stays empty."""
    import ctypes
    from minifem._lib import MF_ERR_ARG, REQ_COEF, REQ_COORD, Request as RequestDesc
    from minifem.serve import Server, make_requests, served_meshes
    meshes = served_meshes(count=2, n_lo=2000, n_hi=3000, seed=1000)
    srv = Server(meshes, hbm_budget=1 << 26, slots=2)
    reqs = make_requests(meshes, count=2, rate_hz=1000.0)
    srv.stage(reqs)
    r = reqs[0]
    lib = mf.lib()
    good = (r.mesh, REQ_COEF if r.kind == "coef" else REQ_COORD, 0, 0, r.host.ptr,
            r.host.nbytes, 0.0, None, None)
    bad = [good[:5] + (r.host.nbytes - 8,) + good[6:],            # short payload
           (5,) + good[1:],                                        # no such mesh
           good[:1] + (7,) + good[2:],                             # no such kind
           good[:1] + (REQ_COORD, 2) + good[3:]]                   # float coefficients on coords
    for b in bad:
        arr = (RequestDesc * 1)(RequestDesc(*b))
        ts, td = np.zeros(1), np.zeros(1)
        assert lib.mf_server_run(srv._srv, ctypes.cast(arr, ctypes.c_void_p), 1,
                                 ts.ctypes.data, td.ctypes.data, None) == MF_ERR_ARG, b
    assert srv.arena_used() == 0
    stats = srv.run(reqs)  # and the server still serves
    assert stats["requests"] == 2 and srv.arena_used() == 0


def test_serving_rejects_fp64_on_float_staged_plans(mf, gpu):
    """A mixed-precision server sizes its plans' tiles by the float stage, and such a plan
    does not launch in FP64: the native loop rejects an FP64 request before it serves
    any, the arena stays empty, and the mixed requests are still served."""
    import ctypes
    from minifem._lib import MF_ERR_ARG, PREC_FP64, PREC_MIXED, REQ_COEF, Request as RequestDesc
    from minifem.serve import Request, Server
    m = mf.Mesh.cube(10, 0.2, 3, mf.OP_ELA, coef_seed=9)
    srv = Server([m], hbm_budget=1 << 26, slots=2, precision=PREC_MIXED)
    assert srv.plans[0].stage_precision == PREC_MIXED
    reqs = [Request(0, 0.0, 0, "coef", m.coef), Request(1, 0.0, 0, "coef", m.coef)]
    f64 = gpu.HostArray(m.coef.size, np.float64)
    f64.array[:] = m.coef
    lib = mf.lib()
    arr = (RequestDesc * 1)(RequestDesc(0, REQ_COEF, PREC_FP64, 0, f64.ptr, f64.nbytes, 0.0,
                                        None, None))
    ts, td = np.zeros(1), np.zeros(1)
    assert lib.mf_server_run(srv._srv, ctypes.cast(arr, ctypes.c_void_p), 1, ts.ctypes.data,
                             td.ctypes.data, None) == MF_ERR_ARG
    assert srv.arena_used() == 0
    srv.stage(reqs)
    stats = srv.run(reqs)
    assert stats["requests"] == 2 and srv.arena_used() == 0


@pytest.mark.parametrize("max_bypass", [0, 3])
def test_serving_bypass_is_bounded(mf, gpu, max_bypass):
    """Admission policy of the native loop. A large request that fits the arena only when
    it is alone arrives second, with twenty small ones behind it, all at once. It waits
    for memory; exactly max_bypass small requests are admitted past it (first fit), then
    admission holds until it fits; every request is served and the arena ends empty."""
    from minifem.serve import Request, Server
    big = mf.Mesh.cube(31, 0.2, 1, mf.OP_ELA, coef_seed=5)
    small = mf.Mesh.cube(8, 0.2, 2, mf.OP_ELA, coef_seed=6)
    fp = lambda m: sum((n + 255) // 256 * 256 for n in (8 * m.coef.size, 72 * m.nnz, 72 * m.nbNodes))
    srv = Server([big, small], hbm_budget=fp(big) + fp(small) // 2, slots=4, max_group=1,
                 max_bypass=max_bypass)
    reqs = [Request(0, 0.0, 1, "coef", small.coef), Request(1, 0.0, 0, "coef", big.coef)]
    reqs += [Request(2 + i, 0.0, 1, "coef", small.coef) for i in range(20)]
    srv.stage(reqs)
    stats = srv.run(reqs)
    assert stats["requests"] == len(reqs) and srv.arena_used() == 0
    assert stats["memory_waits"] >= 1
    t_big = reqs[1].t_start
    passed = [r.rid for r in reqs[2:] if r.t_start < t_big]
    print(f"[serve bypass] max_bypass {max_bypass}: passed {passed}", flush=True)
    assert len(passed) == max_bypass
    assert all(r.t_done >= r.t_start >= 0 for r in reqs)


def test_batched_full_config(mf, orc, gpu):
    """The batched workload at its size: 64 jittered cubes of 5k-200k nodes, lap and ela
    alternating (batch_meshes() defaults). The per-mesh form (one launch per mesh) is the
    reference; its largest lap and largest ela member are bit-exact against the oracle;
    the union form equals it bit for bit on every graph replay (FP64), and its mixed
    form is within 1e-6 relative Frobenius error of FP64 for every mesh."""
    from minifem.batched import Batch, batch_meshes
    print("[batched full] 64 meshes, 5k-200k nodes", flush=True)
    meshes = batch_meshes()
    assert len(meshes) == 64 and max(m.nbNodes for m in meshes) > 150000
    per = Batch(meshes, branches=8)
    per.run_eager()
    want = per.results()
    del per
    for op in (mf.OP_LAP, mf.OP_ELA):
        i = max((k for k, m in enumerate(meshes) if m.operatorID == op),
                key=lambda k: meshes[k].nbNodes)
        m = meshes[i]
        print(f"[batched full] mesh {i} ({m.nbNodes} nodes, op {op}) vs the oracle", flush=True)
        om = orc.OracleMesh(m.coord, m.elemToNode_original, m.boundNodesCode, m.nnz, op,
                            coef=m.coef_original)
        ov = om.assemble()
        assert np.array_equal(want[i][0], ov)
        assert np.array_equal(want[i][1], om.prec(ov)[0])
        del om, ov
    for prec in (mf.PREC_FP64, mf.PREC_MIXED):
        # the mixed batch's tiles are sized by its float stage
        u = Batch(meshes, precision=prec, union=True).capture()
        assert u.launches() == 1  # both operators' unions in one persistent launch
        for rep in range(2):
            for dm in u.dms:
                dm.values.fill(np.nan)
                dm.prec.fill(np.nan)
            u.launch()
            u.stream.sync()
            got = u.results()
            for i, ((v0, p0), (v1, p1)) in enumerate(zip(want, got)):
                if prec == mf.PREC_FP64:
                    assert np.array_equal(v0, v1) and np.array_equal(p0, p1), (rep, i)
                else:
                    ev = np.linalg.norm(v1 - v0) / np.linalg.norm(v0)
                    ep = np.linalg.norm(p1 - p0) / np.linalg.norm(p0)
                    assert ev <= 1e-6 and ep <= 1e-6, (rep, i, ev, ep)
            del got
            u.run_eager()
            gpu.synchronize()
        del u


@pytest.mark.parametrize("precision", [0, 1])
def test_served_full_config(mf, orc, gpu, precision):
    """The served workload at its size: 16 resident meshes of 1k-885k nodes
    (served_meshes() defaults), a seeded Poisson stream of 32 requests at 20,000/s under
    an HBM budget below the sum of the footprints, so admission waits for memory. Every
    request's result in the mix equals the same request served alone, bit for bit; the
    first coefficient request and the first coordinate request are checked against the
    oracle (FP64) or within 1e-6 of the FP64 server (mixed)."""
    from minifem.serve import Server, make_requests, served_meshes
    print(f"[serve full] 16 meshes, precision {precision}", flush=True)
    meshes = served_meshes()
    assert len(meshes) == 16 and max(m.nbNodes for m in meshes) > 800000
    reqs = make_requests(meshes, count=32, rate_hz=20000.0, seed=5000)
    fps = [8 * m.dim * (m.nnz + m.nbNodes) + 8 * max(m.coord.size, m.coef.size) for m in meshes]
    budget = max(int(0.5 * sum(fps)), int(1.05 * max(fps)) + 4096)
    assert budget < sum(fps)
    srv = Server(meshes, budget, slots=4, precision=precision)
    srv.stage(reqs)
    stats, mixed = srv.run(reqs, keep_results=True)
    print(f"[serve full] {stats}", flush=True)
    assert stats["requests"] == len(reqs) and srv.arena_used() == 0
    assert stats["launches"] < len(reqs)  # requests that arrived together shared launches
    assert 0 < stats["p50_ms"] <= stats["p99_ms"]
    for r in reqs:
        v, p = srv.run_alone(r)
        assert np.array_equal(v, mixed[r.rid][0]), r.rid
        assert np.array_equal(p, mixed[r.rid][1]), r.rid
    sample = [next(r for r in reqs if r.kind == "coef"), next(r for r in reqs if r.kind == "coord")]
    if precision == mf.PREC_FP64:
        for r in sample:
            m = meshes[r.mesh]
            print(f"[serve full] request {r.rid} ({r.kind}, {m.nbNodes} nodes) vs the oracle",
                  flush=True)
            coord = r.payload if r.kind == "coord" else m.coord
            coef = m.coef_original if r.kind == "coord" else m.unpermute_elements(r.payload)
            om = orc.OracleMesh(coord, m.elemToNode_original, m.boundNodesCode, m.nnz,
                                m.operatorID, coef=coef)
            ov = om.assemble()
            assert np.array_equal(mixed[r.rid][0], ov)
            assert np.array_equal(mixed[r.rid][1], om.prec(ov)[0])
            del om, ov
    else:
        del srv
        ref = Server(meshes, budget, slots=1)
        ref.stage(sample)
        for r in sample:
            v64, p64 = ref.run_alone(r)
            ev = np.linalg.norm(mixed[r.rid][0] - v64) / np.linalg.norm(v64)
            ep = np.linalg.norm(mixed[r.rid][1] - p64) / np.linalg.norm(p64)
            assert ev <= 1e-6 and ep <= 1e-6, (r.rid, ev, ep)


def test_served_new_mesh_admitted_at_run_time(mf, orc, gpu):
    """A topology that was not resident joins a running server (Server.add_mesh): its
    requests, mixed with requests on the original meshes in one stream, equal the same
    requests served alone, and one of them equals the oracle."""
    from minifem.serve import Server, make_requests, served_meshes
    print("[serve] new mesh at run time", flush=True)
    meshes = served_meshes(count=4, n_lo=1000, n_hi=20000, seed=1000)
    fp = max(8 * m.dim * (m.nnz + m.nbNodes) + 8 * max(m.coord.size, m.coef.size)
             for m in meshes)
    srv = Server(meshes, hbm_budget=int(3.0 * fp), slots=3)
    new = mf.Mesh.cube(18, 0.25, 77, mf.OP_ELA, shuffle_seed=78, coef_seed=79)
    k = srv.add_mesh(new)
    assert k == 4 and len(srv.plans) == 5
    reqs = make_requests(srv.meshes, count=30, rate_hz=20000.0, seed=5001)
    assert any(r.mesh == k for r in reqs)
    srv.stage(reqs)
    stats, mixed = srv.run(reqs, keep_results=True)
    assert stats["requests"] == len(reqs)
    for r in reqs:
        v, p = srv.run_alone(r)
        assert np.array_equal(v, mixed[r.rid][0]) and np.array_equal(p, mixed[r.rid][1]), r.rid
    r = next(r for r in reqs if r.mesh == k)
    coord = r.payload if r.kind == "coord" else new.coord
    coef = new.coef_original if r.kind == "coord" else new.unpermute_elements(r.payload)
    om = orc.OracleMesh(coord, new.elemToNode_original, new.boundNodesCode, new.nnz, new.operatorID,
                        coef=coef)
    ov = om.assemble()
    assert np.array_equal(mixed[r.rid][0], ov) and np.array_equal(mixed[r.rid][1], om.prec(ov)[0])
