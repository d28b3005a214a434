"""Several plans in ONE persistent launch (mf_plans_assemble), through the C ABI.

Every job's values and prec must equal its own single-plan launch (mf_plan_assemble)
bit for bit, and so the oracle: lap and ela jobs mixed in one launch, meshes of
different sizes and shuffles, a row of valence 96 (walk records), a node outside every
element, the same plan twice with different inputs, float coefficients, and graph
replays. Outputs are poisoned with NaN first, so a block a launch leaves unwritten
cannot pass for a match.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fan(mf, n_fan, op, coef_seed):
    ang = np.linspace(0, 2 * np.pi, n_fan + 1)[:-1]
    ring = np.stack([np.cos(ang), np.sin(ang), 0.5 + 0.1 * np.sin(3 * ang)], 1)
    coord = np.concatenate([[[0, 0, 0], [0.05, -0.02, 1]], ring]).astype(np.float64).ravel()
    e2n = np.array([[1, 2, 3 + i, 3 + (i + 1) % n_fan] for i in range(n_fan)], np.int32).ravel()
    code = np.zeros(n_fan + 2, np.int32)
    code[4] = 53
    coef = mf.generate_coef(op, n_fan, coef_seed)
    return mf.Mesh(coord, e2n, code, 2 * (n_fan + 2) + 5 * n_fan, op, coef=coef)


def isolated(mf, op):
    """One tet plus a node no element holds (an empty CSR row)."""
    coord = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [5, 5, 5]], np.float64).ravel()
    return mf.Mesh(coord, np.array([1, 2, 3, 4], np.int32), np.array([0, 0, 0, 0, 53], np.int32),
                   16, op, coef=mf.generate_coef(op, 1, 77))


def mixed_meshes(mf):
    L, E = mf.OP_LAP, mf.OP_ELA
    return [mf.Mesh.cube(10, 0.2, 1, L, shuffle_seed=3, coef_seed=11),
            mf.Mesh.cube(12, 0.25, 2, E, shuffle_seed=5, coef_seed=12),
            fan(mf, 96, E, 13), fan(mf, 70, L, 14), isolated(mf, E),
            mf.Mesh.cube(5, 0.2, 4, E, coef_seed=15), mf.Mesh.cube(16, 0.2, 6, L, coef_seed=16)]


def fresh(gpu, m):
    v = gpu.DeviceArray(m.nnz * m.dim, np.float64)
    p = gpu.DeviceArray(m.nbNodes * m.dim, np.float64)
    v.fill(np.nan)
    p.fill(np.nan)
    return v, p


def alone(gpu, plan, dm, coef, precision):
    v, p = fresh(gpu, dm.mesh)
    plan.assemble(dm.coord, coef, v, p, precision=precision)
    gpu.synchronize()
    return v.numpy(), p.numpy()


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_jobs_equal_single_plan_launches(mf, orc, gpu, precision):
    print(f"[jobs] mixed lap/ela, precision {precision}", flush=True)
    meshes = mixed_meshes(mf)
    dms = [gpu.DeviceMesh(m) for m in meshes]
    plans = [gpu.Plan(m) for m in meshes]
    coefs = [dm.coef if precision != mf.PREC_MIXED_F32COEF else
             gpu.DeviceArray.from_numpy(m.coef.astype(np.float32)) for m, dm in zip(meshes, dms)]
    want = [alone(gpu, p, dm, c, precision) for p, dm, c in zip(plans, dms, coefs)]
    outs = [fresh(gpu, m) for m in meshes]
    jobs = [(p, dm.coord, c, v, pr) for p, dm, c, (v, pr) in zip(plans, dms, coefs, outs)]
    assert gpu.plans_lds_bytes(jobs, precision) is not None
    gpu.plans_assemble(jobs, precision)
    gpu.synchronize()
    for i, ((v, p), (wv, wp)) in enumerate(zip(outs, want)):
        assert np.array_equal(v.numpy(), wv), i
        assert np.array_equal(p.numpy(), wp), i
    if precision == mf.PREC_FP64:  # and the oracle, for the shuffled ela cube and the fan
        for i in (1, 2):
            m = meshes[i]
            om = orc.OracleMesh(m.coord, m.elemToNode_original, m.boundNodesCode, m.nnz,
                                m.operatorID, coef=m.coef_original)
            ov = om.assemble()
            assert np.array_equal(want[i][0], ov) and np.array_equal(want[i][1], om.prec(ov)[0])


def test_jobs_same_plan_twice(mf, gpu):
    """One plan in two jobs with different coefficient fields and outputs (two requests
    on one resident mesh): each equals its own launch."""
    m = mf.Mesh.cube(14, 0.25, 3, mf.OP_ELA, shuffle_seed=9, coef_seed=21)
    dm, plan = gpu.DeviceMesh(m), gpu.Plan(m)
    c2 = gpu.DeviceArray.from_numpy(m.permute_elements(mf.generate_coef(mf.OP_ELA, m.nbElem, 22)))
    want = [alone(gpu, plan, dm, c, 0) for c in (dm.coef, c2)]
    outs = [fresh(gpu, m) for _ in range(2)]
    gpu.plans_assemble([(plan, dm.coord, c, v, p) for c, (v, p) in zip((dm.coef, c2), outs)])
    gpu.synchronize()
    for (v, p), (wv, wp) in zip(outs, want):
        assert np.array_equal(v.numpy(), wv) and np.array_equal(p.numpy(), wp)
    assert not np.array_equal(want[0][0], want[1][0])


def test_jobs_graph_replays(mf, gpu):
    """A captured multi-plan launch, replayed with eager work and synchronisations between
    the replays: every replay writes every block again, equal to the eager result."""
    meshes = mixed_meshes(mf)[:4]
    dms = [gpu.DeviceMesh(m) for m in meshes]
    plans = [gpu.Plan(m) for m in meshes]
    jobs = [(p, dm.coord, dm.coef, dm.values, dm.prec) for p, dm in zip(plans, dms)]
    stream = gpu.Stream()
    gpu.plans_assemble(jobs, 0, stream)
    stream.sync()
    want = [(dm.values.numpy(), dm.prec.numpy()) for dm in dms]
    graph = gpu.Graph(stream, lambda s: gpu.plans_assemble(jobs, 0, s))
    for rep in range(3):
        for dm in dms:
            dm.values.fill(np.nan)
            dm.prec.fill(np.nan)
        graph.launch(stream)
        stream.sync()
        for i, dm in enumerate(dms):
            assert np.array_equal(dm.values.numpy(), want[i][0]), (rep, i)
            assert np.array_equal(dm.prec.numpy(), want[i][1]), (rep, i)
        gpu.plans_assemble(jobs, 0, stream)  # eager work between the replays
        gpu.synchronize()


def test_jobs_argument_errors(mf, gpu):
    from minifem._lib import MF_ERR_ARG, PlanJob
    lib = mf.lib()
    m = mf.Mesh.cube(4, 0.2, 1, mf.OP_LAP)
    dm = gpu.DeviceMesh(m)
    p1024, p512 = gpu.Plan(m, block_threads=1024), gpu.Plan(m, block_threads=512)
    job = lambda p, v=dm.values: PlanJob(p.ptr, dm.coord.ptr, None, v.ptr if v else None,
                                         dm.prec.ptr, None)
    arr = (PlanJob * 17)(*([job(p1024)] * 17))
    ptr = ctypes.cast(arr, ctypes.c_void_p)
    assert lib.mf_plans_assemble(ptr, 0, 0, None) == MF_ERR_ARG
    assert lib.mf_plans_assemble(ptr, 17, 0, None) == MF_ERR_ARG
    assert lib.mf_plans_assemble(ptr, 1, 3, None) == MF_ERR_ARG  # no such precision mode
    two = (PlanJob * 2)(job(p1024), job(p512))  # one workgroup shape per launch
    assert lib.mf_plans_assemble(ctypes.cast(two, ctypes.c_void_p), 2, 0, None) == MF_ERR_ARG
    bad = (PlanJob * 1)(job(p1024, None))
    assert lib.mf_plans_assemble(ctypes.cast(bad, ctypes.c_void_p), 1, 0, None) == MF_ERR_ARG
    with pytest.raises(ValueError):
        gpu.plans_assemble([])
