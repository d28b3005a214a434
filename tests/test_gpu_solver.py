"""The consumer of an iteration's outputs: block-Jacobi PCG (mf_pcg) on the masked
operator K = P A P + (I - P) assembled by the fused plan, against scipy on the same
matrix: x solves K x = K x_true; the residual reported is the true one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def masked_operator(mesh, values):
    import scipy.sparse as sp
    d = 3 if mesh.dim == 9 else 1
    n = mesh.nbNodes
    rows = np.repeat(np.arange(n), np.diff(mesh.row))
    cols = mesh.col - 1
    B = values.reshape(-1, d, d)
    I = (rows[:, None, None] * d + np.arange(d)[None, :, None]) + 0 * np.arange(d)[None, None, :]
    J = (cols[:, None, None] * d + np.arange(d)[None, None, :]) + 0 * np.arange(d)[None, :, None]
    A = sp.csr_matrix((B.ravel(), (I.ravel(), J.ravel())), shape=(n * d, n * d))
    cb = mesh.checkBounds.reshape(3, n)
    m = (cb.T != 0).ravel() if d == 3 else (cb != 0).any(axis=0)
    P = sp.diags((~m).astype(float))
    return (P @ A @ P + sp.diags(m.astype(float))).tocsr(), m


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("n", [10, 40])
def test_pcg_solves_assembled_system(mf, gpu, op, n):
    """At 40^3 (64,000 rows) every 16-lane row group of the operator product walks two
    rows; at 10^3 one."""
    print(f"[pcg] op={op} n={n}", flush=True)
    mesh = mf.Mesh.cube(n, 0.2, 3, op, coef_seed=11)
    plan = gpu.Plan(mesh)
    dm = gpu.DeviceMesh(mesh)
    plan.assemble(dm.coord, dm.coef, dm.values, dm.prec)
    K, m = masked_operator(mesh, dm.values.numpy())
    rng = np.random.default_rng(7)
    x_true = rng.standard_normal(K.shape[0])
    x_true[m] = 0.0
    b = K @ x_true
    db = gpu.DeviceArray.from_numpy(b)
    dx = gpu.DeviceArray.zeros(b.size, np.float64)
    it, rel = gpu.pcg(mesh.operatorID, mesh.nbNodes, dm.row, dm.col, dm.values, dm.prec,
                      dm.checkBounds, db, dx, rtol=1e-11, max_iter=2000)
    x = dx.numpy()
    true_rel = np.linalg.norm(b - K @ x) / np.linalg.norm(b)
    print(f"pcg op={op} n={n}: {it} iterations, reported {rel:.2e}, true {true_rel:.2e}")
    assert rel <= 1e-11 and true_rel <= 1e-10
    # forward error: the residual bound times the condition number, which grows with n
    assert np.linalg.norm(x - x_true) / np.linalg.norm(x_true) < (1e-7 if n == 10 else 1e-5)
    assert 0 < it < 2000


@pytest.mark.parametrize("op", [0, 1])
def test_plan_row_order_is_a_permutation(mf, gpu, op):
    """mf_plan_row_order: every node once, the rows of each tile together; a node with no
    element (an isolated node appended to the cube) comes last."""
    print(f"[row order] op={op}", flush=True)
    coord, e2n, code, nnz = mf.generate_cube(12, 0.2, 3, 17)
    coord = np.concatenate([coord, [2.0, 2.0, 2.0]])
    code = np.concatenate([code, [0]]).astype(np.int32)
    mesh = mf.Mesh(coord, e2n, code, nnz + 1, op)
    plan = gpu.Plan(mesh)
    order = plan.row_order().numpy()[:mesh.nbNodes]
    assert np.array_equal(np.sort(order), np.arange(mesh.nbNodes))
    assert order[-1] == mesh.nbNodes - 1


@pytest.mark.parametrize("op", [0, 1])
def test_pcg_ordered_solves_the_same_system(mf, gpu, op):
    """mf_pcg_ordered with the plan's order on a node-shuffled cube: the same solution as
    mf_pcg within the solve's tolerance and nearly the same iteration count (the dot
    products sum their terms in another order, and the rounding differences grow over
    the ~500 iterations), and x comes back in the caller's numbering."""
    print(f"[pcg ordered] op={op}", flush=True)
    coord, e2n, code, nnz = mf.generate_cube(24, 0.25, 3, 99)
    mesh = mf.Mesh(coord, e2n, code, nnz, op, coef=mf.generate_coef(op, e2n.size // 4, 5))
    plan = gpu.Plan(mesh)
    dm = gpu.DeviceMesh(mesh)
    plan.assemble(dm.coord, dm.coef, dm.values, dm.prec)
    K, m = masked_operator(mesh, dm.values.numpy())
    rng = np.random.default_rng(3)
    x_true = rng.standard_normal(K.shape[0])
    x_true[m] = 0.0
    b = K @ x_true
    db = gpu.DeviceArray.from_numpy(b)
    order = plan.row_order()
    res = []
    for o in (None, order):
        dx = gpu.DeviceArray.zeros(b.size, np.float64)
        it, rel = gpu.pcg(mesh.operatorID, mesh.nbNodes, dm.row, dm.col, dm.values, dm.prec,
                          dm.checkBounds, db, dx, rtol=1e-11, max_iter=3000, order=o)
        x = dx.numpy()
        true_rel = np.linalg.norm(b - K @ x) / np.linalg.norm(b)
        print(f"pcg op={op} order={o is not None}: {it} iterations, reported {rel:.2e}, "
              f"true {true_rel:.2e}")
        assert rel <= 1e-11 and true_rel <= 1e-10
        res.append((it, x))
    assert abs(res[0][0] - res[1][0]) <= max(2, res[0][0] // 50)
    assert np.linalg.norm(res[0][1] - res[1][1]) / np.linalg.norm(res[0][1]) < 1e-6


def test_pcg_ordered_rejects_a_bad_order(mf, gpu):
    """An order that repeats a node or leaves the node range is refused with MF_ERR_ARG
    before anything is written: x stays as it was."""
    mesh = mf.Mesh.cube(6, 0.2, 3, 1, coef_seed=2)
    plan = gpu.Plan(mesh)
    dm = gpu.DeviceMesh(mesh)
    plan.assemble(dm.coord, dm.coef, dm.values, dm.prec)
    n = mesh.nbNodes
    b = gpu.DeviceArray.from_numpy(np.ones(3 * n))
    for bad in (np.r_[np.arange(n - 1), 0], np.r_[np.arange(n - 1), n], np.r_[-1, np.arange(1, n)]):
        dx = gpu.DeviceArray.from_numpy(np.full(3 * n, 7.0))
        with pytest.raises(mf.MiniFemError) as exc:
            gpu.pcg(1, n, dm.row, dm.col, dm.values, dm.prec, dm.checkBounds, b, dx,
                    order=gpu.DeviceArray.from_numpy(bad.astype(np.int32)))
        assert exc.value.status == -1
        assert np.all(dx.numpy() == 7.0)
