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
def test_pcg_solves_assembled_system(mf, gpu, op):
    print(f"[pcg] op={op}", flush=True)
    mesh = mf.Mesh.cube(10, 0.2, 3, op, coef_seed=11)
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
    print(f"pcg op={op}: {it} iterations, reported {rel:.2e}, true {true_rel:.2e}")
    assert rel <= 1e-11 and true_rel <= 1e-10
    assert np.linalg.norm(x - x_true) / np.linalg.norm(x_true) < 1e-7
    assert 0 < it < 2000
