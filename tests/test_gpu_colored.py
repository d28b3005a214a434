"""GPU parity of the drop-in entry points (mf_assembly[_ex], mf_prec_init,
mf_prec_inversion, mf_norm) against the CPU oracle, called through the C ABI.

Bar: bit-exact values and prec against the oracle's COLORING-order restatement (same
element order, same per-entry operation order); <= 1e-12 relative Frobenius error
against the REF-order oracle, compared per (row, col) key since the column orders
differ; mixed precision bit-exact against the oracle's mixed restatement and
<= 1e-6 relative Frobenius error against FP64.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def build(mf, orc, n, op, jitter=0.2, seed=1, shuffle=0, coef_seed=None):
    coord, e2n, code, nnz = mf.generate_cube(n, jitter, seed, shuffle)
    coef = None if coef_seed is None else mf.generate_coef(op, e2n.size // 4, coef_seed)
    mesh = mf.Mesh(coord, e2n, code, nnz, op, coef=coef)
    om = orc.OracleMesh(coord, e2n, code, nnz, op, coef=coef)
    assert np.array_equal(mesh.elemToNode, om.e2n) and np.array_equal(mesh.col, om.col)
    mesh.raw = (coord, e2n, code, nnz, coef)
    return mesh, om


def keyed(values, row, col, dim):
    n = len(row) - 1
    rows = np.repeat(np.arange(n), np.diff(row))
    key = rows.astype(np.int64) * (n + 1) + (col - 1)
    order = np.argsort(key)
    return key[order], values.reshape(-1, dim)[order]


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("coef_seed", [None, 101])
@pytest.mark.parametrize("use_e2e", [False, True])
def test_colored_bit_exact_30_cube(mf, orc, gpu, op, coef_seed, use_e2e):
    print(f"[colored] op={op} coef={coef_seed} e2e={use_e2e}", flush=True)
    mesh, om = build(mf, orc, 30, op, seed=1 + op, coef_seed=coef_seed)
    dm = gpu.DeviceMesh(mesh, with_elem_to_edge=use_e2e)
    dm.iterate_colored()
    vals = dm.values.numpy()
    prec = dm.prec.numpy()
    ov = om.assemble()
    op_, ns = om.prec(ov)
    assert np.array_equal(vals, ov)
    assert np.array_equal(prec, op_)
    # against the REF (unpermuted) order: same matrix to rounding
    coord, e2n, code, nnz, coef = mesh.raw
    ref = orc.OracleMesh(coord, e2n, code, nnz, op, coef=coef, colored=False)
    rv = ref.assemble()
    k1, a = keyed(vals, mesh.row, mesh.col, mesh.dim)
    k2, b = keyed(rv, ref.row, ref.col, mesh.dim)
    assert np.array_equal(k1, k2)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-12
    # norms (src/FEM.cc:59-98)
    assert abs(gpu.norm(dm.values) - orc.norm(ov)) <= 1e-13 * orc.norm(ov)
    assert abs(gpu.norm(dm.prec) - orc.norm(op_)) <= 1e-13 * orc.norm(op_)


@pytest.mark.parametrize("op", [0, 1])
def test_colored_shuffled_and_deterministic(mf, orc, gpu, op):
    print(f"[colored shuffled] op={op}", flush=True)
    mesh, om = build(mf, orc, 14, op, jitter=0.25, seed=3, shuffle=203, coef_seed=7)
    dm = gpu.DeviceMesh(mesh)
    dm.iterate_colored()
    v1 = dm.values.numpy()
    p1 = dm.prec.numpy()
    dm.iterate_colored()
    assert np.array_equal(v1, dm.values.numpy())
    assert np.array_equal(p1, dm.prec.numpy())
    ov = om.assemble()
    assert np.array_equal(v1, ov)
    assert np.array_equal(p1, om.prec(ov)[0])


@pytest.mark.parametrize("op", [0, 1])
def test_mixed_precision(mf, orc, gpu, op):
    print(f"[colored mixed] op={op}", flush=True)
    mesh, om = build(mf, orc, 20, op, coef_seed=55)
    dm = gpu.DeviceMesh(mesh)
    dm.iterate_colored(precision=mf.PREC_MIXED)
    vals = dm.values.numpy()
    assert np.array_equal(vals, om.assemble(mixed=True))
    ov = om.assemble()
    rel = np.linalg.norm(vals - ov) / np.linalg.norm(ov)
    print(f"mixed vs fp64 relative Frobenius error: {rel:.3e}")
    assert rel <= 1e-6
    # prec stays FP64 on the mixed values
    assert np.array_equal(dm.prec.numpy(), om.prec(vals)[0])


def test_golden_cube4(mf, orc, gpu):
    with open(os.path.join(GOLDEN, "cube4_values.json")) as f:
        gold = json.load(f)
    for op in (0, 1):
        coord, e2n, code, nnz = mf.generate_cube(4, 0.2, 7, 0)
        coef = mf.generate_coef(op, e2n.size // 4, 107)
        mesh = mf.Mesh(coord, e2n, code, nnz, op, coef=coef)
        g = gold[f"op{op}"]
        assert mesh.row.tolist() == g["row"] and mesh.col.tolist() == g["col"]
        dm = gpu.DeviceMesh(mesh)
        dm.iterate_colored()
        assert np.array_equal(dm.values.numpy(), np.array(g["values"]))
        assert np.array_equal(dm.prec.numpy(), np.array(g["prec"]))


@pytest.mark.parametrize("op", [0, 1])
def test_tiny_meshes(mf, orc, gpu, op):
    # single unit tet and the two-tet pair, no coefficients
    cases = [
        (np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0]], np.float64).ravel(),
         np.array([1, 2, 3, 4], np.int32), np.zeros(4, np.int32), 16),
        (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], np.float64).ravel(),
         np.array([1, 2, 3, 4, 2, 3, 4, 5], np.int32), np.array([52, 53, 54, 10, 0], np.int32), 23),
    ]
    for coord, e2n, code, nnz in cases:
        mesh = mf.Mesh(coord, e2n, code, nnz, op)
        om = orc.OracleMesh(coord, e2n, code, nnz, op)
        dm = gpu.DeviceMesh(mesh)
        dm.iterate_colored()
        ov = om.assemble()
        assert np.array_equal(dm.values.numpy(), ov)
        assert np.array_equal(dm.prec.numpy(), om.prec(ov)[0])
    # empty element set: values zeroed, nothing launched per color
    vals = gpu.DeviceArray.from_numpy(np.ones(10))
    gpu.assembly(None, vals, None, None, None, None, 0, 10, 1, 0, np.array([0], np.int32))
    gpu.synchronize()
    assert np.all(vals.numpy() == 0.0)


def test_singular_block_counted(mf, gpu):
    # one node, zero diagonal block, no mask: DGETRF reports a singular block
    D = gpu.DeviceArray.from_numpy
    prec = D(np.zeros(9))
    row = D(np.array([0, 1], np.int32))
    col = D(np.array([1], np.int32))
    cb = D(np.zeros(3, np.int32))
    sing = D(np.zeros(1, np.int32))
    gpu.prec_inversion(prec, row, col, cb, 1, 1, singular=sing)
    gpu.synchronize()
    assert int(sing.numpy()[0]) == 1
