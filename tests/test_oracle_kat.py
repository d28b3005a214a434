"""Pins the CPU oracle against facts fixed independently of its own formulas:
closed-form known answers on the unit tet, brute-force dense assembly from barycentric
gradients computed with numpy's 4x4 inverse, invariants (symmetry, exact power-of-two
scaling, zero row sums), the two-tet CSR/coloring answers, the closed-form 3x3 inverse,
numpy inverses of masked blocks, and the Dirichlet code table.

The reference ships no tests and no golden vectors; these checks are chosen so that a
dropped term, a swapped b/c, a wrong sign, a transposed block or an off-by-one index
fails at least one of them.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

UNIT_TET = dict(
    coord=np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, 0]], np.float64).ravel(),
    e2n=np.array([1, 2, 3, 4], np.int32),
)


def kat():
    with open(os.path.join(GOLDEN, "kat.json")) as f:
        return json.load(f)


def dense_from_csr(values, row, col, n, dim):
    b = int(round(dim ** 0.5))
    K = np.zeros((n * b, n * b))
    for i in range(n):
        for l in range(row[i], row[i + 1]):
            j = col[l] - 1
            K[i * b:(i + 1) * b, j * b:(j + 1) * b] = values[l * dim:(l + 1) * dim].reshape(b, b)
    return K


def brute_force(coord, e2n, op, coef=None):
    """Dense assembly from barycentric gradients: grad(phi_i) = column i of inv(M)[:3],
    M rows [x y z 1]; lap k(gi.gj); ela mu(gi.gj)I + (lambda+mu) gi gj^T."""
    coord = coord.reshape(-1, 3)
    e2n = e2n.reshape(-1, 4) - 1
    n = coord.shape[0]
    b = 3 if op == 1 else 1
    K = np.zeros((n * b, n * b))
    for e, nodes in enumerate(e2n):
        M = np.hstack([coord[nodes], np.ones((4, 1))])
        G = np.linalg.inv(M)[:3].T  # G[i] = grad phi_i
        for i in range(4):
            for j in range(4):
                if op == 0:
                    k = 1.0 if coef is None else coef[e]
                    blk = k * G[i] @ G[j]
                else:
                    lam, mu = (0.25, 1.0) if coef is None else coef[2 * e:2 * e + 2]
                    blk = mu * (G[i] @ G[j]) * np.eye(3) + (lam + mu) * np.outer(G[i], G[j])
                I, J = nodes[i], nodes[j]
                K[I * b:(I + 1) * b, J * b:(J + 1) * b] += blk
    return K


def test_unit_tet_gradients(orc):
    g = orc.elem_coef(UNIT_TET["coord"], UNIT_TET["e2n"], 0)
    assert np.array_equal(g, np.array(kat()["unit_tet"]["gradients"]))


@pytest.mark.parametrize("op", [0, 1])
def test_unit_tet_matrix_closed_form(orc, op):
    k = kat()["unit_tet"]
    m = orc.OracleMesh(UNIT_TET["coord"], UNIT_TET["e2n"], np.zeros(4, np.int32), 16, op)
    vals = m.assemble()
    dim = 9 if op else 1
    K = dense_from_csr(vals, m.row, m.col, 4, dim)
    if op == 0:
        assert np.array_equal(K, np.array(k["lap_K"]))
    else:
        assert np.array_equal(K[0:3, 0:3], np.array(k["ela_block_00"]).reshape(3, 3))
        assert np.array_equal(K[9:12, 9:12], np.array(k["ela_block_33"]).reshape(3, 3))
        assert np.array_equal(K[0:3, 9:12], np.array(k["ela_block_03"]).reshape(3, 3))


@pytest.mark.parametrize("op", [0, 1])
def test_unit_tet_prec_closed_form(orc, op):
    k = kat()["unit_tet"]
    m = orc.OracleMesh(UNIT_TET["coord"], UNIT_TET["e2n"], np.zeros(4, np.int32), 16, op)
    p, ns = m.prec(m.assemble())
    assert ns == 0
    if op == 0:
        assert np.allclose(p, k["lap_prec"], rtol=0, atol=1e-15)
    else:
        inv33 = p[27:36].reshape(3, 3)
        ex = np.full((3, 3), -5.0 / 81.0) + np.eye(3) * (27.0 / 81.0)
        assert np.allclose(inv33, ex, rtol=0, atol=1e-15)
        inv00 = p[0:9].reshape(3, 3)
        assert np.allclose(inv00, np.diag([1 / 2.25, 1.0, 1.0]), rtol=0, atol=1e-16)


def test_two_tet_csr_and_coloring(orc):
    k = kat()["two_tet"]
    e2n = np.array(k["elemToNode"], np.int32)
    row, col = orc.csr(e2n, 5, 64)
    assert row.tolist() == k["row_ptr"]
    assert col.tolist() == k["col_idx"]
    part, perm, c2e = orc.coloring(e2n, 5)
    assert part.tolist() == k["colors"]
    assert c2e.tolist() == k["colorToElem"]
    assert perm.tolist() == [0, 1]


def test_csr_capacity_overflow(orc):
    e2n = np.array(kat()["two_tet"]["elemToNode"], np.int32)
    with pytest.raises(ValueError):
        orc.csr(e2n, 5, 22)  # true nnz is 23


def small_mesh(mf, n=4, op=1, jitter=0.2, seed=7, shuffle=0, coef_seed=None):
    coord, e2n, code, nnz = mf.generate_cube(n, jitter, seed, shuffle)
    coef = None if coef_seed is None else mf.generate_coef(op, e2n.size // 4, coef_seed)
    return coord, e2n, code, nnz, coef


@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("with_coef", [False, True])
@pytest.mark.parametrize("colored", [False, True])
def test_brute_force_dense(orc, mf, op, with_coef, colored):
    coord, e2n, code, nnz, coef = small_mesh(mf, n=4, op=op, coef_seed=11 if with_coef else None)
    m = orc.OracleMesh(coord, e2n, code, nnz, op, coef=coef, colored=colored)
    vals = m.assemble()
    K = dense_from_csr(vals, m.row, m.col, m.nN, m.dim)
    ref = brute_force(m.coord, m.e2n, op, m.coef)
    err = np.linalg.norm(K - ref) / np.linalg.norm(ref)
    assert err < 1e-13, err


@pytest.mark.parametrize("op", [0, 1])
def test_symmetry_bitwise_and_row_sums(orc, mf, op):
    coord, e2n, code, nnz, coef = small_mesh(mf, n=6, op=op, coef_seed=3)
    m = orc.OracleMesh(coord, e2n, code, nnz, op, coef=coef)
    vals = m.assemble()
    K = dense_from_csr(vals, m.row, m.col, m.nN, m.dim)
    assert np.array_equal(K, K.T)  # (i,j) and (j,i) get the same terms in the same order
    rs = np.abs(K.reshape(K.shape[0], -1, 3 if op else 1)).sum()
    if op == 0:
        assert np.max(np.abs(K.sum(axis=1))) < 1e-11 * np.abs(K).max()
    else:
        # ela: sum over column blocks of each row block vanishes (rigid translations)
        n = m.nN
        T = K.reshape(n, 3, n, 3).sum(axis=2)
        assert np.max(np.abs(T)) < 1e-11 * np.abs(K).max()
    assert rs > 0


@pytest.mark.parametrize("op", [0, 1])
def test_power_of_two_scaling_exact(orc, mf, op):
    coord, e2n, code, nnz, coef = small_mesh(mf, n=5, op=op, coef_seed=5)
    m1 = orc.OracleMesh(coord, e2n, code, nnz, op, coef=coef)
    m2 = orc.OracleMesh(coord * 8.0, e2n, code, nnz, op, coef=coef)
    assert np.array_equal(m2.assemble(), m1.assemble() / 64.0)


def test_coefficient_extension_reduces_to_reference(orc, mf):
    for op in (0, 1):
        coord, e2n, code, nnz, _ = small_mesh(mf, n=5, op=op)
        nE = e2n.size // 4
        unit = np.ones(nE) if op == 0 else np.tile([0.25, 1.0], nE)
        m = orc.OracleMesh(coord, e2n, code, nnz, op, coef=unit)
        assert np.array_equal(m.assemble(use_coef=True), m.assemble(use_coef=False))


def test_coloring_order_matches_ref_order_within_tolerance(orc, mf):
    coord, e2n, code, nnz, coef = small_mesh(mf, n=8, op=1, coef_seed=2)
    ref = orc.OracleMesh(coord, e2n, code, nnz, 1, coef=coef, colored=False)
    col = orc.OracleMesh(coord, e2n, code, nnz, 1, coef=coef, colored=True)
    assert np.array_equal(ref.row, col.row)  # row_ptr identical, columns reordered
    Kr = dense_from_csr(ref.assemble(), ref.row, ref.col, ref.nN, 9)
    Kc = dense_from_csr(col.assemble(), col.row, col.col, col.nN, 9)
    assert np.linalg.norm(Kr - Kc) / np.linalg.norm(Kr) < 1e-14
    # the coloring-order sum done color by color equals the plain loop over the permuted
    # elements, serial or parallel inside a color
    v1 = col.assemble()
    v2 = orc.assembly_coloring(col.coord, col.row, col.col, col.e2n, 1, col.c2e, coef=col.coef)
    v3 = orc.assembly_coloring(col.coord, col.row, col.col, col.e2n, 1, col.c2e, coef=col.coef,
                               parallel=True)
    assert np.array_equal(v1, v2) and np.array_equal(v1, v3)


def test_elem_to_edge_and_search_agree(orc, mf):
    coord, e2n, code, nnz, coef = small_mesh(mf, n=6, op=1, coef_seed=4)
    m = orc.OracleMesh(coord, e2n, code, nnz, 1, coef=coef)
    e2e = orc.elemToEdge(m.row, m.col, m.e2n)
    E = m.e2n.reshape(-1, 4)
    pos = e2e.reshape(-1, 4, 4)
    for j in range(4):
        for k in range(4):
            p = pos[:, j, k]
            assert np.all(p >= m.row[E[:, j] - 1]) and np.all(p < m.row[E[:, j]])
            assert np.array_equal(m.col[p], E[:, k])
    v1 = orc.assembly(m.coord, m.row, m.col, m.e2n, 1, coef=m.coef, e2e=e2e)
    assert np.array_equal(v1, m.assemble())


def test_coloring_valid_and_first_fit(orc, mf):
    coord, e2n, code, nnz, _ = small_mesh(mf, n=9, op=0)
    nN = coord.size // 3
    part, perm, c2e = orc.coloring(e2n, nN)
    E = e2n.reshape(-1, 4) - 1
    n2e = [[] for _ in range(nN)]
    for e, vs in enumerate(E):
        for v in vs:
            n2e[v].append(e)
    for e, vs in enumerate(E):
        nb = set()
        for v in vs:
            nb.update(n2e[v])
        nb.discard(e)
        used = {part[f] for f in nb if f < e}
        assert part[e] not in {part[f] for f in nb}  # valid
        assert part[e] == min(set(range(129)) - used)  # first fit
    # stable counting sort by color
    order = np.argsort(part, kind="stable")
    assert np.array_equal(np.argsort(perm), order)
    assert c2e[0] == 0 and c2e[-1] == E.shape[0]
    assert np.array_equal(np.diff(c2e), np.bincount(part))


def test_csr_closed_form_nnz_and_first_seen_order(orc, mf):
    for n in (3, 5, 8):
        coord, e2n, code, nnz, _ = small_mesh(mf, n=n, op=0)
        nN = coord.size // 3
        row, col = orc.csr(e2n, nN, nnz)
        assert int(row[-1]) == nnz  # closed form N + 2(3mn^2 + 3nm^2 + m^3)
        lens = np.diff(row)
        assert lens.max() == (15 if n > 2 else lens.max())
        # first-seen order: brute force over the node->element lists
        E = e2n.reshape(-1, 4)
        for i in range(0, nN, max(1, nN // 40)):
            seen = []
            for e in np.nonzero((E == i + 1).any(axis=1))[0]:
                for v in E[e]:
                    if v not in seen:
                        seen.append(v)
            assert col[row[i]:row[i + 1]].tolist() == seen


def test_boundary_code_table(orc):
    k = kat()["mask"]
    code = np.array(k["codes"], np.int32)
    cb = orc.boundary_mask(code)
    assert cb.reshape(3, -1).T.tolist() == k["mask"]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ela_prec_against_numpy_inverse(orc, mf, seed):
    coord, e2n, code, nnz, coef = small_mesh(mf, n=5, op=1, seed=seed, coef_seed=seed + 9)
    m = orc.OracleMesh(coord, e2n, code, nnz, 1, coef=coef)
    vals = m.assemble()
    p, ns = m.prec(vals)
    assert ns == 0
    cb = m.cb.reshape(3, -1)
    for i in range(m.nN):
        l = next(l for l in range(m.row[i], m.row[i + 1]) if m.col[l] == i + 1)
        A = vals[l * 9:(l + 1) * 9].reshape(3, 3).T.copy()  # Fortran view
        for k in range(3):
            if cb[k, i]:
                A[k, :] = 0
                A[:, k] = 0
                A[k, k] = 1
        inv = np.linalg.inv(A)
        got = p[i * 9:(i + 1) * 9].reshape(3, 3).T
        assert np.allclose(got, inv, rtol=1e-13, atol=1e-13 * np.abs(inv).max())
        assert np.allclose(got @ A, np.eye(3), atol=1e-13)


def test_singular_block_reported(orc):
    # a zero diagonal block with no mask: DGETRF reports info != 0
    row = np.array([0, 1], np.int32)
    col = np.array([1], np.int32)
    vals = np.zeros(9)
    p, ns = orc.prec(vals, row, col, np.zeros(3, np.int32), 1, 1)
    assert ns == 1


def test_norm_matches_numpy(orc):
    x = np.random.default_rng(0).standard_normal(10001)
    assert abs(orc.norm(x) - np.sqrt(np.sum(x * x))) <= 1e-14 * orc.norm(x)


def test_lap_prec_reciprocal(orc, mf):
    coord, e2n, code, nnz, coef = small_mesh(mf, n=5, op=0, coef_seed=1)
    m = orc.OracleMesh(coord, e2n, code, nnz, 0, coef=coef)
    vals = m.assemble()
    p, _ = m.prec(vals)
    diag = np.array([vals[next(l for l in range(m.row[i], m.row[i + 1]) if m.col[l] == i + 1)]
                     for i in range(m.nN)])
    assert np.array_equal(p, 1.0 / diag)
