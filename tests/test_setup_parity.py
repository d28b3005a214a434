"""The product's host setup (libminifem, C++) against the CPU oracle, bit for bit:
node->element lists, CSR row_ptr/col_idx, colors, color permutation and offsets,
elemToEdge, the Dirichlet mask, plus properties of the synthetic generator and the mesh
and norm file round trips."""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

CASES = [  # (n, jitter, geomSeed, shuffleSeed)
    (2, 0.0, 1, 0),
    (5, 0.2, 1, 0),
    (7, 0.25, 3, 203),
    (30, 0.2, 1, 0),
]


@pytest.mark.parametrize("case", CASES)
def test_setup_bit_exact(mf, orc, case):
    n, jit, gs, ss = case
    coord, e2n, code, nnz = mf.generate_cube(n, jit, gs, ss)
    nN = coord.size // 3
    # nodeToElem
    i1, v1 = mf.create_nodeToElem(e2n, nN)
    i2, v2 = orc.nodeToElem(e2n, nN)
    assert np.array_equal(i1, i2) and np.array_equal(v1, v2)
    # REF-build CSR (unpermuted elements)
    r1, c1 = mf.create_nodeToNode(e2n, nN, nnz)
    r2, c2 = orc.csr(e2n, nN, nnz)
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
    # COLORING build: colors, permutation, permuted CSR
    p1, q1, o1 = mf.coloring_creation(e2n, nN)
    p2, q2, o2 = orc.coloring(e2n, nN)
    assert np.array_equal(p1, p2) and np.array_equal(q1, q2) and np.array_equal(o1, o2)
    pe1 = mf.permute_2d(e2n, q1, 4)
    pe2 = orc.permute(e2n, q2, 4)
    assert np.array_equal(pe1, pe2)
    r1, c1 = mf.create_nodeToNode(pe1, nN, nnz)
    r2, c2 = orc.csr(pe2, nN, nnz)
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
    assert int(r1[-1]) == nnz
    e1 = mf.create_elemToEdge(r1, c1, pe1)
    e2 = orc.elemToEdge(r2, c2, pe2)
    assert np.array_equal(e1, e2)
    assert np.array_equal(mf.boundary_mask(code), orc.boundary_mask(code))


def test_generator_properties(mf):
    n, J = 9, 0.25
    coord, e2n, code, nnz = mf.generate_cube(n, J, 5, 0)
    nE, nN, nnz2 = mf.cube_sizes(n)
    assert (e2n.size // 4, coord.size // 3, nnz) == (nE, nN, nnz2)
    h = 1.0 / (n - 1)
    X = coord.reshape(-1, 3)
    ijk = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1)
    ijk = ijk.transpose(2, 1, 0, 3).reshape(-1, 3)  # node id = i + n(j + n k)
    d = X - ijk * h
    interior = np.all((ijk > 0) & (ijk < n - 1), axis=1)
    assert np.all(d[~interior] == 0)
    assert np.all(np.abs(d[interior]) <= J * h) and np.abs(d[interior]).max() > 0.8 * J * h
    # boundary codes with precedence 10 > 52 > 53 > 54
    exp = np.zeros(nN, np.int32)
    exp[ijk[:, 2] == 0] = 54
    exp[ijk[:, 1] == 0] = 53
    exp[ijk[:, 0] == 0] = 52
    exp[(ijk == n - 1).any(axis=1)] = 10
    assert np.array_equal(code, exp)
    # conforming: every interior triangular face is shared by exactly two tets, and the
    # tets tile the cube (volumes of the unjittered mesh sum to 1)
    E = e2n.reshape(-1, 4) - 1
    faces = np.sort(np.concatenate([E[:, [0, 1, 2]], E[:, [0, 1, 3]], E[:, [0, 2, 3]],
                                    E[:, [1, 2, 3]]]), axis=1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    assert set(cnt.tolist()) <= {1, 2}
    c0, _, _, _ = mf.generate_cube(n, 0.0, 5, 0)
    P = c0.reshape(-1, 3)[E]
    vol = np.abs(np.einsum("ij,ij->i", P[:, 0] - P[:, 3],
                           np.cross(P[:, 1] - P[:, 3], P[:, 2] - P[:, 3]))) / 6
    assert abs(vol.sum() - 1.0) < 1e-12


def test_shuffle_is_a_relabeling(mf):
    c0, e0, b0, nnz0 = mf.generate_cube(6, 0.2, 9, 0)
    c1, e1, b1, nnz1 = mf.generate_cube(6, 0.2, 9, 209)
    assert nnz0 == nnz1
    X0, X1 = c0.reshape(-1, 3), c1.reshape(-1, 3)
    # recover the permutation from coordinates and check it maps everything consistently
    key0 = {tuple(x): i for i, x in enumerate(X0)}
    new_of_old = np.empty(len(X0), np.int64)
    for i, x in enumerate(X1):
        new_of_old[key0[tuple(x)]] = i
    assert not np.array_equal(new_of_old, np.arange(len(X0)))
    assert np.array_equal(new_of_old[e0 - 1] + 1, e1)
    assert np.array_equal(b1[new_of_old], b0)


def _h(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def test_config_hashes_match_golden(mf):
    """Pins the 30^3 generator output and its setup (pattern, colors, permutation) against
    hashes committed when the oracle first produced them (tests/golden/make_golden.py)."""
    with open(os.path.join(GOLDEN, "setup_hashes.json")) as f:
        gold = json.load(f)
    for name, g in gold.items():
        coord, e2n, code, nnz = mf.generate_cube(g["n"], g["jitter"], g["seed"], g["shuffle"])
        part, perm, c2e = mf.coloring_creation(e2n, coord.size // 3)
        pe = mf.permute_2d(e2n, perm, 4)
        row, col = mf.create_nodeToNode(pe, coord.size // 3, nnz)
        got = dict(coord=_h(coord), elemToNode=_h(e2n), code=_h(code), colors=_h(part),
                   perm=_h(perm), colorToElem=c2e.tolist(), row=_h(row), col=_h(col), nnz=nnz)
        for k, v in got.items():
            assert g[k] == v, (name, k)


def test_mesh_file_round_trip(mf, orc, tmp_path):
    coord, e2n, code, nnz = mf.generate_cube(6, 0.2, 4, 0)
    p = str(tmp_path / "ela_1_0")
    mf.store_input_data(p, coord, e2n, code, nnz)
    d = mf.read_input_data(p)
    assert (d["nbElem"], d["nbNodes"], d["nbEdges"], d["nbIntf"], d["nbIntfNodes"]) == \
        (e2n.size // 4, coord.size // 3, nnz, 0, 0)
    assert d["nbBoundNodes"] == np.count_nonzero(code)
    assert np.array_equal(d["coord"], coord) and np.array_equal(d["elemToNode"], e2n)
    assert np.array_equal(d["boundNodesCode"], code)
    # the oracle's independent reader agrees, byte layout included
    import ctypes
    h = np.zeros(6, np.int32)
    assert orc.lib().or_read_header(p.encode(), h.ctypes.data) == 0
    assert h.tolist() == [e2n.size // 4, coord.size // 3, nnz, 0, 0, int(np.count_nonzero(code))]
    size = 24 + coord.nbytes + e2n.nbytes + 12 + 4 + 0 + code.nbytes
    assert os.path.getsize(p) == size
    del ctypes


def test_truncated_and_missing_files(mf, tmp_path):
    coord, e2n, code, nnz = mf.generate_cube(3, 0.0, 1, 0)
    p = tmp_path / "lap_1_0"
    mf.store_input_data(str(p), coord, e2n, code, nnz)
    data = p.read_bytes()
    p.write_bytes(data[:-8])
    with pytest.raises(mf.MiniFemError):
        mf.read_input_data(str(p))
    with pytest.raises(mf.MiniFemError):
        mf.read_input_data(str(tmp_path / "missing"))


def test_ref_norm_file_round_trip(mf, tmp_path):
    p = str(tmp_path / "ela_1_0")
    mf.store_ref_assembly(p, 1234.5678901234567, 0.1)
    assert mf.read_ref_assembly(p) == (1234.5678901234567, 0.1)
    txt = open(p).read().split()
    assert txt == ["1234.5678901234567", "0.10000000000000001"]


def test_nnz_capacity_error(mf):
    coord, e2n, code, nnz = mf.generate_cube(4, 0.1, 1, 0)
    with pytest.raises(mf.MiniFemError) as ei:
        mf.create_nodeToNode(e2n, coord.size // 3, nnz - 1)
    assert ei.value.status == -4


def test_too_many_colors(mf):
    # 130 tets sharing node 1: every pair shares a node, so first fit needs 130 colors
    e2n = np.array([[1, 2 + 3 * i, 3 + 3 * i, 4 + 3 * i] for i in range(130)], np.int32).ravel()
    with pytest.raises(mf.MiniFemError) as ei:
        mf.coloring_creation(e2n, 1 + 3 * 130)
    assert ei.value.status == -3


def test_host_norm(mf, orc):
    x = np.random.default_rng(1).standard_normal(5000)
    assert abs(mf.compute_double_norm(x) - orc.norm(x)) <= 1e-14 * orc.norm(x)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_coloring_any_element_order(mf, orc, seed):
    """Greedy first fit depends on the element order: random element orders of a node-
    shuffled cube, the product's per-node color masks against the oracle's walk over
    every element sharing a node (src/coloring.cc:46-81), bit for bit."""
    coord, e2n, code, nnz = mf.generate_cube(14, 0.2, seed, 100 + seed)
    order = np.random.default_rng(seed).permutation(e2n.size // 4)
    e2n = np.ascontiguousarray(e2n.reshape(-1, 4)[order].ravel())
    nN = coord.size // 3
    got, want = mf.coloring_creation(e2n, nN), orc.coloring(e2n, nN)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    assert len(got[2]) - 1 > 20  # a random order needs many more colors than the cube's 30s


def test_coloring_60_cube_shuffled(mf, orc):
    coord, e2n, code, nnz = mf.generate_cube(60, 0.25, 3, 203)
    nN = coord.size // 3
    got, want = mf.coloring_creation(e2n, nN), orc.coloring(e2n, nN)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))


@pytest.mark.parametrize("threads", ["1", "3", "8"])
def test_setup_independent_of_host_threads(mf, orc, monkeypatch, threads):
    """nodeToElem and the CSR pattern run on several host threads (chunks of elements,
    ranges of rows); the result is the oracle's, whatever the count."""
    monkeypatch.setenv("MF_PLAN_THREADS", threads)
    coord, e2n, code, nnz = mf.generate_cube(40, 0.25, 3, 203)
    nN = coord.size // 3
    i1, v1 = mf.create_nodeToElem(e2n, nN)
    i2, v2 = orc.nodeToElem(e2n, nN)
    assert np.array_equal(i1, i2) and np.array_equal(v1, v2)
    r1, c1 = mf.create_nodeToNode(e2n, nN, nnz, n2e=(i1, v1))
    r2, c2 = orc.csr(e2n, nN, nnz)
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
    with pytest.raises(mf.MiniFemError) as ei:  # capacity one short
        mf.create_nodeToNode(e2n, nN, nnz - 1, n2e=(i1, v1))
    assert ei.value.status == -4
