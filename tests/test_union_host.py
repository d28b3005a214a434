"""The disjoint-union mesh of the batched form (no GPU): each part's CSR segment is the
part's own CSR shifted by its node offset, the union's color-permuted elements keep every
part's own permuted order, and the oracle's COLORING assembly and prec of the union, cut
into the parts' ranges, equal each part's own bit for bit. The union's tiled schedule
(with the parts set apart for partitioning) builds and validates."""
import numpy as np
import pytest

from minifem.batched import UnionMesh, batch_meshes
from minifem.setup import plan_host_stats


def parts(op, count=5):
    ms = batch_meshes(count=2 * count, n_lo=60, n_hi=3000, seed=77)
    return [m for m in ms if m.operatorID == op]


@pytest.mark.parametrize("op", [0, 1])
def test_union_csr_and_element_order(mf, op):
    ps = parts(op)
    u = UnionMesh(ps)
    m = u.mesh
    assert m.nbNodes == sum(p.nbNodes for p in ps) and m.nnz == sum(p.nnz for p in ps)
    assert m.nbColors == max(p.nbColors for p in ps)
    # the part each permuted element belongs to, from its first vertex
    owner = np.searchsorted(u.node_off, m.elemToNode.reshape(-1, 4)[:, 0] - 1, side="right") - 1
    for k, p in enumerate(ps):
        a, b = u.node_off[k], u.node_off[k + 1]
        r = m.row[a:b + 1] - m.row[a]
        assert np.array_equal(r, p.row)
        assert np.array_equal(m.col[m.row[a]:m.row[b]] - a, p.col)
        assert np.array_equal(m.elemToNode.reshape(-1, 4)[owner == k] - a,
                              p.elemToNode.reshape(-1, 4))
        # the Dirichlet mask is dof-major: checkBounds[c * N + node]
        assert np.array_equal(m.checkBounds.reshape(3, -1)[:, a:b],
                              p.checkBounds.reshape(3, -1))
        va, vb = u.value_range(k)
        assert (vb - va) == p.nnz * p.dim


@pytest.mark.parametrize("op", [0, 1])
def test_union_oracle_results_equal_parts(mf, orc, op):
    ps = parts(op, count=4)
    u = UnionMesh(ps)
    m = u.mesh
    uv = orc.assembly_coloring(m.coord, m.row, m.col, m.elemToNode, op, m.colorToElem,
                               coef=m.coef)
    up, _ = orc.prec(uv, m.row, m.col, m.checkBounds, m.nbNodes, op)
    for k, p in enumerate(ps):
        pv = orc.assembly_coloring(p.coord, p.row, p.col, p.elemToNode, op, p.colorToElem,
                                   coef=p.coef)
        pp, _ = orc.prec(pv, p.row, p.col, p.checkBounds, p.nbNodes, op)
        va, vb = u.value_range(k)
        pa, pb = u.prec_range(k)
        assert np.array_equal(uv[va:vb], pv), k
        assert np.array_equal(up[pa:pb], pp), k


@pytest.mark.parametrize("op", [0, 1])
def test_union_plan_validates(mf, op):
    u = UnionMesh(parts(op))
    s = plan_host_stats(u.mesh, partition_coord=u.partition_coord)
    assert s["tiles"] >= len(u.parts)  # parts are set apart: no tile spans all of them
    assert s["tile_elems"] >= u.mesh.nbElem


def test_union_of_one_mesh_is_the_mesh(mf):
    p = parts(1, count=2)[0]
    u = UnionMesh([p])
    for a in ("row", "col", "elemToNode", "checkBounds", "colorToElem", "coef"):
        assert np.array_equal(getattr(u.mesh, a), getattr(p, a)), a
    assert u.value_range(0) == (0, p.nnz * 9) and u.prec_range(0) == (0, p.nbNodes * 9)


def test_partition_coord_size_is_checked(mf):
    u = UnionMesh(parts(0, count=2))
    with pytest.raises(ValueError):
        plan_host_stats(u.mesh, partition_coord=u.partition_coord[:-3])


def test_union_rejects_mixed_operators(mf):
    ms = batch_meshes(count=2, n_lo=60, n_hi=200, seed=77)
    with pytest.raises(ValueError):
        UnionMesh(ms)
    with pytest.raises(ValueError):
        UnionMesh([])
