"""The tiled schedule's host builder and validator (no GPU): every CSR entry is covered by
exactly one task, every index the kernel follows is in bounds, on several meshes and
tuning parameters; invalid inputs are rejected."""
import numpy as np
import pytest

from minifem.setup import plan_host_stats


@pytest.mark.parametrize("n,shuffle", [(2, 0), (6, 0), (9, 209), (20, 0)])
@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("params", [
    dict(), dict(tile_lds_bytes=8192), dict(tile_lds_bytes=150 * 1024, max_tile_rows=254),
    dict(max_tile_rows=1), dict(use_coord=False),
])
def test_plan_builds_and_validates(mf, n, shuffle, op, params):
    coord, e2n, code, nnz = mf.generate_cube(n, 0.2, 3, shuffle)
    mesh = mf.Mesh(coord, e2n, code, nnz, op)
    s = plan_host_stats(mesh, **params)
    assert s["tiles"] >= 1
    assert s["tile_elems"] >= mesh.nbElem  # every element is staged at least once
    # one diagonal task per non-empty row plus one row-chunk lane per CSR entry
    n_rows = int(np.count_nonzero(np.diff(mesh.row)))
    assert s["tasks"] == n_rows + mesh.nnz
    # every (row, element, vertex) incidence is one contribution, diagonals twice
    # (their own sum and the zero-contribution row lane is padding)
    assert s["contribs"] >= 16 * mesh.nbElem


def test_plan_rejects_bad_parameters(mf):
    coord, e2n, code, nnz = mf.generate_cube(4, 0.2, 3, 0)
    mesh = mf.Mesh(coord, e2n, code, nnz, 1)
    for bad in (dict(tile_lds_bytes=100), dict(tile_lds_bytes=10 ** 6),
                dict(max_tile_rows=300)):
        with pytest.raises(mf.MiniFemError):
            plan_host_stats(mesh, **bad)


def test_plan_rejects_inconsistent_pattern(mf):
    coord, e2n, code, nnz = mf.generate_cube(3, 0.0, 3, 0)
    mesh = mf.Mesh(coord, e2n, code, nnz, 0)
    mesh.row = mesh.row.copy()
    mesh.row[-1] -= 1  # row_ptr no longer ends at nnz
    with pytest.raises(mf.MiniFemError):
        plan_host_stats(mesh)
