"""The tiled schedule's host builder and validator (no GPU): every CSR entry is covered by
exactly one entry lane with an exact element mask, every index the kernel follows is in
bounds, on several meshes and tuning parameters; invalid inputs are rejected."""
import numpy as np
import pytest

from minifem.setup import plan_host_stats


@pytest.mark.parametrize("n,shuffle", [(2, 0), (6, 0), (9, 209), (20, 0)])
@pytest.mark.parametrize("op", [0, 1])
@pytest.mark.parametrize("params", [
    dict(), dict(tile_lds_bytes=8192), dict(tile_lds_bytes=150 * 1024, max_tile_rows=254),
    dict(max_tile_rows=1), dict(use_coord=False), dict(block_threads=512),
    dict(block_threads=1024), dict(id_chunks=3), dict(id_chunks=8, tile_lds_bytes=40000),
    dict(stage_precision=1), dict(stage_precision=2, tile_lds_bytes=8192),
])
def test_plan_builds_and_validates(mf, n, shuffle, op, params):
    coord, e2n, code, nnz = mf.generate_cube(n, 0.2, 3, shuffle)
    mesh = mf.Mesh(coord, e2n, code, nnz, op)
    s = plan_host_stats(mesh, **params)
    assert s["tiles"] >= 1
    assert s["tile_elems"] >= mesh.nbElem  # every element is staged at least once
    # one task per off-diagonal pair computed once: at most every off-diagonal block, at
    # least half of them (in-tile pairs always share a task)
    off = mesh.nnz - mesh.nbNodes
    assert (off + 1) // 2 <= s["tasks"] <= off
    if s["tiles"] == 1:
        assert s["tasks"] == off // 2  # every pair inside the one tile: computed once
    # every (row, incident element) pair once; each row's list padded to 4, each tile's
    # to 8
    assert 4 * mesh.nbElem <= s["incidences"] <= 4 * mesh.nbElem + 3 * mesh.nbNodes + 4 * s["tiles"]
    assert s["incidences"] % 4 == 0
    if "max_tile_rows" in params:
        assert s["max_rows"] <= params["max_tile_rows"]


def test_plan_rejects_bad_parameters(mf):
    coord, e2n, code, nnz = mf.generate_cube(4, 0.2, 3, 0)
    mesh = mf.Mesh(coord, e2n, code, nnz, 1)
    for bad in (dict(tile_lds_bytes=100), dict(tile_lds_bytes=10 ** 6),
                dict(max_tile_rows=300), dict(block_threads=128), dict(block_threads=256),
                dict(block_threads=384), dict(stage_precision=3), dict(stage_precision=-1)):
        with pytest.raises(mf.MiniFemError):
            plan_host_stats(mesh, **bad)


@pytest.mark.parametrize("op", [0, 1])
def test_float_stage_sizes_larger_tiles(mf, op):
    """Tiles sized by the float stage of the mixed modes hold more elements than those
    sized by the FP64 stage under the same LDS budget: fewer tiles, fewer staged elements
    (fewer halo copies), and the LDS a workgroup takes is the float layout's, within one
    CU's. Both mixed modes size alike."""
    coord, e2n, code, nnz = mf.generate_cube(24, 0.25, 3, 17)
    mesh = mf.Mesh(coord, e2n, code, nnz, op)
    s64 = plan_host_stats(mesh, tile_lds_bytes=40000)
    s32 = plan_host_stats(mesh, tile_lds_bytes=40000, stage_precision=1)
    assert plan_host_stats(mesh, tile_lds_bytes=40000, stage_precision=2) == s32
    assert s32["tiles"] < s64["tiles"]
    assert s32["tile_elems"] < s64["tile_elems"]
    assert 0 < s32["lds_bytes"] <= 163840 and 0 < s64["lds_bytes"] <= 163840


def test_plan_rejects_inconsistent_pattern(mf):
    coord, e2n, code, nnz = mf.generate_cube(3, 0.0, 3, 0)
    mesh = mf.Mesh(coord, e2n, code, nnz, 0)
    mesh.row = mesh.row.copy()
    mesh.row[-1] -= 1  # row_ptr no longer ends at nnz
    with pytest.raises(mf.MiniFemError):
        plan_host_stats(mesh)


def test_plan_accepts_valence_between_32_and_64(mf):
    """Valence 33..64 selects the two-word entry masks."""
    n_fan = 40
    ang = np.linspace(0, 2 * np.pi, n_fan + 1)[:-1]
    ring = np.stack([np.cos(ang), np.sin(ang), np.full(n_fan, 0.5)], 1)
    coord = np.concatenate([[[0, 0, 0], [0, 0, 1]], ring]).astype(np.float64)
    e2n = []
    for i in range(n_fan):
        e2n += [1, 2, 3 + i, 3 + (i + 1) % n_fan]
    e2n = np.array(e2n, np.int32)
    code = np.zeros(coord.shape[0], np.int32)
    nnz = 2 * (n_fan + 2) + 5 * n_fan
    mesh = mf.Mesh(coord.ravel(), e2n, code, nnz, 0)
    s = plan_host_stats(mesh)
    assert s["tasks"] <= mesh.nnz - mesh.nbNodes


@pytest.mark.parametrize("use_coord", [True, False])
def test_plan_independent_of_host_threads(mf, monkeypatch, use_coord):
    """The partition and the tile emission run on several host threads; the plan is the
    same with one (MF_PLAN_THREADS overrides the count read at each build)."""
    coord, e2n, code, nnz = mf.generate_cube(30, 0.2, 3, 209)
    mesh = mf.Mesh(coord, e2n, code, nnz, 1)
    got = []
    for threads in ("1", "4"):
        monkeypatch.setenv("MF_PLAN_THREADS", threads)
        got.append(plan_host_stats(mesh, use_coord=use_coord))
    assert got[0]["tiles"] >= 128  # enough tiles for several emission chunks
    assert got[0] == got[1]


def fan(mf, n_fan, op):
    """n_fan tets around the edge between two hub nodes (valence n_fan at both hubs)."""
    ang = np.linspace(0, 2 * np.pi, n_fan + 1)[:-1]
    ring = np.stack([np.cos(ang), np.sin(ang), 0.5 + 0.1 * np.sin(3 * ang)], 1)
    coord = np.concatenate([[[0, 0, 0], [0.05, -0.02, 1]], ring]).astype(np.float64).ravel()
    e2n = np.array([[1, 2, 3 + i, 3 + (i + 1) % n_fan] for i in range(n_fan)], np.int32).ravel()
    code = np.zeros(n_fan + 2, np.int32)
    return mf.Mesh(coord, e2n, code, 2 * (n_fan + 2) + 5 * n_fan, op)


@pytest.mark.parametrize("n_fan", [64, 65, 96, 127])
@pytest.mark.parametrize("op", [0, 1])
def test_plan_high_valence(mf, n_fan, op):
    """Rows with more incident elements than a mask has bits (64) build and validate: the
    hub-hub task takes a walk record; 127 tets at one edge need 127 colors, the most a
    mesh can have at a node and still color within 128."""
    mesh = fan(mf, n_fan, op)
    assert mesh.nbColors == n_fan
    s = plan_host_stats(mesh)
    assert s["tiles"] >= 1 and s["tile_elems"] == n_fan


def test_plan_rejects_valence_above_128(mf):
    """129 elements at one node cannot be colored within 128 colors, and the plan refuses
    them too (the mesh is built without coloring: elements in file order)."""
    n_fan = 129
    ang = np.linspace(0, 2 * np.pi, n_fan + 1)[:-1]
    ring = np.stack([np.cos(ang), np.sin(ang), 0.5 + 0.1 * np.sin(3 * ang)], 1)
    coord = np.concatenate([[[0, 0, 0], [0.05, -0.02, 1]], ring]).astype(np.float64).ravel()
    e2n = np.array([[1, 2, 3 + i, 3 + (i + 1) % n_fan] for i in range(n_fan)], np.int32).ravel()
    with pytest.raises(mf.MiniFemError):
        mf.coloring_creation(e2n, n_fan + 2)
    row, col = mf.create_nodeToNode(e2n, n_fan + 2, 2 * (n_fan + 2) + 5 * n_fan)
    nnz = col.size

    class M:  # the fields plan_host_stats reads
        pass
    m = M()
    m.operatorID, m.nbElem, m.nbNodes, m.nnz = 1, n_fan, n_fan + 2, nnz
    m.elemToNode, m.row, m.col, m.coord = e2n, row, col, coord
    with pytest.raises(mf.MiniFemError):
        plan_host_stats(m)
