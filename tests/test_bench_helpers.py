"""bench.py's own arithmetic (no GPU): the algorithmic byte count the roofline divides by
(the closed form at the 30^3 and 100^3 cube sizes) and the relative Frobenius error every
bench line states."""
import importlib.util
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_algorithmic_bytes_at_the_cube_sizes(mf):
    b = load_bench()
    # (nE, N, nnz) of the Kuhn cubes, checked against the generator's own counts
    for n, op, want in ((30, 0, 9_089_116), (30, 1, 36_872_300), (100, 1, 1_420_148_460),
                        (100, 0, 352_857_436)):
        nE, N, nnz = mf.cube_sizes(n)
        assert b.algorithmic_bytes(nE, N, nnz, op) == want, (n, op)
    nE, N, nnz = mf.cube_sizes(100)
    assert (nE, N, nnz) == (5_821_794, 1_000_000, 14_761_198)


def test_rel_fro():
    b = load_bench()
    ref = np.array([3.0, 4.0, np.inf, 0.0])
    assert b.rel_fro(ref.copy(), ref) == 0.0
    got = np.array([3.0, 4.0 + 5e-3, np.inf, 0.0])
    assert abs(b.rel_fro(got, ref) - 1e-3) < 1e-15  # inf entries are left out
    assert b.rel_fro(np.array([1e-3]), np.array([0.0])) == 1e-3  # absolute when ref is 0


def test_delaunay_workload_mesh(mf):
    """bench.py --delaunay: the nnz it returns is the exact CSR size the setup builds (the
    diagonal plus both directions of every edge), every point is used, and only hull
    nodes carry boundary codes."""
    from scipy.spatial import Delaunay
    b = load_bench()
    coord, e2n, code, nnz = b.delaunay_arrays(2000, seed=3)
    n = coord.size // 3
    assert n == 2000 and e2n.min() == 1 and e2n.max() == n
    assert np.unique(e2n).size == n
    mesh = mf.Mesh(coord, e2n, code, nnz, mf.OP_ELA)
    assert mesh.nnz == nnz
    hull = np.unique(Delaunay(coord.reshape(-1, 3)).convex_hull)
    assert set(np.flatnonzero(code)) <= set(hull)
    assert set(np.unique(code[hull])) <= {10, 52, 53, 54}
