"""Device-side setup (mf_dev_create_*) bit-exact against the host setup, which the
oracle pins (tests/test_setup_parity.py): nodeToElem, the CSR pattern in first-seen
order and elemToEdge, on ordered, shuffled and the 100^3 benchmark meshes; the nnz
capacity error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,shuffle", [(2, 0), (9, 0), (17, 217), (30, 0), (100, 203)])
def test_device_setup_equals_host(mf, gpu, n, shuffle):
    print(f"[dev setup] n={n} shuffle={shuffle}", flush=True)
    coord, e2n, code, nnz = mf.generate_cube(n, 0.25, 3, shuffle)
    mesh = mf.Mesh(coord, e2n, code, nnz, 1)
    index, value, row, col, e2e, dnnz = gpu.device_setup(mesh.elemToNode, mesh.nbNodes, nnz)
    hi, hv = mf.create_nodeToElem(mesh.elemToNode, mesh.nbNodes)
    assert np.array_equal(index.numpy(), hi)
    assert np.array_equal(value.numpy()[:hv.size], hv)
    assert dnnz == mesh.nnz
    assert np.array_equal(row.numpy(), mesh.row)
    assert np.array_equal(col.numpy()[:mesh.nnz], mesh.col)
    if n <= 30:
        he = mf.create_elemToEdge(mesh.row, mesh.col, mesh.elemToNode)
        assert np.array_equal(e2e.numpy()[:he.size], he)


def test_device_setup_nnz_capacity(mf, gpu):
    coord, e2n, code, nnz = mf.generate_cube(4, 0.0, 3, 0)
    mesh = mf.Mesh(coord, e2n, code, nnz, 0)
    with pytest.raises(mf.MiniFemError):
        gpu.device_setup(mesh.elemToNode, mesh.nbNodes, nnz - 1)


def _host_coloring(mf, e2n, N):
    return mf.coloring_creation(np.ascontiguousarray(e2n, np.int32).ravel(), N)


@pytest.mark.parametrize("case", ["cube2", "cube9", "cube17s", "cube30", "cube58s", "cube100s",
                                  "cube20_random_order", "cube60_random_order", "random_dense",
                                  "random_wide"])
def test_device_coloring_equals_host(mf, gpu, case):
    """mf_dev_coloring_creation against the host scan (which the oracle pins): colors,
    permutation and color offsets equal, on ordered and shuffled cubes, cubes whose
    elements come in random order (few, wide levels: the 60^3 one has levels of ~10^4
    elements), and random tets on few nodes (valence up to ~120, repeated vertices)."""
    import time
    print(f"[dev coloring] {case}", flush=True)
    rng = np.random.default_rng(11)
    if case.endswith("_random_order"):
        coord, e2n, _, _ = mf.generate_cube(int(case[4:6]), 0.2, 3, 0)
        e2n = e2n.reshape(-1, 4)[rng.permutation(e2n.size // 4)]
        N = coord.size // 3
    elif case.startswith("cube"):
        n = int(case[4:].rstrip("s"))
        coord, e2n, _, _ = mf.generate_cube(n, 0.25, 3, 200 + n if case.endswith("s") else 0)
        N = coord.size // 3
    elif case == "random_dense":
        N, e2n = 40, rng.integers(1, 41, (300, 4))
    else:
        N, e2n = 1000, rng.integers(1, 1001, (20000, 4))
    e2n = np.ascontiguousarray(e2n, np.int32).ravel()
    hp, hperm, hc2e = _host_coloring(mf, e2n, N)
    d_e2n = gpu.DeviceArray.from_numpy(e2n)
    gpu.synchronize()
    t = time.perf_counter()
    dp, dperm, dc2e = gpu.coloring_creation(d_e2n, N)
    dt = time.perf_counter() - t
    nE = e2n.size // 4
    print(f"[dev coloring] {case}: {nE} elements, {len(hc2e) - 1} colors, {dt * 1e3:.1f} ms",
          flush=True)
    assert np.array_equal(dp.numpy()[:nE], hp)
    assert np.array_equal(dperm.numpy()[:nE], hperm)
    assert np.array_equal(dc2e, hc2e)


def test_device_coloring_errors(mf, gpu):
    """More than 128 colors (a fan of 130 tets around one edge) fails as on the host; an
    empty mesh gives no colors; a misaligned element array and bad node ids are refused."""
    import ctypes
    from minifem._lib import MF_ERR_ARG, MF_ERR_COLORS
    lib = mf.lib()
    fan = np.array([[1, 2, 3 + i, 3 + (i + 1) % 130] for i in range(130)], np.int32).ravel()
    with pytest.raises(mf.MiniFemError):
        _host_coloring(mf, fan, 132)
    d = gpu.DeviceArray.from_numpy(fan)
    part = gpu.DeviceArray(130, np.int32)
    perm = gpu.DeviceArray(130, np.int32)
    c2e = gpu.DeviceArray(129, np.int32)
    nc = ctypes.c_int(-1)
    assert lib.mf_dev_coloring_creation(d.ptr, part.ptr, perm.ptr, 130, 132, c2e.ptr,
                                        ctypes.byref(nc), None) == MF_ERR_COLORS
    assert lib.mf_dev_coloring_creation(d.ptr + 4, part.ptr, perm.ptr, 129, 132, c2e.ptr,
                                        ctypes.byref(nc), None) == MF_ERR_ARG
    assert lib.mf_dev_coloring_creation(d.ptr, part.ptr, perm.ptr, 130, 100, c2e.ptr,
                                        ctypes.byref(nc), None) == MF_ERR_ARG  # ids above N
    assert lib.mf_dev_coloring_creation(None, None, perm.ptr, 0, 0, c2e.ptr,
                                        ctypes.byref(nc), None) == 0
    assert nc.value == 0 and c2e.numpy()[0] == 0


def test_device_permute(mf, gpu):
    rng = np.random.default_rng(5)
    n = 1000
    perm = rng.permutation(n).astype(np.int32)
    dperm = gpu.DeviceArray.from_numpy(perm)
    for dtype, dim in ((np.int32, 4), (np.float64, 2), (np.float64, 1)):
        a = rng.integers(0, 1 << 20, n * dim).astype(dtype)
        d = gpu.DeviceArray.from_numpy(a)
        gpu.permute_2d(d, dperm, dim)
        assert np.array_equal(d.numpy(), mf.permute_2d(a, perm, dim))
    bad = perm.copy()
    bad[7] = n
    a = np.arange(n * 4, dtype=np.int32)
    d = gpu.DeviceArray.from_numpy(a)
    with pytest.raises(mf.MiniFemError):
        gpu.permute_2d(d, gpu.DeviceArray.from_numpy(bad), 4)
    assert np.array_equal(d.numpy(), a)  # left unchanged


@pytest.mark.parametrize("n,op,shuffle", [(12, 0, 0), (30, 1, 230), (100, 1, 203)])
def test_mesh_device_setup_equals_host(mf, gpu, n, op, shuffle):
    """The whole setup on the GPU (coloring, element and coefficient permutation, CSR):
    every array equal to the host setup's."""
    import time
    print(f"[dev mesh setup] n={n} op={op}", flush=True)
    coord, e2n, code, nnz = mf.generate_cube(n, 0.25, 3, shuffle)
    coef = mf.generate_coef(op, e2n.size // 4, 103)
    t = time.perf_counter()
    h = mf.Mesh(coord, e2n, code, nnz, op, coef=coef)
    th = time.perf_counter() - t
    t = time.perf_counter()
    d = mf.Mesh(coord, e2n, code, nnz, op, coef=coef, setup="device")
    td = time.perf_counter() - t
    print(f"[dev mesh setup] n={n}: host {th:.3f} s, device {td:.3f} s", flush=True)
    for name in ("elemToNode", "coef", "colorPerm", "colorToElem", "row", "col", "checkBounds"):
        assert np.array_equal(getattr(h, name), getattr(d, name)), name
    assert h.nnz == d.nnz
