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
