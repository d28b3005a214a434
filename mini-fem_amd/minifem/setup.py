"""Host-side mirror of the reference's setup interface (src/main.cc:140-351) over the C
ABI: mesh files, the synthetic generator, the CSR pattern, coloring, permutation,
element->CSR positions and the Dirichlet mask. Arrays are numpy; every call checks the
C status and raises MiniFemError on failure, where the reference would exit().
"""
import numpy as np

from ._lib import OP_ELA, OP_LAP, MF_MAX_COLOR, check, lib


def _ptr(a):
    return a.ctypes.data if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# --- files: src/IO.cc ---------------------------------------------------------------

def read_header(path):
    h = np.zeros(6, np.int32)
    check(lib().mf_read_header(str(path).encode(), _ptr(h)), f"read_header {path}")
    return [int(x) for x in h]


def read_input_data(path):
    """read_input_data (src/IO.cc:61-96). Returns a dict of the header counts and the six
    arrays."""
    nE, nN, nnz, nI, nIN, nB = read_header(path)
    out = dict(nbElem=nE, nbNodes=nN, nbEdges=nnz, nbIntf=nI, nbIntfNodes=nIN,
               nbBoundNodes=nB,
               coord=np.zeros(nN * 3, np.float64), elemToNode=np.zeros(nE * 4, np.int32),
               neighborsList=np.zeros(max(nI, 1) * 3, np.int32),
               intfIndex=np.zeros(nI + 1, np.int32), intfNodes=np.zeros(max(nIN, 1), np.int32),
               boundNodesCode=np.zeros(nN, np.int32))
    check(lib().mf_read_input_data(str(path).encode(), _ptr(out["coord"]), _ptr(out["elemToNode"]),
                                   _ptr(out["neighborsList"]), _ptr(out["intfIndex"]),
                                   _ptr(out["intfNodes"]), _ptr(out["boundNodesCode"])),
          f"read_input_data {path}")
    out["intfNodes"] = out["intfNodes"][:nIN]
    return out


def store_input_data(path, coord, elemToNode, boundNodesCode, nbEdges):
    """store_input_data_ (src/IO.cc:99-127) for one rank (no interface)."""
    coord = _c(coord, np.float64)
    e2n = _c(elemToNode, np.int32)
    code = _c(boundNodesCode, np.int32)
    h = np.array([e2n.size // 4, code.size, int(nbEdges), 0, 0, int(np.count_nonzero(code))],
                 np.int32)
    check(lib().mf_store_input_data(str(path).encode(), _ptr(h), _ptr(coord), _ptr(e2n),
                                    None, None, None, _ptr(code)), f"store_input_data {path}")


def read_ref_assembly(path):
    m = np.zeros(1)
    p = np.zeros(1)
    check(lib().mf_read_ref_assembly(str(path).encode(), _ptr(m), _ptr(p)), f"read {path}")
    return float(m[0]), float(p[0])


def store_ref_assembly(path, matrix_norm, prec_norm):
    check(lib().mf_store_ref_assembly(str(path).encode(), float(matrix_norm), float(prec_norm)),
          f"store {path}")


# --- synthetic meshes ---------------------------------------------------------------

def cube_sizes(n):
    nE = np.zeros(1, np.int32)
    nN = np.zeros(1, np.int32)
    nnz = np.zeros(1, np.int64)
    check(lib().mf_cube_sizes(int(n), _ptr(nE), _ptr(nN), _ptr(nnz)), "cube_sizes")
    return int(nE[0]), int(nN[0]), int(nnz[0])


def generate_cube(n, jitter, geom_seed, shuffle_seed=0):
    nE, nN, nnz = cube_sizes(n)
    coord = np.zeros(nN * 3, np.float64)
    e2n = np.zeros(nE * 4, np.int32)
    code = np.zeros(nN, np.int32)
    check(lib().mf_generate_cube(int(n), float(jitter), int(geom_seed), int(shuffle_seed),
                                 _ptr(coord), _ptr(e2n), _ptr(code)), "generate_cube")
    return coord, e2n, code, nnz


def generate_coef(operator_id, nb_elem, seed):
    coef = np.zeros(nb_elem * (2 if operator_id == OP_ELA else 1), np.float64)
    check(lib().mf_generate_coef(int(operator_id), int(nb_elem), int(seed), _ptr(coef)),
          "generate_coef")
    return coef


# --- setup: src/matrix.cc, src/coloring.cc, Fortran mask ----------------------------

def create_nodeToElem(elemToNode, nbNodes):
    e2n = _c(elemToNode, np.int32)
    idx = np.zeros(nbNodes + 1, np.int32)
    val = np.zeros(max(e2n.size, 1), np.int32)
    check(lib().mf_create_nodeToElem(_ptr(e2n), e2n.size // 4, int(nbNodes), _ptr(idx), _ptr(val)),
          "create_nodeToElem")
    return idx, val[:e2n.size]


def create_nodeToNode(elemToNode, nbNodes, nbEdges, n2e=None):
    e2n = _c(elemToNode, np.int32)
    idx, val = n2e if n2e is not None else create_nodeToElem(e2n, nbNodes)
    row = np.zeros(nbNodes + 1, np.int32)
    col = np.zeros(max(int(nbEdges), 1), np.int32)
    nnz = check(lib().mf_create_nodeToNode(_ptr(row), _ptr(col), _ptr(idx), _ptr(val), _ptr(e2n),
                                           int(nbNodes), int(nbEdges)), "create_nodeToNode")
    return row, col[:nnz]


def create_elemToEdge(row, col, elemToNode):
    e2n = _c(elemToNode, np.int32)
    out = np.zeros(max(e2n.size * 4, 1), np.int32)
    check(lib().mf_create_elemToEdge(_ptr(_c(row, np.int32)), _ptr(_c(col, np.int32)), _ptr(e2n),
                                     _ptr(out), e2n.size // 4), "create_elemToEdge")
    return out[:e2n.size * 4]


def coloring_creation(elemToNode, nbNodes):
    """Returns (colorPart, colorPerm (old->new), colorToElem[nColors+1])."""
    e2n = _c(elemToNode, np.int32)
    nE = e2n.size // 4
    part = np.zeros(max(nE, 1), np.int32)
    perm = np.zeros(max(nE, 1), np.int32)
    c2e = np.zeros(MF_MAX_COLOR + 1, np.int32)
    nc = np.zeros(1, np.int32)
    check(lib().mf_coloring_creation(_ptr(e2n), _ptr(part), _ptr(perm), nE, int(nbNodes),
                                     _ptr(c2e), _ptr(nc)), "coloring_creation")
    return part[:nE], perm[:nE], c2e[:int(nc[0]) + 1].copy()


def permute_2d(tab, perm, dim):
    tab = np.array(tab, copy=True)
    n = tab.size // dim
    if tab.dtype == np.float64:
        check(lib().mf_permute_double_2d(_ptr(tab), _ptr(_c(perm, np.int32)), n, dim), "permute")
    else:
        tab = _c(tab, np.int32)
        check(lib().mf_permute_int_2d(_ptr(tab), _ptr(_c(perm, np.int32)), n, dim), "permute")
    return tab


def boundary_mask(boundNodesCode, nbBoundNodes=None):
    code = _c(boundNodesCode, np.int32)
    nb = int(np.count_nonzero(code)) if nbBoundNodes is None else int(nbBoundNodes)
    cb = np.zeros(max(code.size * 3, 1), np.int32)
    check(lib().mf_boundary_mask(_ptr(code), code.size, nb, _ptr(cb)), "boundary_mask")
    return cb[:code.size * 3]


def compute_double_norm(x):
    x = _c(x, np.float64)
    return float(lib().mf_compute_double_norm(_ptr(x), x.size))


class Mesh:
    """A mesh after the reference's COLORING-build setup (src/main.cc:209-351):
    elements permuted by color, CSR built from the permuted elements, Dirichlet mask.

    Attributes: nbElem, nbNodes, nnz, coord, elemToNode (permuted), coef (permuted or
    None), colorPerm, colorToElem, row, col, checkBounds, operatorID, dim;
    elemToNode_original / coef_original keep the input element order.

    setup="device" runs the coloring, the permutations and the CSR pattern on the GPU
    (mf_dev_coloring_creation, mf_dev_permute_*, mf_dev_create_*) and copies the results
    back; they equal the host setup's bit for bit.
    """

    def __init__(self, coord, elemToNode, boundNodesCode, nbEdges, operator_id, coef=None,
                 colored=True, setup="host"):
        if setup not in ("host", "device"):
            raise ValueError(f"setup must be 'host' or 'device', got {setup!r}")
        self.operatorID = int(operator_id)
        self.dim = 9 if self.operatorID == OP_ELA else 1
        self.coord = _c(coord, np.float64)
        e2n = _c(elemToNode, np.int32)
        self.elemToNode_original = e2n
        self.coef_original = None if coef is None else _c(coef, np.float64)
        self.nbElem = e2n.size // 4
        self.nbNodes = self.coord.size // 3
        if colored and setup == "device":
            self._device_setup(e2n, coef, nbEdges)
        elif colored:
            _, self.colorPerm, self.colorToElem = coloring_creation(e2n, self.nbNodes)
            e2n = permute_2d(e2n, self.colorPerm, 4)
            if coef is not None:
                coef = permute_2d(_c(coef, np.float64), self.colorPerm, 2 if self.dim == 9 else 1)
        else:
            self.colorPerm = np.arange(self.nbElem, dtype=np.int32)
            self.colorToElem = np.array([0, self.nbElem], np.int32)
        if not (colored and setup == "device"):
            self.elemToNode = e2n
            self.coef = None if coef is None else _c(coef, np.float64)
            self.row, self.col = create_nodeToNode(e2n, self.nbNodes, nbEdges)
        self.nnz = int(self.row[-1])
        self.boundNodesCode = _c(boundNodesCode, np.int32)
        self.checkBounds = boundary_mask(self.boundNodesCode)

    def _device_setup(self, e2n, coef, nbEdges):
        from . import gpu
        d_e2n = gpu.DeviceArray.from_numpy(e2n)
        _, d_perm, self.colorToElem = gpu.coloring_creation(d_e2n, self.nbNodes)
        gpu.permute_2d(d_e2n, d_perm, 4)
        self.coef = None
        if coef is not None:
            c = _c(coef, np.float64)
            d_coef = gpu.DeviceArray.from_numpy(c)
            gpu.permute_2d(d_coef, d_perm, c.size // self.nbElem)
            self.coef = d_coef.numpy()
        _, _, row, col, _, nnz = gpu.device_setup(d_e2n, self.nbNodes, nbEdges,
                                                  with_elem_to_edge=False)
        self.colorPerm = d_perm.numpy()[:self.nbElem].copy()
        self.elemToNode = d_e2n.numpy()
        self.row = row.numpy()
        self.col = col.numpy()[:nnz].copy()

    @property
    def nbColors(self):
        return len(self.colorToElem) - 1

    def permute_elements(self, per_element):
        """A per-element array (lap: nE values; ela: nE (lambda, mu) pairs) given in the
        original element order, moved to this mesh's color-permuted order."""
        a = _c(per_element, np.float64)
        return permute_2d(a, self.colorPerm, a.size // self.nbElem)

    def unpermute_elements(self, per_element):
        """The inverse of permute_elements: color-permuted order back to the original."""
        a = _c(per_element, np.float64)
        w = a.size // self.nbElem
        return np.ascontiguousarray(a.reshape(self.nbElem, w)[self.colorPerm].ravel())

    @classmethod
    def cube(cls, n, jitter, geom_seed, operator_id, shuffle_seed=0, coef_seed=None,
             colored=True):
        coord, e2n, code, nnz = generate_cube(n, jitter, geom_seed, shuffle_seed)
        coef = None
        if coef_seed is not None:
            coef = generate_coef(operator_id, e2n.size // 4, coef_seed)
        return cls(coord, e2n, code, nnz, operator_id, coef=coef, colored=colored)


def plan_host_stats(mesh, tile_lds_bytes=0, block_threads=0, max_tile_rows=0, use_coord=True,
                    partition_coord=None, id_chunks=0, stage_precision=0):
    """Builds and validates the tiled gather schedule on the host (mf_plan_host_stats).
    partition_coord: coordinates used only to group nodes into tiles (default: the mesh's).
    stage_precision: the precision mode the tiles are sized for (0: FP64)."""
    import ctypes
    from ._lib import PlanParams
    params = PlanParams(int(tile_lds_bytes), int(block_threads), int(max_tile_rows),
                        int(id_chunks), int(stage_precision))
    out = np.zeros(8, np.int64)
    pc = mesh.coord if partition_coord is None else _c(partition_coord, np.float64).ravel()
    if pc.size != 3 * mesh.nbNodes:
        raise ValueError(f"partition_coord has {pc.size} values, the mesh needs "
                         f"{3 * mesh.nbNodes} (x, y, z per node)")
    check(lib().mf_plan_host_stats(mesh.operatorID, mesh.nbElem, mesh.nbNodes, mesh.nnz,
                                   _ptr(mesh.elemToNode), _ptr(mesh.row), _ptr(mesh.col),
                                   _ptr(pc) if use_coord else None,
                                   ctypes.byref(params), out.ctypes.data), "mf_plan_host_stats")
    keys = ["tiles", "tile_elems", "tasks", "incidences", "plan_bytes", "lds_bytes",
            "block_threads", "max_rows"]
    return dict(zip(keys, (int(x) for x in out)))
