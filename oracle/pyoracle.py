"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker / host-core baseline, never by the product.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
BIN_PATH = os.path.join(HERE, "bin", "minifem_oracle")

_i, _i64, _d, _p, _s = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_char_p
_SIG = {
    "or_create_nodeToElem": (_i, [_p, _i, _i, _p, _p]),
    "or_create_nodeToNode": (_i, [_p, _p, _p, _p, _p, _i, _i]),
    "or_create_elemToEdge": (None, [_p, _p, _p, _p, _i]),
    "or_coloring_creation": (_i, [_p, _p, _p, _i, _i, _p, _p]),
    "or_permute_int_2d": (None, [_p, _p, _i, _i]),
    "or_permute_double_2d": (None, [_p, _p, _i, _i]),
    "or_dqmrd4": (_i, [_i, _p, _p, _p]),
    "or_e_essbcm": (None, [_i, _i, _p, _p, _p]),
    "or_elem_coef": (None, [_p, _p, _p, _i]),
    "or_assembly": (None, [_p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i]),
    "or_assembly_coloring": (None, [_p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i]),
    "or_assembly_mixed": (None, [_p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i]),
    "or_prec_init": (None, [_p, _p, _p, _p, _i, _i, _i]),
    "or_prec_inversion": (_i, [_p, _p, _p, _p, _i, _i, _i]),
    "or_ela_invert_prec": (_i, [_i, _p, _p, _p, _p, _i]),
    "or_compute_double_norm": (_d, [_p, _i64]),
    "or_read_header": (_i, [_s, _p]),
    "or_read_input_data": (_i, [_s, _p, _p, _p, _p, _p, _p]),
    "or_fem_loop": (_i, [_p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _i, _i, _p, _p]),
    "or_num_threads": (_i, []),
}
_LIB = None


def build():
    subprocess.run(["make", "-B", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _LIB = h
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def nodeToElem(e2n, nN):
    e2n = _c(e2n, np.int32)
    idx = np.zeros(nN + 1, np.int32)
    val = np.zeros(max(e2n.size, 1), np.int32)
    assert lib().or_create_nodeToElem(_ptr(e2n), e2n.size // 4, nN, _ptr(idx), _ptr(val)) == 0
    return idx, val[:e2n.size]


def csr(e2n, nN, cap):
    e2n = _c(e2n, np.int32)
    idx, val = nodeToElem(e2n, nN)
    row = np.zeros(nN + 1, np.int32)
    col = np.zeros(max(cap, 1), np.int32)
    nnz = lib().or_create_nodeToNode(_ptr(row), _ptr(col), _ptr(idx), _ptr(val), _ptr(e2n), nN, cap)
    if nnz < 0:
        raise ValueError(f"or_create_nodeToNode status {nnz}")
    return row, col[:nnz]


def elemToEdge(row, col, e2n):
    e2n = _c(e2n, np.int32)
    out = np.full(max(e2n.size * 4, 1), -1, np.int32)
    lib().or_create_elemToEdge(_ptr(_c(row, np.int32)), _ptr(_c(col, np.int32)), _ptr(e2n), _ptr(out),
                               e2n.size // 4)
    return out[:e2n.size * 4]


def coloring(e2n, nN):
    e2n = _c(e2n, np.int32)
    nE = e2n.size // 4
    part = np.zeros(max(nE, 1), np.int32)
    perm = np.zeros(max(nE, 1), np.int32)
    c2e = np.zeros(129, np.int32)
    nc = np.zeros(1, np.int32)
    st = lib().or_coloring_creation(_ptr(e2n), _ptr(part), _ptr(perm), nE, nN, _ptr(c2e), _ptr(nc))
    if st != 0:
        raise ValueError(f"or_coloring_creation status {st}")
    return part[:nE], perm[:nE], c2e[:nc[0] + 1].copy()


def permute(tab, perm, dim):
    tab = np.array(tab, copy=True)
    n = tab.size // dim
    if tab.dtype == np.float64:
        lib().or_permute_double_2d(_ptr(tab), _ptr(_c(perm, np.int32)), n, dim)
    else:
        tab = _c(tab, np.int32)
        lib().or_permute_int_2d(_ptr(tab), _ptr(_c(perm, np.int32)), n, dim)
    return tab


def boundary_mask(code, nb=None):
    code = _c(code, np.int32)
    N = code.size
    cap = np.array([int(np.count_nonzero(code)) if nb is None else nb], np.int32)
    blist = np.zeros(max(int(cap[0]), 1), np.int32)
    lib().or_dqmrd4(N, _ptr(code), _ptr(cap), _ptr(blist))
    cb = np.zeros(max(3 * N, 1), np.int32)
    lib().or_e_essbcm(N, int(min(cap[0], blist.size)), _ptr(blist), _ptr(code), _ptr(cb))
    return cb[:3 * N]


def elem_coef(coord, e2n, elem):
    g = np.zeros((4, 3))
    lib().or_elem_coef(_ptr(g), _ptr(_c(coord, np.float64)), _ptr(_c(e2n, np.int32)), int(elem))
    return g


def assembly(coord, row, col, e2n, op, coef=None, e2e=None, mixed=False):
    """REF-order assembly (src/assembly.cc:647-659) over the given element order."""
    dim = 9 if op == 1 else 1
    nnz = int(row[-1])
    vals = np.zeros(max(nnz * dim, 1))
    fn = lib().or_assembly_mixed if mixed else lib().or_assembly
    fn(_ptr(_c(coord, np.float64)), _ptr(vals), _ptr(_c(row, np.int32)), _ptr(_c(col, np.int32)),
       _ptr(_c(e2n, np.int32)), _ptr(_c(e2e, np.int32)), _ptr(_c(coef, np.float64)),
       np.asarray(e2n).size // 4, nnz, dim, op)
    return vals[:nnz * dim]


def assembly_coloring(coord, row, col, e2n, op, c2e, coef=None, parallel=False):
    dim = 9 if op == 1 else 1
    nnz = int(row[-1])
    vals = np.zeros(max(nnz * dim, 1))
    c2e = _c(c2e, np.int32)
    lib().or_assembly_coloring(_ptr(_c(coord, np.float64)), _ptr(vals), _ptr(_c(row, np.int32)),
                               _ptr(_c(col, np.int32)), _ptr(_c(e2n, np.int32)), None,
                               _ptr(_c(coef, np.float64)), _ptr(c2e), c2e.size - 1, nnz, dim, op,
                               int(parallel))
    return vals[:nnz * dim]


def prec(values, row, col, cb, nN, op):
    """prec_init + prec_inversion (src/preconditioner.cc:25-87). Returns (prec, nsingular)."""
    dim = 9 if op == 1 else 1
    p = np.zeros(max(nN * dim, 1))
    lib().or_prec_init(_ptr(p), _ptr(_c(values, np.float64)), _ptr(_c(row, np.int32)),
                       _ptr(_c(col, np.int32)), nN, dim, 0)
    ns = lib().or_prec_inversion(_ptr(p), _ptr(_c(row, np.int32)), _ptr(_c(col, np.int32)),
                                 _ptr(_c(cb, np.int32)), nN, op, 0)
    return p[:nN * dim], ns


def norm(x):
    x = _c(x, np.float64)
    return float(lib().or_compute_double_norm(_ptr(x), x.size))


def fem_loop(coord, row, col, e2n, cb, c2e, op, nN, iters, mode, coef=None):
    """Timed host FEM loop (src/FEM.cc:177-257). mode 0 ref, 1 coloring-omp as written,
    2 coloring-omp with per-color parallel elements. Returns (values, prec, times_us[4])."""
    dim = 9 if op == 1 else 1
    nnz = int(row[-1])
    vals = np.zeros(nnz * dim)
    p = np.zeros(nN * dim)
    t = np.zeros(4)
    cy = np.zeros(4, np.uint64)
    c2e = _c(c2e, np.int32)
    e2n = _c(e2n, np.int32)
    lib().or_fem_loop(_ptr(_c(coord, np.float64)), _ptr(vals), _ptr(p), _ptr(_c(row, np.int32)),
                      _ptr(_c(col, np.int32)), _ptr(e2n), None, _ptr(_c(coef, np.float64)),
                      _ptr(_c(cb, np.int32)), _ptr(c2e), c2e.size - 1, e2n.size // 4, nN, nnz, op,
                      iters, mode, _ptr(t), _ptr(cy))
    return vals, p, t


def num_threads():
    return int(lib().or_num_threads())


class OracleMesh:
    """The oracle's own setup pipeline (src/main.cc:209-351) for a mesh given as arrays."""

    def __init__(self, coord, e2n, code, nnz_cap, op, coef=None, colored=True):
        self.op = op
        self.dim = 9 if op == 1 else 1
        self.coord = np.ascontiguousarray(coord, np.float64)
        self.nN = self.coord.size // 3
        e2n = np.ascontiguousarray(e2n, np.int32)
        self.nE = e2n.size // 4
        if colored:
            self.part, self.perm, self.c2e = coloring(e2n, self.nN)
            e2n = permute(e2n, self.perm, 4)
            if coef is not None:
                coef = permute(np.asarray(coef, np.float64), self.perm, 2 if op == 1 else 1)
        else:
            self.perm = np.arange(self.nE, dtype=np.int32)
            self.c2e = np.array([0, self.nE], np.int32)
        self.e2n = e2n
        self.coef = coef
        self.row, self.col = csr(e2n, self.nN, nnz_cap)
        self.nnz = int(self.row[-1])
        self.cb = boundary_mask(code)

    def assemble(self, use_coef=True, mixed=False):
        return assembly(self.coord, self.row, self.col, self.e2n, self.op,
                        coef=self.coef if use_coef else None, mixed=mixed)

    def prec(self, values):
        return prec(values, self.row, self.col, self.cb, self.nN, self.op)
