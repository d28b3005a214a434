"""GPU entry points over the C ABI.

Device memory, streams and events come from libminifem itself (mf_dev_alloc, ...), so
the process holds exactly one HIP runtime: the one libminifem is linked against. Do not
initialise another GPU runtime (e.g. torch.cuda) in the same process.

Mirrors the reference's per-iteration interface: assembly
(src/headers/assembly.h:61-74), prec_init / prec_inversion
(src/headers/preconditioner.h:27-32) and compute_double_norm (src/headers/FEM.h:26).
"""
import ctypes

import numpy as np

from ._lib import OP_ELA, PREC_FP64, check, lib

_dev_ready = False


def device_count():
    n = lib().mf_device_count()
    return max(n, 0)


def init(device=0):
    global _dev_ready
    check(lib().mf_device_init(int(device)), "mf_device_init")
    _dev_ready = True


def _ensure():
    if not _dev_ready:
        init(0)


class DeviceArray:
    """A flat device buffer of one numpy dtype, freed when garbage collected."""

    def __init__(self, n, dtype):
        _ensure()
        self.dtype = np.dtype(dtype)
        self.size = int(n)
        self.nbytes = self.size * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().mf_dev_alloc(ctypes.byref(p), self.nbytes), "mf_dev_alloc")
        self.ptr = p.value
        self._lib = lib()

    @classmethod
    def from_numpy(cls, a, dtype=None, stream=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.size, a.dtype)
        check(lib().mf_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes, _sp(stream)), "h2d")
        return d

    @classmethod
    def zeros(cls, n, dtype, stream=None):
        return cls.from_numpy(np.zeros(int(n), dtype), stream=stream)

    def numpy(self, stream=None):
        out = np.empty(self.size, self.dtype)
        check(lib().mf_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, _sp(stream)), "d2h")
        return out

    def copy_from(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.size
        check(lib().mf_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, _sp(stream)), "h2d")

    def fill(self, value, stream=None):
        self.copy_from(np.full(self.size, value, self.dtype), stream)

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p:
            self._lib.mf_dev_free(p)
            self.ptr = None


def _sp(stream):
    if stream is None:
        return None
    return stream.ptr if isinstance(stream, Stream) else stream


def _p(a):
    if a is None:
        return None
    if isinstance(a, (DeviceArray, DeviceView)):
        return a.ptr
    raise TypeError(f"expected a DeviceArray, got {type(a)}")


class DeviceView:
    """A typed window [offset, offset + n) into device memory another object owns (an
    arena sub-allocation): same interface as DeviceArray, never freed by itself."""

    def __init__(self, ptr, n, dtype, owner=None):
        self.ptr = int(ptr)
        self.size = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.size * self.dtype.itemsize
        self._owner = owner

    def numpy(self, stream=None):
        out = np.empty(self.size, self.dtype)
        check(lib().mf_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, _sp(stream)), "d2h")
        return out

    def fill(self, value, stream=None):
        a = np.full(self.size, value, self.dtype)
        check(lib().mf_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes, _sp(stream)), "h2d")


class HostArray:
    """Pinned host memory (mf_host_alloc) as a numpy array, for asynchronous copies."""

    def __init__(self, n, dtype):
        _ensure()
        self.dtype = np.dtype(dtype)
        self.size = int(n)
        self.nbytes = self.size * self.dtype.itemsize
        p = ctypes.c_void_p()
        check(lib().mf_host_alloc(ctypes.byref(p), max(self.nbytes, 1)), "mf_host_alloc")
        self.ptr = p.value
        self._lib = lib()
        buf = (ctypes.c_byte * max(self.nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=self.size)

    def __del__(self):
        if getattr(self, "ptr", None):
            self.array = None
            self._lib.mf_host_free(self.ptr)
            self.ptr = None


def copy_h2d_async(dst, src, stream):
    """dst: DeviceArray/DeviceView; src: HostArray (pinned), same byte size."""
    assert dst.nbytes == src.nbytes
    check(lib().mf_memcpy_h2d_async(dst.ptr, src.ptr, src.nbytes, _sp(stream)), "h2d async")


def copy_d2h_async(dst, src, stream):
    """dst: HostArray (pinned); src: DeviceArray/DeviceView, same byte size."""
    assert dst.nbytes == src.nbytes
    check(lib().mf_memcpy_d2h_async(dst.ptr, src.ptr, src.nbytes, _sp(stream)), "d2h async")


class Stream:
    def __init__(self):
        _ensure()
        p = ctypes.c_void_p()
        check(lib().mf_stream_create(ctypes.byref(p)), "mf_stream_create")
        self.ptr = p.value
        self._lib = lib()

    def sync(self):
        check(self._lib.mf_stream_sync(self.ptr), "mf_stream_sync")

    def wait(self, event):
        """Work enqueued on this stream after the call waits for `event`."""
        check(self._lib.mf_stream_wait_event(self.ptr, event.ptr), "mf_stream_wait_event")

    def __del__(self):
        if getattr(self, "ptr", None):
            self._lib.mf_stream_destroy(self.ptr)
            self.ptr = None


class Event:
    def __init__(self):
        _ensure()
        p = ctypes.c_void_p()
        check(lib().mf_event_create(ctypes.byref(p)), "mf_event_create")
        self.ptr = p.value
        self._lib = lib()

    def record(self, stream=None):
        check(self._lib.mf_event_record(self.ptr, _sp(stream)), "mf_event_record")

    def query(self):
        """True once the work recorded before the event has completed."""
        done = ctypes.c_int()
        check(self._lib.mf_event_query(self.ptr, ctypes.byref(done)), "mf_event_query")
        return bool(done.value)

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        check(self._lib.mf_event_elapsed_ms(self.ptr, end.ptr, ctypes.byref(ms)), "elapsed")
        return float(ms.value)

    def __del__(self):
        if getattr(self, "ptr", None):
            self._lib.mf_event_destroy(self.ptr)
            self.ptr = None


def synchronize():
    check(lib().mf_device_sync(), "mf_device_sync")


class Graph:
    """Work captured from a stream into an executable HIP graph, replayed as one
    submission (mf_graph_begin / mf_graph_end / mf_graph_launch)."""

    def __init__(self, stream, enqueue):
        """Captures everything `enqueue(stream)` puts on `stream`."""
        self._lib = lib()
        check(self._lib.mf_graph_begin(_sp(stream)), "mf_graph_begin")
        try:
            enqueue(stream)
        finally:
            p = ctypes.c_void_p()
            st = self._lib.mf_graph_end(_sp(stream), ctypes.byref(p))
        check(st, "mf_graph_end")
        self.ptr = p.value

    def launch(self, stream):
        check(self._lib.mf_graph_launch(self.ptr, _sp(stream)), "mf_graph_launch")

    def __del__(self):
        if getattr(self, "ptr", None):
            self._lib.mf_graph_destroy(self.ptr)
            self.ptr = None


def assembly(coord, values, row, col, elemToNode, elemToEdge, nbElem, nbEdges, operatorDim,
             operatorID, colorToElem, coef=None, precision=PREC_FP64, stream=None):
    """assembly(...) of the COLORING build on the GPU; colorToElem is a host array."""
    _ensure()
    c2e = np.ascontiguousarray(colorToElem, dtype=np.int32)
    check(lib().mf_assembly_ex(_p(coord), _p(coef), _p(values), _p(row), _p(col), _p(elemToNode),
                               _p(elemToEdge), int(nbElem), int(nbEdges), int(operatorDim),
                               int(operatorID), c2e.ctypes.data, len(c2e) - 1, int(precision),
                               _sp(stream)), "mf_assembly_ex")


def prec_init(prec, values, row, col, nbNodes, operatorDim, stream=None):
    _ensure()
    check(lib().mf_prec_init(_p(prec), _p(values), _p(row), _p(col), int(nbNodes),
                             int(operatorDim), _sp(stream)), "mf_prec_init")


def prec_inversion(prec, row, col, checkBounds, nbNodes, operatorID, singular=None, stream=None):
    _ensure()
    check(lib().mf_prec_inversion(_p(prec), _p(row), _p(col), _p(checkBounds), int(nbNodes),
                                  int(operatorID), _p(singular), _sp(stream)),
          "mf_prec_inversion")


def norm(x, stream=None):
    _ensure()
    out = np.zeros(1)
    check(lib().mf_norm(_p(x), x.size, out.ctypes.data, _sp(stream)), "mf_norm")
    return float(out[0])


def device_setup(elemToNode, nbNodes, nbEdges, stream=None, with_elem_to_edge=True):
    """nodeToElem, the CSR pattern and elemToEdge built on the GPU from a (color-permuted,
    1-based) elemToNode: bit-exact twins of the host setup. Returns DeviceArrays
    (index, value, row, col, elemToEdge or None) and the nnz."""
    _ensure()
    e2n = elemToNode if isinstance(elemToNode, (DeviceArray, DeviceView)) else \
        DeviceArray.from_numpy(np.asarray(elemToNode, np.int32))
    nE = e2n.size // 4
    index = DeviceArray(nbNodes + 1, np.int32)
    value = DeviceArray(max(4 * nE, 1), np.int32)
    check(lib().mf_dev_create_nodeToElem(e2n.ptr, nE, int(nbNodes), index.ptr, value.ptr,
                                         _sp(stream)), "mf_dev_create_nodeToElem")
    row = DeviceArray(nbNodes + 1, np.int32)
    col = DeviceArray(max(int(nbEdges), 1), np.int32)
    nnz = check(lib().mf_dev_create_nodeToNode(row.ptr, col.ptr, index.ptr, value.ptr, e2n.ptr,
                                               int(nbNodes), int(nbEdges), _sp(stream)),
                "mf_dev_create_nodeToNode")
    if not with_elem_to_edge:
        return index, value, row, col, None, nnz
    e2e = DeviceArray(max(16 * nE, 1), np.int32)
    check(lib().mf_dev_create_elemToEdge(row.ptr, col.ptr, e2n.ptr, e2e.ptr, nE, _sp(stream)),
          "mf_dev_create_elemToEdge")
    return index, value, row, col, e2e, nnz


def coloring_creation(elemToNode, nbNodes, stream=None):
    """Greedy first-fit coloring on the GPU (mf_dev_coloring_creation), equal to the host
    scan: (colorPart, colorPerm) as DeviceArrays, colorToElem as a numpy array of
    nbColors + 1 offsets."""
    _ensure()
    e2n = elemToNode if isinstance(elemToNode, (DeviceArray, DeviceView)) else \
        DeviceArray.from_numpy(np.asarray(elemToNode, np.int32))
    nE = e2n.size // 4
    part = DeviceArray(max(nE, 1), np.int32)
    perm = DeviceArray(max(nE, 1), np.int32)
    c2e = DeviceArray(129, np.int32)
    nc = ctypes.c_int()
    check(lib().mf_dev_coloring_creation(e2n.ptr, part.ptr, perm.ptr, nE, int(nbNodes), c2e.ptr,
                                         ctypes.byref(nc), _sp(stream)), "mf_dev_coloring_creation")
    return part, perm, c2e.numpy()[:nc.value + 1].copy()


def permute_2d(tab, perm, dim, stream=None):
    """Row i of the device array tab moves to row perm[i] (mf_dev_permute_*_2d)."""
    n = tab.size // dim
    fn = lib().mf_dev_permute_int_2d if np.dtype(tab.dtype) == np.int32 else \
        lib().mf_dev_permute_double_2d
    check(fn(tab.ptr, perm.ptr, n, int(dim), _sp(stream)), "mf_dev_permute_2d")


def pcg(operator_id, nbNodes, row, col, values, prec, checkBounds, b, x, rtol=1e-10,
        max_iter=1000, stream=None, order=None):
    """Block-Jacobi preconditioned CG on the assembled operator (mf_pcg); x is updated in
    place. order (device int32, e.g. Plan.row_order()): solve in that node numbering
    (mf_pcg_ordered). Returns (iterations, relative residual)."""
    _ensure()
    it = ctypes.c_int()
    rr = ctypes.c_double()
    if order is None:
        check(lib().mf_pcg(int(operator_id), int(nbNodes), _p(row), _p(col), _p(values), _p(prec),
                           _p(checkBounds), _p(b), _p(x), float(rtol), int(max_iter),
                           ctypes.byref(it), ctypes.byref(rr), _sp(stream)), "mf_pcg")
    else:
        check(lib().mf_pcg_ordered(int(operator_id), int(nbNodes), _p(row), _p(col), _p(values),
                                   _p(prec), _p(checkBounds), _p(order), _p(b), _p(x), float(rtol),
                                   int(max_iter), ctypes.byref(it), ctypes.byref(rr), _sp(stream)),
              "mf_pcg_ordered")
    return it.value, rr.value


class DeviceMesh:
    """A Mesh's arrays resident on the GPU, with value/prec buffers."""

    def __init__(self, mesh, with_elem_to_edge=False):
        _ensure()
        self.mesh = mesh
        up = DeviceArray.from_numpy
        self.coord = up(mesh.coord)
        self.elemToNode = up(mesh.elemToNode)
        self.row = up(mesh.row)
        self.col = up(mesh.col)
        self.checkBounds = up(mesh.checkBounds)
        self.coef = None if mesh.coef is None else up(mesh.coef)
        self.elemToEdge = None
        if with_elem_to_edge:
            from .setup import create_elemToEdge
            self.elemToEdge = up(create_elemToEdge(mesh.row, mesh.col, mesh.elemToNode))
        self.values = DeviceArray(mesh.nnz * mesh.dim, np.float64)
        self.prec = DeviceArray(mesh.nbNodes * mesh.dim, np.float64)

    def iterate_colored(self, precision=PREC_FP64, use_coef=True, stream=None):
        """One FEM iteration (src/FEM.cc:183-232) with the drop-in colored kernels."""
        m = self.mesh
        assembly(self.coord, self.values, self.row, self.col, self.elemToNode, self.elemToEdge,
                 m.nbElem, m.nnz, m.dim, m.operatorID, m.colorToElem,
                 coef=self.coef if use_coef else None, precision=precision, stream=stream)
        prec_init(self.prec, self.values, self.row, self.col, m.nbNodes, m.dim, stream=stream)
        prec_inversion(self.prec, self.row, self.col, self.checkBounds, m.nbNodes, m.operatorID,
                       stream=stream)


def _job_table(jobs):
    from ._lib import MAX_PLAN_JOBS, PlanJob
    jobs = list(jobs)
    if not 1 <= len(jobs) <= MAX_PLAN_JOBS:
        raise ValueError(f"a multi-plan launch takes 1..{MAX_PLAN_JOBS} jobs, got {len(jobs)}")
    arr = (PlanJob * len(jobs))()
    for k, j in enumerate(jobs):
        plan, coord, coef, values, prec = j[:5]
        singular = j[5] if len(j) > 5 else None
        arr[k] = PlanJob(plan.ptr, _p(coord), _p(coef), _p(values), _p(prec), _p(singular))
    return arr, len(jobs)


def plans_assemble(jobs, precision=PREC_FP64, stream=None):
    """One persistent launch over the tiles of several plans (mf_plans_assemble), lap and
    ela mixed. jobs: (plan, coord, coef, values, prec[, singular]) per plan, the arguments
    Plan.assemble takes; each job's results equal its own Plan.assemble bit for bit."""
    arr, n = _job_table(jobs)
    check(lib().mf_plans_assemble(ctypes.cast(arr, ctypes.c_void_p), n, int(precision),
                                  _sp(stream)), "mf_plans_assemble")


def plans_lds_bytes(jobs, precision=PREC_FP64):
    """LDS bytes per workgroup of that launch, or None when the jobs cannot run as one
    launch in this mode (their tiles do not fit one CU together, their workgroup sizes
    differ, or a plan sized for the float stage is asked for FP64)."""
    arr, n = _job_table(jobs)
    r = lib().mf_plans_lds_bytes(ctypes.cast(arr, ctypes.c_void_p), n, int(precision))
    if r == -1:
        return None
    return check(r, "mf_plans_lds_bytes")


class Plan:
    """A mesh topology resident on the GPU with its tiled gather schedule
    (mf_plan_create). assemble() runs one fused iteration: values + inverted prec."""

    def __init__(self, mesh, tile_lds_bytes=0, block_threads=0, max_tile_rows=0,
                 use_coord=True, partition_coord=None, id_chunks=0, stage_precision=PREC_FP64):
        """partition_coord: coordinates used only to group nodes into tiles (default: the
        mesh's own). stage_precision: the precision mode the tiles are sized for; a plan
        sized for a mixed mode (float stage, larger tiles) runs the mixed modes only."""
        from ._lib import PlanParams
        _ensure()
        self._lib = lib()
        self.mesh = mesh
        self.stage_precision = int(stage_precision)
        params = PlanParams(int(tile_lds_bytes), int(block_threads), int(max_tile_rows),
                            int(id_chunks), self.stage_precision)
        pc = None
        if use_coord:
            pc = mesh.coord if partition_coord is None else np.ascontiguousarray(
                partition_coord, np.float64).ravel()
            if pc.size != 3 * mesh.nbNodes:
                raise ValueError(f"partition_coord has {pc.size} values, the mesh needs "
                                 f"{3 * mesh.nbNodes} (x, y, z per node)")
        p = ctypes.c_void_p()
        check(self._lib.mf_plan_create(
            ctypes.byref(p), mesh.operatorID, mesh.nbElem, mesh.nbNodes, mesh.nnz,
            mesh.elemToNode.ctypes.data, mesh.row.ctypes.data, mesh.col.ctypes.data,
            mesh.checkBounds.ctypes.data, None if pc is None else pc.ctypes.data,
            ctypes.byref(params)), "mf_plan_create")
        self.ptr = p.value

    def stats(self):
        out = np.zeros(8, np.int64)
        check(self._lib.mf_plan_stats(self.ptr, out.ctypes.data), "mf_plan_stats")
        keys = ["tiles", "tile_elems", "tasks", "incidences", "plan_bytes", "lds_bytes",
                "block_threads", "max_rows"]
        return dict(zip(keys, (int(x) for x in out)))

    def row_order(self, stream=None):
        """The plan's node order as a device int32 array (mf_plan_row_order)."""
        out = DeviceArray(max(self.mesh.nbNodes, 1), np.int32)
        check(self._lib.mf_plan_row_order(self.ptr, out.ptr, _sp(stream)), "mf_plan_row_order")
        return out

    def assemble(self, coord, coef, values, prec, precision=PREC_FP64, singular=None,
                 stream=None):
        check(self._lib.mf_plan_assemble(self.ptr, _p(coord), _p(coef), _p(values), _p(prec),
                                         int(precision), _p(singular), _sp(stream)),
              "mf_plan_assemble")

    def write_probe(self, values, mode, stream=None):
        """Zeros over `values` with the assembly's output stream alone (mode 0: tile rows
        in the kernel's order; mode 1: front to back). Measurement only."""
        check(self._lib.mf_plan_write_probe(self.ptr, _p(values), int(mode), _sp(stream)),
              "mf_plan_write_probe")

    def __del__(self):
        if getattr(self, "ptr", None):
            self._lib.mf_plan_destroy(self.ptr)
            self.ptr = None
