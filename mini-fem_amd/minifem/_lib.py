"""ctypes binding of libminifem.so (the C ABI declared in include/minifem.h).

The library is built in-tree by ``make -C mini-fem_amd`` (``__graft_entry__.build()``)
and loaded from this directory only; a missing library is an error, never a fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libminifem.so")

MF_OK = 0
MF_ERR_ARG = -1
MF_ERR_IO = -2
MF_ERR_COLORS = -3
MF_ERR_NNZ = -4
MF_ERR_HIP = -5
MF_ERR_NOMEM = -6
MF_ERR_BUDGET = -7
MF_ERR_STATE = -8
MF_MAX_COLOR = 128
OP_LAP = 0
OP_ELA = 1
PREC_FP64 = 0
PREC_MIXED = 1
PREC_MIXED_F32COEF = 2  # mixed, with the coefficient field given as float32 (plans only)

_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_d = ctypes.c_double
_p = ctypes.c_void_p
_s = ctypes.c_char_p

# name -> (restype, argtypes)
SIGNATURES = {
    "mf_status_string": (_s, [_i]),
    "mf_version": (_i, []),
    "mf_read_header": (_i, [_s, _p]),
    "mf_read_input_data": (_i, [_s, _p, _p, _p, _p, _p, _p]),
    "mf_store_input_data": (_i, [_s, _p, _p, _p, _p, _p, _p, _p]),
    "mf_read_ref_assembly": (_i, [_s, _p, _p]),
    "mf_store_ref_assembly": (_i, [_s, _d, _d]),
    "mf_cube_sizes": (_i, [_i, _p, _p, _p]),
    "mf_generate_cube": (_i, [_i, _d, _u64, _u64, _p, _p, _p]),
    "mf_generate_coef": (_i, [_i, _i, _u64, _p]),
    "mf_create_nodeToElem": (_i, [_p, _i, _i, _p, _p]),
    "mf_create_nodeToNode": (_i, [_p, _p, _p, _p, _p, _i, _i]),
    "mf_create_elemToEdge": (_i, [_p, _p, _p, _p, _i]),
    "mf_coloring_creation": (_i, [_p, _p, _p, _i, _i, _p, _p]),
    "mf_permute_int_2d": (_i, [_p, _p, _i, _i]),
    "mf_permute_double_2d": (_i, [_p, _p, _i, _i]),
    "mf_boundary_mask": (_i, [_p, _i, _i, _p]),
    "mf_compute_double_norm": (_d, [_p, _i64]),
    "mf_device_count": (_i, []),
    "mf_device_init": (_i, [_i]),
    "mf_assembly": (_i, [_p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _p, _i, _p]),
    "mf_assembly_ex": (_i, [_p, _p, _p, _p, _p, _p, _p, _i, _i, _i, _i, _p, _i, _i, _p]),
    "mf_prec_init": (_i, [_p, _p, _p, _p, _i, _i, _p]),
    "mf_prec_inversion": (_i, [_p, _p, _p, _p, _i, _i, _p, _p]),
    "mf_norm": (_i, [_p, _i64, _p, _p]),
    "mf_norm_noalloc": (_i, [_p, _i64, _p, _p]),
    "mf_dev_alloc": (_i, [_p, ctypes.c_size_t]),
    "mf_dev_free": (_i, [_p]),
    "mf_memcpy_h2d": (_i, [_p, _p, ctypes.c_size_t, _p]),
    "mf_memcpy_d2h": (_i, [_p, _p, ctypes.c_size_t, _p]),
    "mf_memcpy_d2d": (_i, [_p, _p, ctypes.c_size_t, _p]),
    "mf_stream_create": (_i, [_p]),
    "mf_stream_destroy": (_i, [_p]),
    "mf_stream_sync": (_i, [_p]),
    "mf_device_sync": (_i, []),
    "mf_event_create": (_i, [_p]),
    "mf_event_destroy": (_i, [_p]),
    "mf_event_record": (_i, [_p, _p]),
    "mf_event_elapsed_ms": (_i, [_p, _p, _p]),
    "mf_hbm_copy": (_i, [_p, _p, ctypes.c_size_t, _p]),
    "mf_gather_probe": (_i, [_p, _p, _p, _i64, _p]),
    "mf_dev_create_nodeToElem": (_i, [_p, _i, _i, _p, _p, _p]),
    "mf_dev_create_nodeToNode": (_i, [_p, _p, _p, _p, _p, _i, _i, _p]),
    "mf_dev_create_elemToEdge": (_i, [_p, _p, _p, _p, _i, _p]),
    "mf_dev_coloring_creation": (_i, [_p, _p, _p, _i, _i, _p, _p, _p]),
    "mf_dev_permute_int_2d": (_i, [_p, _p, _i, _i, _p]),
    "mf_dev_permute_double_2d": (_i, [_p, _p, _i, _i, _p]),
    "mf_pcg": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p, ctypes.c_double, _i, _p, _p, _p]),
    "mf_pcg_ordered": (_i, [_i, _i, _p, _p, _p, _p, _p, _p, _p, _p, ctypes.c_double, _i, _p, _p,
                            _p]),
    "mf_host_alloc": (_i, [_p, ctypes.c_size_t]),
    "mf_host_free": (_i, [_p]),
    "mf_memcpy_h2d_async": (_i, [_p, _p, ctypes.c_size_t, _p]),
    "mf_memcpy_d2h_async": (_i, [_p, _p, ctypes.c_size_t, _p]),
    "mf_event_query": (_i, [_p, _p]),
    "mf_stream_wait_event": (_i, [_p, _p]),
    "mf_graph_begin": (_i, [_p]),
    "mf_graph_end": (_i, [_p, _p]),
    "mf_graph_launch": (_i, [_p, _p]),
    "mf_graph_destroy": (_i, [_p]),
    "mf_plan_create": (_i, [_p, _i, _i, _i, _i, _p, _p, _p, _p, _p, _p]),
    "mf_plan_destroy": (_i, [_p]),
    "mf_plan_stats": (_i, [_p, _p]),
    "mf_plan_row_order": (_i, [_p, _p, _p]),
    "mf_plan_assemble": (_i, [_p, _p, _p, _p, _p, _i, _p, _p]),
    "mf_plan_write_probe": (_i, [_p, _p, _i, _p]),
    "mf_plan_host_stats": (_i, [_i, _i, _i, _i, _p, _p, _p, _p, _p, _p]),
    "mf_plans_assemble": (_i, [_p, _i, _i, _p]),
    "mf_plans_lds_bytes": (_i, [_p, _i, _i]),
    "mf_server_create": (_i, [_p, _p, _i64, _i, _i, _i]),
    "mf_server_add_mesh": (_i, [_p, _p]),
    "mf_server_run": (_i, [_p, _p, _i, _p, _p, _p]),
    "mf_server_arena_used": (_i64, [_p]),
    "mf_server_destroy": (_i, [_p]),
}

MAX_PLAN_JOBS = 16
REQ_COEF = 0
REQ_COORD = 1


class PlanJob(ctypes.Structure):
    """mf_plan_job: one plan's iteration inside a multi-plan launch."""
    _fields_ = [("plan", _p), ("coord", _p), ("coef", _p), ("values", _p), ("prec", _p),
                ("singular", _p)]


class ServerMesh(ctypes.Structure):
    """mf_server_mesh: a resident mesh of the native request loop."""
    _fields_ = [("plan", _p), ("coord", _p), ("coef", _p), ("nvalues", _i64), ("nprec", _i64),
                ("ncoord", _i64), ("ncoef", _i64)]


class Request(ctypes.Structure):
    """mf_request: one served request."""
    _fields_ = [("mesh", ctypes.c_int), ("kind", ctypes.c_int), ("mode", ctypes.c_int),
                ("pad", ctypes.c_int), ("payload", _p), ("payload_bytes", _i64),
                ("t_arrival", _d), ("values_out", _p), ("prec_out", _p)]


class ServerStats(ctypes.Structure):
    _fields_ = [("launches", _i64), ("memory_waits", _i64), ("makespan_s", _d)]


class PlanParams(ctypes.Structure):
    _fields_ = [("tile_lds_bytes", ctypes.c_int), ("block_threads", ctypes.c_int),
                ("max_tile_rows", ctypes.c_int), ("id_chunks", ctypes.c_int),
                ("stage_precision", ctypes.c_int)]


class MiniFemError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = _LIB.mf_status_string(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else msg)


_LIB = None


def lib():
    """Load libminifem.so (once). Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C mini-fem_amd` "
                "(or __graft_entry__.build())")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def check(status, what=""):
    if status < 0:
        raise MiniFemError(status, what)
    return status
