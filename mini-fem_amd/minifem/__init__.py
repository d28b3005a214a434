"""minifem: MI355X-native Mini-FEM assembly + block-Jacobi preconditioner.

Host setup (minifem.setup) and GPU entry points (minifem.gpu, imported on demand) over
libminifem.so's C ABI.
"""
from ._lib import (LIB_PATH, MF_MAX_COLOR, OP_ELA, OP_LAP, PREC_FP64, PREC_MIXED,
                   PREC_MIXED_F32COEF, MiniFemError, check, lib)
from .setup import (Mesh, boundary_mask, coloring_creation, compute_double_norm, create_elemToEdge,
                    create_nodeToElem, create_nodeToNode, cube_sizes, generate_coef, generate_cube,
                    permute_2d, read_header, read_input_data, read_ref_assembly, store_input_data,
                    store_ref_assembly)

__all__ = [
    "LIB_PATH", "MF_MAX_COLOR", "OP_ELA", "OP_LAP", "PREC_FP64", "PREC_MIXED",
    "PREC_MIXED_F32COEF", "MiniFemError",
    "check", "lib", "Mesh", "boundary_mask", "coloring_creation", "compute_double_norm",
    "create_elemToEdge", "create_nodeToElem", "create_nodeToNode", "cube_sizes", "generate_coef",
    "generate_cube", "permute_2d", "read_header", "read_input_data", "read_ref_assembly",
    "store_input_data", "store_ref_assembly",
]
