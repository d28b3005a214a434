"""The C-ABI library loads, exports every function include/minifem.h declares, the
ctypes signatures cover the same set, and argument errors are reported as status codes
before any GPU work (runs without a GPU)."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(mf_\w+)\s*\(", src))
    return names


def test_exports_every_declared_symbol(mf):
    from minifem._lib import LIB_PATH, SIGNATURES
    names = declared_functions()
    assert len(names) > 20
    raw = ctypes.CDLL(LIB_PATH)
    for n in sorted(names):
        assert hasattr(raw, n), n
    assert names == set(SIGNATURES), names ^ set(SIGNATURES)


def test_status_strings(mf):
    lib = mf.lib()
    assert lib.mf_status_string(0) == b"ok"
    assert b"colors" in lib.mf_status_string(-3)
    assert lib.mf_version() >= 1


def test_device_entry_points_reject_bad_arguments(mf):
    lib = mf.lib()
    c2e = np.array([0, 4], np.int32)
    # bad operatorDim/operatorID pairing
    assert lib.mf_assembly(None, None, None, None, None, None, 4, 10, 9, 0, c2e.ctypes.data, 1,
                           None) == -1
    # color offsets that do not cover nbElem
    dummy = ctypes.c_void_p(16)
    assert lib.mf_assembly(dummy, dummy, dummy, dummy, dummy, None, 5, 10, 1, 0, c2e.ctypes.data,
                           1, None) == -1
    # too many colors
    assert lib.mf_assembly(dummy, dummy, dummy, dummy, dummy, None, 4, 10, 1, 0, c2e.ctypes.data,
                           129, None) == -1
    assert lib.mf_prec_init(None, None, None, None, -1, 1, None) == -1
    assert lib.mf_prec_inversion(None, None, None, None, 5, 3, None, None) == -1
    assert lib.mf_norm(None, -1, None, None) == -1
    # empty work is fine
    assert lib.mf_prec_init(None, None, None, None, 0, 1, None) == 0
