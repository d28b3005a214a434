"""The allocation-free norm (mf_norm_noalloc) through the C ABI: the same result as
mf_norm bit for bit (same kernels, same order), on two calls in a row on the shared
per-GPU scratch, the second with a smaller input, and within 1e-13 of numpy's norm."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_norm_noalloc_equals_norm(mf, gpu):
    lib = mf.lib()
    rng = np.random.default_rng(11)
    for n in (3_000_017, 1_001):  # the second call is smaller: no stale partial may remain
        print(f"[noalloc] n={n}", flush=True)
        x = rng.standard_normal(n) * 10.0
        d = gpu.DeviceArray.from_numpy(x)
        a, b = ctypes.c_double(-1.0), ctypes.c_double(-1.0)
        mf.check(lib.mf_norm(d.ptr, n, ctypes.byref(a), None), "mf_norm")
        mf.check(lib.mf_norm_noalloc(d.ptr, n, ctypes.byref(b), None), "mf_norm_noalloc")
        assert a.value == b.value, (n, a.value, b.value)
        ref = float(np.linalg.norm(x))
        assert abs(b.value - ref) <= 1e-13 * ref, (n, b.value, ref)
        del d
    z = ctypes.c_double(-1.0)
    mf.check(lib.mf_norm_noalloc(None, 0, ctypes.byref(z), None), "mf_norm_noalloc")
    assert z.value == 0.0
    from minifem._lib import MF_ERR_ARG
    assert lib.mf_norm_noalloc(None, -1, ctypes.byref(z), None) == MF_ERR_ARG
