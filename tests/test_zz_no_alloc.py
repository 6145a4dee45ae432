"""ep_grf_gauss_noalloc (the entry point that allocates nothing) against
ep_grf_gauss, through the C ABI, on two calls in a row, the second smaller.
Device buffers for the points and values come from the HIP runtime the
library is bound to (hipMalloc resolved through the library's handle), not
from the library's own entry points."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from elastoplasm import grf, hip

pytestmark = pytest.mark.gpu

EMU = "cpucheck" in os.environ.get("EP_HIP_LIB", "")
_dp = C.POINTER(C.c_double)


def _hip_runtime():
    # the HIP runtime the library itself is bound to (dlsym through the
    # library's handle searches its dependencies), never a second copy
    rt = hip.load()
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return rt


def _noalloc(L, rt, x, waves, ab):
    dim, n = x.shape
    nh = len(ab)
    dx, dg = C.c_void_p(), C.c_void_p()
    assert rt.hipMalloc(C.byref(dx), x.nbytes) == 0 and rt.hipMalloc(C.byref(dg), 8 * n) == 0
    try:
        assert rt.hipMemcpy(dx, x.ctypes.data, x.nbytes, 1) == 0          # hipMemcpyHostToDevice
        hip._check(L.ep_grf_gauss_noalloc(dim, n, C.cast(dx, _dp), nh, waves.ctypes.data_as(_dp),
                                          ab.ctypes.data_as(_dp), C.cast(dg, _dp)))
        g = np.zeros(n)
        assert rt.hipMemcpy(g.ctypes.data, dg, 8 * n, 2) == 0             # hipMemcpyDeviceToHost
    finally:
        rt.hipFree(dx)
        rt.hipFree(dg)
    return g


@pytest.mark.skipif(EMU, reason="device buffers from the HIP runtime: for the GPU")
def test_grf_noalloc_matches_grf_gauss_twice_then_smaller():
    L = hip.load()
    L.ep_grf_gauss_noalloc.argtypes = [C.c_int32, C.c_int64, _dp, C.c_int32, _dp, _dp, _dp]
    rt = _hip_runtime()
    rng = np.random.default_rng(5)
    for dim, n, nh in ((3, 100_003, 5000), (2, 7_001, 300)):
        x = np.asfortranarray(rng.uniform(0.0, 50.0, size=(dim, n)))
        want = grf.gauss_field(x, seed=n, ell=5.0, nh=nh)
        waves, ab = grf.harmonics(n, dim, 5.0, nh)
        got = _noalloc(L, rt, x, waves, ab)
        # the same kernel on the same inputs: identical bits
        np.testing.assert_array_equal(got, want)
        assert np.isfinite(got).all() and 0.5 < got.std() < 1.5
    # sizes beyond the static table and host pointers are refused
    x = np.asfortranarray(rng.uniform(size=(3, 10)))
    waves, ab = grf.harmonics(1, 3, 5.0, 16385)
    g = np.zeros(10)
    with pytest.raises(hip.EpError, match="EP_E_ARG"):
        hip._check(L.ep_grf_gauss_noalloc(3, 10, x.ctypes.data_as(_dp), 16385, waves.ctypes.data_as(_dp),
                                          ab.ctypes.data_as(_dp), g.ctypes.data_as(_dp)))
    waves, ab = grf.harmonics(1, 3, 5.0, 10)
    with pytest.raises(hip.EpError, match="device memory"):
        hip._check(L.ep_grf_gauss_noalloc(3, 10, x.ctypes.data_as(_dp), 10, waves.ctypes.data_as(_dp),
                                          ab.ctypes.data_as(_dp), g.ctypes.data_as(_dp)))
