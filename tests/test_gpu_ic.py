"""The particle lattice and slump geometry built on the device
(ep_slump_points: mpts_populate.jl:26-56 + get_slump.jl:18-95) against the
host restatement: the same points, bit for bit, in the same order."""
from __future__ import annotations

import numpy as np
import pytest

import elastoplasm as ep
from cases import L2, L3

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,nel,geometry", [
    (2, [40, 10], "wedge"), (2, [160, 40], "wedge"), (3, [40, 10, 10], "wedge"), (2, [40, 10], "column"),
    (3, [64, 16, 32], "column")])
def test_device_lattice_matches_host(dim, nel, geometry):
    L = L2 if dim == 2 else L3
    kw = dict(geometry=geometry, column_width=0.4)
    a, _ = ep.ic_slump(L, nel, **kw)
    b, _ = ep.ic_slump(L, nel, device=True, **kw)
    assert a.mpts.nmp == b.mpts.nmp > 0
    np.testing.assert_array_equal(a.mpts.x, b.mpts.x)
    np.testing.assert_array_equal(a.mpts.s.phi, b.mpts.s.phi)
    if dim == 2 and nel == [40, 10] and geometry == "wedge":
        assert a.mpts.nmp == 591           # the reference's 2D slump (test_slump.jl:62-64)


def test_device_lattice_reports_count_beyond_cap():
    import ctypes as C
    from elastoplasm import hip
    ic, _ = ep.ic_slump(L2, [40, 10])
    Lb = hip.load()
    dp = C.POINTER(C.c_double)
    Lb.ep_slump_points.argtypes = [C.c_int32, C.POINTER(dp), C.POINTER(C.c_int64), C.c_int32, dp, C.c_int64, dp,
                                   C.POINTER(C.c_int64)]
    ax = [np.linspace(0.5, 9.5, 10), np.linspace(0.5, 4.5, 5)]
    ap = (dp * 3)(ax[0].ctypes.data_as(dp), ax[1].ctypes.data_as(dp), None)
    na = (C.c_int64 * 3)(10, 5, 1)
    par = np.array([100.0, 100.0, 0, 0, 0])
    n = C.c_int64()
    out = np.zeros((2, 10))
    r = Lb.ep_slump_points(2, ap, na, 1, par.ctypes.data_as(dp), 10, out.ctypes.data_as(dp), C.byref(n))
    assert r == -1 and n.value == 50          # EP_E_ARG, with the count
