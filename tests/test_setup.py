"""Host-side setup mirror: Julia range semantics, ic_slump geometry, Dirichlet
masks, element-to-node table and B-spline node classification."""
from __future__ import annotations

import math

import numpy as np
import pytest

import elastoplasm as ep
from cases import slump
from elastoplasm.jrange import frange, rat


def test_julia_float_ranges():
    # Julia: collect(0:0.1:0.3)[end] === 0.3 and length(0.0:0.1:1.0) == 11
    assert frange(0.0, 0.1, 0.3)[-1] == 0.3
    assert len(frange(0.0, 0.1, 1.0)) == 11
    assert frange(0.0, 0.1, 1.0)[3] == 0.3
    assert list(frange(0.1, 0.1, 0.3)) == [0.1, 0.2, 0.3]
    # integral steps are exact
    assert np.array_equal(frange(0.0, 1.0, 5.0), np.arange(6.0))
    assert rat(0.75) == (3, 4)
    # the slump mesh: h = 64.1584/40, last node = L exactly
    x = frange(0.0, 64.1584 / 40, 64.1584)
    assert len(x) == 41 and x[-1] == 64.1584 and x[-2] == x[-1] - 64.1584 / 40


def test_ic_slump_2d_matches_reference_geometry():
    ic, cfg = slump(2)
    me, mp = ic.mesh, ic.mpts
    assert mp.nmp == 591                           # get_slump.jl:18-95 wedge for 40x10
    assert list(me.nno) == [41, 11, 451]
    assert me.h[0] == me.h[1] == 64.1584 / 40
    assert (mp.x[1] <= 12.8 - me.h[1] / 4).all()
    # weak layer: ϕ = ϕr below 2*wl (get_slump.jl:92)
    wl = 0.15 * 12.8
    assert np.all(mp.s.phi[mp.x[1] <= 2 * wl] == ic.cmpr.phir)
    assert np.all(mp.s.phi[mp.x[1] > 2 * wl] == ic.cmpr.phi0)
    # Ω₀ = (h/2)^2, ℓ₀ = h/4 (setup_mpts.jl:37-38)
    assert np.allclose(mp.Om0, (me.h[0] / 2) ** 2)
    assert np.allclose(mp.l0, me.h[0] / 4)
    # roller walls: normal component fixed on each wall (get_bc.jl)
    bc = me.bcs.status
    assert bc[0, me.x[0] == 0.0].all() and bc[0, me.x[0] == me.x[0].max()].all()
    assert bc[1, me.x[1] == 0.0].all() and not bc[1, me.x[1] == me.x[1][5]].any()
    assert bc.sum() == 2 * 11 + 2 * 41
    assert ic.cmpr.c == pytest.approx(math.sqrt((1e6 / 1.2 + 4 / 3 * 1e6 / 2.6) / 2700.0))


def test_ic_slump_3d_headroom_and_counts():
    ic, cfg = slump(3)
    me = ic.mesh
    assert list(me.nno[:3]) == [41, 11, 13]        # z gets +2h (get_coords.jl:42)
    assert ic.mpts.nmp == 11820
    ic2, _ = ep.ic_slump([128.0, 32.0, 32.0], [128, 32, 32], geometry="column", column_width=0.4,
                         basis={"which": "gimpm", "how": "undeformed", "ghost": True})
    assert ic2.mpts.nmp == 102 * 64 * 64           # x <= 51.2 m, z <= 32 m, 2x2x2 per cell


def _types_by_equality(axis, lo, hi, h):
    out = []
    for xn in axis:
        if xn == lo or xn == hi:
            out.append(1)
        elif xn == lo + h:
            out.append(2)
        elif lo + h < xn < hi - h:
            out.append(3)
        elif xn == hi - h:
            out.append(4)
        else:
            out.append(0)
    return out


@pytest.mark.parametrize("L,nel", [([64.1584, 16.0396], [40, 10]), ([64.1584, 16.0396, 16.0396], [40, 10, 10]),
                                   ([128.0, 32.0, 32.0], [128, 32, 32]), ([256.0, 64.0, 128.0], [256, 64, 128])])
def test_bspline_classification_by_equality_equals_by_index(L, nel):
    """bsmpm.jl:63-76 classifies nodes by FP equality with xB; the device
    looks the type up by node index from a table built the same way.  For
    every configured mesh the two agree with the index pattern 1,2,3..3,4,1."""
    instr = ep.kwargser("instr", {}, dim=len(L))
    me = ep.setup_mesh(instr, ep.get_geom(nel, L, instr))
    for i, ax in enumerate(me.axes):
        t = _types_by_equality(ax, me.xB[i, 0], me.xB[i, 1], me.h[i])
        n = len(ax)
        want = [1, 2] + [3] * (n - 4) + [4, 1]
        assert t == want


def test_e2n_brute_force():
    ic, _ = slump(2)
    me = ic.mesh
    e2n = ep.mesh_e2n(me)
    nx, nz = int(me.nno[0]), int(me.nno[1])
    for el in range(0, int(me.nel[2]), 17):
        ex, ez = divmod(el, int(me.nel[1]))
        k = 0
        for i in range(-1, 3):
            for j in range(-1, 3):
                ix, iz = ex + i, ez + j
                want = 1 + iz + nz * ix if (0 <= ix < nx and 0 <= iz < nz) else 0
                assert e2n[k, el] == want
                k += 1


def test_kwargser_errors_match_reference():
    with pytest.raises(ValueError, match="not a supported shape function basis"):
        ep.kwargser("instr", {"basis": {"which": "foo", "how": None, "ghost": False}})
    with pytest.raises(ValueError, match="unsupported transfer scheme"):
        ep.kwargser("instr", {"fwrk": {"deform": "finite", "trsfr": "musl", "C_pf": 1.0, "musl": True,
                                       "locking": True, "damping": 0.1}})
    with pytest.raises(ValueError, match="InvalidReturnMapping"):
        ep.kwargser("instr", {"plast": {"status": True, "constitutive": "camC"}})
    with pytest.warns(UserWarning, match="miscellaneous kwargs"):
        ep.kwargser("instr", {"bogus": 1})
