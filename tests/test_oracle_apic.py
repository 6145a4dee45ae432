"""Pins of the oracle's APIC transfer (Dij_nd misc.jl:15-27, apic_{2,3}d_p2n
p2n.jl:231-280, Bij Bij.jl:14-26, n2p.jl:59-74) by what the cubic B-spline
fixes: zero first moment and second moment h²/3 per axis at interior
particles, so a linear grid velocity v = L x + c gathers to B = L·diag(h²/3),
and an APIC scatter returns B through the same moment."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from cases import clone, relerr, slump
from oracle.oracle import Oracle, load

_dp = C.POINTER(C.c_double)


def _oracle(dim, **kw):
    ic, cfg = slump(dim, trsfr="apic", **kw)
    return ic, cfg, Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)


def _interior(mp, me):
    """Particles whose 4 stencil nodes per axis are all interior B-spline
    nodes (none at xB1, xB1+h, xB2-h, xB2)."""
    x, h = mp.x, me.h[:, None]
    lo, hi = me.xB[:, 0][:, None], me.xB[:, 1][:, None]
    return ((x - lo >= 3.0 * h) & (hi - x >= 4.0 * h)).all(axis=0)


@pytest.mark.parametrize("dim", [2, 3])
def test_bij_of_linear_velocity(dim):
    ic, cfg, o = _oracle(dim)
    mp, me = ic.mpts, ic.mesh
    L = load()
    assert L.or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    L.or_p2n(C.byref(o.pt), C.byref(o.m), np.zeros(3).ctypes.data_as(_dp), C.byref(o.ins))
    rng = np.random.default_rng(21)
    A = rng.normal(size=(dim, dim))
    c = rng.normal(size=dim)
    me.v[:] = A @ me.x + c[:, None]
    me.a[:] = 0.0
    L.or_n2p(C.byref(o.pt), C.byref(o.m), 0.01, C.byref(o.ins))
    inner = _interior(mp, me)
    assert inner.sum() >= 20
    want = A @ np.diag(me.h[:dim] ** 2 / 3.0)
    err = np.abs(mp.Bij[:, :, inner] - want[:, :, None]).max() / np.abs(want).max()
    assert err < 1e-12
    # a transposed gather would give Aᵀ diag(h²/3), which differs
    assert np.abs(A - A.T).max() > 0.1


@pytest.mark.parametrize("dim", [2, 3])
def test_apic_scatter_returns_b(dim):
    """One interior particle with velocity 0 and affine B: Σ_n p_n = 0 and
    Σ_n p_n δ_nᵀ = m B diag(h²/3) (D⁻¹ = I, since D is rank one)."""
    ic, cfg, o = _oracle(dim)
    mp, me = ic.mpts, ic.mesh
    inner = np.flatnonzero(_interior(mp, me))
    k = inner[len(inner) // 2]
    rng = np.random.default_rng(22)
    Bm = rng.normal(size=(dim, dim))
    mp.s.v[:] = 0.0
    mp.Bij[:] = 0.0
    mp.Bij[:, :, k] = Bm
    mp.s.tau[:] = 0.0
    L = load()
    assert L.or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    L.or_p2n(C.byref(o.pt), C.byref(o.m), np.zeros(3).ctypes.data_as(_dp), C.byref(o.ins))
    m = mp.s.rho[k] * mp.Om[k]
    d = me.x - mp.x[:, k][:, None]
    assert np.abs(me.p.sum(axis=1)).max() < 1e-12 * m * np.abs(Bm).max()
    mom = me.p @ d.T
    want = m * Bm @ np.diag(me.h[:dim] ** 2 / 3.0)
    assert np.abs(mom - want).max() < 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("dim", [2, 3])
def test_dij_is_last_existing_node(dim):
    """Dᵢⱼ is N δδᵀ of the last existing stencil node (assignment, not a sum),
    so it is rank one and the identity stands in for D⁻¹."""
    ic, cfg, o = _oracle(dim)
    mp, me = ic.mpts, ic.mesh
    L = load()
    assert L.or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    for p in range(0, mp.nmp, max(1, mp.nmp // 200)):
        ks = np.flatnonzero(mp.p2n[:, p])
        k = ks[-1]
        no = mp.p2n[k, p] - 1
        dl = me.x[:, no] - mp.x[:, p]
        want = mp.phi[k, p, 0] * np.outer(dl, dl)
        assert np.array_equal(mp.Dij[:, :, p], want)
        assert abs(L.or_det(dim, mp.Dij[:, :, p].ravel(order="F").ctypes.data_as(_dp))) <= 1e-12


def test_det_and_inv_against_numpy():
    L = load()
    rng = np.random.default_rng(23)
    for d in (2, 3):
        for _ in range(50):
            A = np.asfortranarray(rng.normal(size=(d, d)))
            Ai = np.zeros((d, d), order="F")
            det = L.or_det(d, A.ctypes.data_as(_dp))
            L.or_inv(d, A.ctypes.data_as(_dp), Ai.ctypes.data_as(_dp))
            assert det == pytest.approx(np.linalg.det(A), rel=1e-12, abs=1e-14)
            assert np.allclose(Ai, np.linalg.inv(A), rtol=1e-10, atol=1e-12)
        T = np.asfortranarray(np.triu(rng.normal(size=(d, d))))
        assert L.or_det(d, T.ctypes.data_as(_dp)) == np.prod(np.diag(T))


@pytest.mark.parametrize("dim", [2, 3])
def test_first_step_matches_standard_transfer(dim):
    """B starts at zero, so the first APIC step is the standard step up to
    the association of N·m·v; the second is not (B ≠ 0 by then)."""
    ic, cfg = slump(dim, trsfr="apic")
    ic_s, cfg_s = slump(dim, trsfr="std")
    a, s = clone(ic), clone(ic_s)
    oa = Oracle(a.mesh, a.mpts, a.cmpr, cfg.instr)
    os_ = Oracle(s.mesh, s.mpts, s.cmpr, cfg_s.instr)
    oa.run(1, 100.0)
    os_.run(1, 100.0)
    assert relerr(a.mpts.s.v, s.mpts.s.v) < 1e-13
    assert relerr(a.mpts.s.tau, s.mpts.s.tau) < 1e-12
    assert np.abs(a.mpts.Bij).max() > 0
    oa.run(1, 100.0)
    os_.run(1, 100.0)
    assert relerr(a.mpts.s.v, s.mpts.s.v) > 1e-8
