"""Pins of the oracle's shape functions (bsmpm.jl:1-135, gimpm.jl:1-70,
smpm.jl:1-66) by what the mathematics fixes: partition of unity, linear
completeness, derivative = finite difference, and the reference's own
invariants (docs/src/misc/test/check_bsmpm.png: PoU 1 and linear
completeness error ~1e-14; shpTest.jl:107 PoU bounds)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from cases import slump
from oracle.oracle import Oracle, load

_dp = C.POINTER(C.c_double)


def bs(xi, t):
    f, d = C.c_double(), C.c_double()
    load().or_bsmpm_1d(xi, t, C.byref(f), C.byref(d))
    return f.value, d.value


def gimp(dx, h, lp):
    f, d = C.c_double(), C.c_double()
    load().or_gimp_1d(dx, h, lp, C.byref(f), C.byref(d))
    return f.value, d.value


def lin(dx, h):
    f, d = C.c_double(), C.c_double()
    load().or_smpm_1d(dx, h, C.byref(f), C.byref(d))
    return f.value, d.value


def test_bspline_interior_is_textbook_cubic():
    # type 3 = uniform cubic B-spline: N(0)=2/3, N(±1)=1/6, N(±2)=0
    assert bs(0.0, 3)[0] == pytest.approx(2 / 3, abs=1e-15)
    assert bs(1.0, 3)[0] == pytest.approx(1 / 6, abs=1e-15)
    assert bs(-1.0, 3)[0] == pytest.approx(1 / 6, abs=1e-15)
    assert bs(2.0, 3)[0] == pytest.approx(0.0, abs=1e-15)
    # closed form (2-|x|)^3/6 on 1<=|x|<=2, 2/3 - x^2 + |x|^3/2 on |x|<=1
    for x in np.linspace(-2, 2, 41):
        a = abs(x)
        want = (2 - a) ** 3 / 6 if a >= 1 else 2 / 3 - a * a + a ** 3 / 2
        assert bs(x, 3)[0] == pytest.approx(want, abs=1e-14)


@pytest.mark.parametrize("t", [1, 2, 3, 4])
def test_bspline_derivative_matches_fd(t):
    # boundary-modified types have kinks at integer offsets (the wall node);
    # the derivative is checked away from them
    for x in np.linspace(-1.97, 1.97, 37):
        if abs(x - round(x)) < 1e-3:
            continue
        f1, d = bs(x, t)
        e = 1e-6
        fd = (bs(x + e, t)[0] - bs(x - e, t)[0]) / (2 * e)
        assert d == pytest.approx(fd, abs=1e-7)


def test_gimp_and_linear_derivatives_match_fd():
    h, lp = 1.3, 0.3
    for x in np.linspace(-1.55, 1.55, 41):
        if min(abs(abs(x) - lp), abs(abs(x) - (h - lp)), abs(abs(x) - (h + lp))) < 1e-3:
            continue
        e = 1e-7
        assert gimp(x, h, lp)[1] == pytest.approx((gimp(x + e, h, lp)[0] - gimp(x - e, h, lp)[0]) / (2 * e), abs=1e-6)
        if abs(x) > 1e-3 and abs(abs(x) - h) > 1e-3:
            assert lin(x, h)[1] == pytest.approx((lin(x + e, h)[0] - lin(x - e, h)[0]) / (2 * e), abs=1e-6)


def test_gimp_pou_1d():
    h, lp = 1.0, 0.25
    for xp in np.linspace(2.3, 3.7, 15):
        nodes = np.arange(0, 7) * h
        s = sum(gimp(xp - xn, h, lp)[0] for xn in nodes)
        ds = sum(gimp(xp - xn, h, lp)[1] for xn in nodes)
        assert s == pytest.approx(1.0, abs=1e-14)
        assert ds == pytest.approx(0.0, abs=1e-13)


def _stencil_sums(dim, basis):
    ic, cfg = slump(dim, basis=basis)
    o = Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)
    assert load().or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    mp, me = ic.mpts, ic.mesh
    nn = me.nn
    ids = mp.p2n
    N = mp.phi[:, :, 0]
    pou = N.sum(axis=0)
    xn = np.zeros((nn, mp.nmp, dim))
    for i in range(dim):
        xn[:, :, i] = np.where(ids > 0, me.x[i, np.maximum(ids - 1, 0)], 0.0)
    lin = np.einsum("kp,kpi->ip", N, xn)
    grads = mp.phi[:, :, 1:].sum(axis=0)                       # Σ ∂N = 0
    dlin = np.einsum("kpj,kpi->pij", mp.phi[:, :, 1:], xn)      # Σ ∂_j N x_i = δ_ij
    return pou, lin, mp.x, grads, dlin


@pytest.mark.parametrize("dim", [2, 3])
def test_bsmpm_pou_and_linear_completeness_on_slump_mesh(dim):
    pou, lin, x, grads, dlin = _stencil_sums(dim, "bsmpm")
    assert np.abs(pou - 1).max() < 1e-13
    assert np.abs(lin - x).max() < 1e-12
    assert np.abs(grads).max() < 1e-12
    assert np.abs(dlin - np.eye(dim)[None]).max() < 1e-12


@pytest.mark.parametrize("basis", ["gimpm", "smpm"])
def test_gimp_smpm_pou_on_slump_mesh(basis):
    pou, lin, x, grads, dlin = _stencil_sums(2, basis)
    assert np.abs(pou - 1).max() < 1e-13
    assert np.abs(lin - x).max() < 1e-12
    assert np.abs(grads).max() < 1e-12


def test_p2e2n_element_ids_brute_force():
    ic, cfg = slump(2)
    o = Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)
    load().or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins))
    me, mp = ic.mesh, ic.mpts
    xs, zs = me.axes
    for p in range(0, mp.nmp, 7):
        ex = np.searchsorted(xs, mp.x[0, p], side="right") - 1
        ez = np.searchsorted(zs, mp.x[1, p], side="right") - 1
        ex, ez = min(max(ex, 0), me.nel[0] - 1), min(max(ez, 0), me.nel[1] - 1)
        assert mp.p2e[p] == 1 + ez + me.nel[1] * ex
