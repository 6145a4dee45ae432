"""Pins of the oracle's transfers and deformation update by conservation laws
and exact reproduction of simple fields (p2n.jl, solve.jl, n2p.jl, augm.jl,
deform.jl, volumetric.jl, elast.jl)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sl

from cases import slump
from oracle.oracle import Oracle, load

_dp = C.POINTER(C.c_double)


def _oracle(dim, **kw):
    ic, cfg = slump(dim, **kw)
    o = Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)
    return ic, cfg, o


def _g(dim, G=9.81):
    g = np.zeros(3)
    g[dim - 1] = -G
    return g


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("trsfr", ["std", "tpic"])
def test_p2g_conservation(dim, trsfr):
    """Σ m_n = Σ ρΩ, Σ p_n = Σ m_p v_p (+ Σ_n N m ∇v Δnp = 0 for TPIC by
    linear completeness), Σ oobf = Σ m_p g since Σ_n ∂N = 0."""
    ic, cfg, o = _oracle(dim, trsfr=trsfr, musl=trsfr == "std")
    mp, me = ic.mpts, ic.mesh
    rng = np.random.default_rng(11)
    mp.s.v[:] = rng.normal(size=mp.s.v.shape)
    mp.s.tau[:] = rng.normal(size=mp.s.tau.shape) * 1e4
    mp.s.gradv[:] = rng.normal(size=mp.s.gradv.shape)
    mp.J[:] = 1 + 0.01 * rng.random(mp.nmp)
    mp.Om[:] = mp.J * mp.Om0
    mp.s.rho[:] = mp.s.rho0 / mp.J
    L = load()
    assert L.or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    g = _g(dim)
    L.or_p2n(C.byref(o.pt), C.byref(o.m), g.ctypes.data_as(_dp), C.byref(o.ins))
    ms = mp.s.rho * mp.Om
    assert me.m.sum() == pytest.approx(ms.sum(), rel=1e-13)
    assert np.allclose(me.p.sum(axis=1), (ms * mp.s.v).sum(axis=1), rtol=1e-11, atol=1e-9 * ms.sum())
    want_f = np.zeros(dim)
    want_f[-1] = -9.81 * ms.sum()
    scale = np.abs(me.oobf).sum()
    assert np.allclose(me.oobf.sum(axis=1), want_f, rtol=1e-11, atol=1e-11 * scale)
    # σ = τ/J (transform, p2n.jl:12-21)
    assert np.allclose(mp.s.sig, mp.s.tau / mp.J, rtol=1e-15)


def test_rigid_translation_through_step():
    """Uniform velocity, zero stress, no gravity, no walls: the grid velocity
    is uniform (PoU), particles move by dt v and keep v; mUSL reproduces v;
    ∇v = 0 so F stays I."""
    ic, cfg, o = _oracle(2, C_pf=0.9)
    mp, me = ic.mpts, ic.mesh
    me.bcs.status[:] = False
    o._bcs[:] = 0
    v = np.array([0.3, -0.2])
    mp.s.v[:] = v[:, None]
    x0 = mp.x.copy()
    o.step(np.zeros(2), 0.01, False)
    assert np.allclose(mp.s.v, v[:, None], atol=1e-14)
    assert np.allclose(mp.x, x0 + 0.01 * v[:, None], atol=1e-13)
    assert np.allclose(mp.s.F, np.eye(2)[:, :, None], atol=1e-13)
    mask = me.m > 0
    assert np.allclose(me.v[:, mask], v[:, None], atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_affine_velocity_gives_exact_gradient_and_fbar_identity(dim):
    """Grid velocity v_n = A x_n + c: ∇v_p = A exactly (linear completeness),
    ΔF = I + dt A, F = ΔF F, J = det F; a uniform ΔJ leaves F-bar's factor at
    1 (volumetric.jl:23-35)."""
    ic, cfg, o = _oracle(dim, locking=True)
    mp, me = ic.mpts, ic.mesh
    L = load()
    assert L.or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    L.or_p2n(C.byref(o.pt), C.byref(o.m), np.zeros(3).ctypes.data_as(_dp), C.byref(o.ins))
    rng = np.random.default_rng(12)
    A = 0.05 * rng.normal(size=(dim, dim))
    c = rng.normal(size=dim)
    me.v[:] = A @ me.x + c[:, None]
    F0 = np.eye(dim) + 0.01 * rng.normal(size=(dim, dim))
    mp.s.F[:] = F0[:, :, None]
    dt = 0.02
    L.or_update(C.byref(o.pt), C.byref(o.m), dt, C.byref(o.ins))
    assert np.abs(mp.s.gradv - A[:, :, None]).max() < 1e-12
    dF = np.eye(dim) + dt * A
    Fn = dF @ F0
    assert np.abs(mp.s.F - Fn[:, :, None]).max() < 1e-13
    assert np.allclose(mp.J, np.linalg.det(Fn), rtol=1e-13)
    assert np.allclose(mp.dJ, np.linalg.det(dF), rtol=1e-13)
    assert np.abs(mp.s.dF - dF[:, :, None]).max() < 1e-12      # F-bar factor 1
    assert np.allclose(mp.Om, mp.J * mp.Om0, rtol=1e-15)
    assert np.allclose(mp.s.rho, mp.s.rho0 / mp.J, rtol=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
def test_finite_elastic_predictor_against_logm(dim):
    """b ← ΔF b ΔFᵀ, τ = Del · voigt(½ log b, 2) (elast.jl:45-56)."""
    ic, cfg, o = _oracle(dim)
    mp = ic.mpts
    rng = np.random.default_rng(13)
    n = min(mp.nmp, 50)
    for p in range(n):
        M = 0.05 * rng.normal(size=(dim, dim))
        mp.s.dF[:, :, p] = np.eye(dim) + M
        B = np.eye(dim) + 0.05 * (lambda S: S @ S.T)(rng.normal(size=(dim, dim)))
        mp.s.b[:, :, p] = B
    b0 = mp.s.b.copy()
    load().or_elast(C.byref(o.pt), C.byref(o.ins))
    Del = ic.cmpr.Del
    for p in range(n):
        dF = mp.s.dF[:, :, p]
        bn = dF @ b0[:, :, p] @ dF.T
        assert np.allclose(mp.s.b[:, :, p], bn, rtol=1e-14, atol=1e-15)
        e = 0.5 * sl.logm(bn).real
        ev = [e[0, 0], e[1, 1], 2 * e[0, 1]] if dim == 2 else \
            [e[0, 0], e[1, 1], e[2, 2], 2 * e[1, 2], 2 * e[0, 2], 2 * e[0, 1]]
        assert np.allclose(mp.s.tau[:, p], Del @ np.array(ev), rtol=1e-10, atol=1e-9 * np.abs(Del).max() * 1e-4)


def test_infinitesimal_elastic_increment():
    """σ += Del · voigt(sym ΔF - I, 2); the Jaumann term vanishes
    (elast.jl:69-80)."""
    ic, cfg, o = _oracle(2, deform="infinitesimal")
    mp = ic.mpts
    rng = np.random.default_rng(14)
    mp.s.dF[:] = np.eye(2)[:, :, None] + 1e-3 * rng.normal(size=mp.s.dF.shape)
    mp.s.sig[:] = rng.normal(size=mp.s.sig.shape) * 1e3
    s0 = mp.s.sig.copy()
    load().or_elast(C.byref(o.pt), C.byref(o.ins))
    for p in range(0, mp.nmp, 13):
        dF = mp.s.dF[:, :, p]
        e = 0.5 * (dF + dF.T) - np.eye(2)
        want = s0[:, p] + ic.cmpr.Del @ np.array([e[0, 0], e[1, 1], 2 * e[0, 1]])
        assert np.allclose(mp.s.sig[:, p], want, rtol=1e-13, atol=1e-9)


def test_get_dt_rule():
    """dt = min(0.5 max_d h_d/(vmax_d + c), T - t); vmax reset (physics.jl:21-26)."""
    ic, cfg, o = _oracle(2)
    o.pt.vmax[0], o.pt.vmax[1] = 3.0, 1.0
    h, c = ic.mesh.h, ic.cmpr.c
    dt = o.get_dt(100.0)
    assert dt == 0.5 * max(h[0] / (3.0 + c), h[1] / (1.0 + c))
    assert o.pt.vmax[0] == 0.0 and o.pt.vmax[1] == 0.0
    o.t = 99.99
    assert o.get_dt(100.0) == pytest.approx(100.0 - 99.99)


def _solve_nodes(dim, eta, nodes):
    """Run or_solve on the slump mesh with every node massless except
    `nodes`: {node: (m, p[dim], f[dim], bc[dim])}; returns (a, v, oobf)."""
    ic, cfg, o = _oracle(dim, damping=eta)
    me = ic.mesh
    me.m[:] = 0.0
    me.p[:] = 7.0       # junk on massless nodes must not leak into a, v
    me.oobf[:] = -5.0
    o._bcs[:] = 0
    for no, (m, p, f, bc) in nodes.items():
        me.m[no] = m
        me.p[:, no] = p
        me.oobf[:, no] = f
        o._bcs[:, no] = bc
    load().or_solve(C.byref(o.m), 0.01, C.byref(o.ins))
    return me.a.copy(), me.v.copy(), me.oobf.copy()


@pytest.mark.parametrize("dim", [2, 3])
def test_grid_solve_undamped_closed_form(dim):
    """η = 0 (solve.jl:14-40): a = f/m and v = (p + dt f)/m at every free
    component; Dirichlet components and massless nodes give a = v = 0."""
    rng = np.random.default_rng(21)
    nodes = {}
    for no in (3, 40, 111):
        nodes[no] = (rng.uniform(0.5, 2.0), rng.normal(size=dim), rng.normal(size=dim) * 10, np.zeros(dim, np.uint8))
    nodes[77] = (1.5, np.full(dim, 0.4), np.full(dim, 3.0), np.eye(dim, dtype=np.uint8)[0])   # Dirichlet in x
    a, v, f = _solve_nodes(dim, 0.0, nodes)
    for no, (m, p, ff, bc) in nodes.items():
        free = bc == 0
        np.testing.assert_allclose(a[free, no], ff[free] / m, rtol=1e-15)
        np.testing.assert_allclose(v[free, no], (p[free] + 0.01 * ff[free]) / m, rtol=1e-14)
        assert (a[~free, no] == 0).all() and (v[~free, no] == 0).all()
        np.testing.assert_array_equal(f[:, no], ff)       # undamped: oobf untouched
    massless = np.setdiff1d(np.arange(a.shape[1]), list(nodes))
    assert (a[:, massless] == 0).all() and (v[:, massless] == 0).all()


def test_grid_solve_gate_and_sequential_norm():
    """Damping D = η‖oobf‖ sign(p/m) acts only where |p/m| >= 1e-3, opposes the
    nodal velocity, and axis 2 sees the norm after axis 1 was damped
    (solve.jl:20-31).  Hand values for η = 0.1, m = 2, f = (3, 4):
      p = (2, 2):   f1 = 3 - 0.1*5 = 2.5; f2 = 4 - 0.1*sqrt(2.5^2 + 4^2) = 3.528301...
      p = (-2, 2):  f1 = 3 + 0.5 = 3.5;   f2 = 4 - 0.1*sqrt(3.5^2 + 4^2)
      p = (1e-3, 2): axis 1 below the gate (|p/m| = 5e-4), untouched;
                     f2 = 4 - 0.1*5 = 3.5
    A norm taken once before damping would give f2 = 3.5 in the first case."""
    z = np.zeros(2, np.uint8)
    f0 = np.array([3.0, 4.0])
    nodes = {5: (2.0, np.array([2.0, 2.0]), f0, z), 9: (2.0, np.array([-2.0, 2.0]), f0, z),
             13: (2.0, np.array([1e-3, 2.0]), f0, z)}
    a, v, f = _solve_nodes(2, 0.1, nodes)
    assert f[0, 5] == pytest.approx(2.5, rel=1e-15)
    assert f[1, 5] == pytest.approx(4.0 - 0.1 * 4.716990566028302, rel=1e-15)
    assert abs(f[1, 5] - 3.5) > 1e-3                       # not the norm-before-damping form
    assert f[0, 9] == pytest.approx(3.5, rel=1e-15)         # damping opposes v_x < 0
    assert f[1, 9] == pytest.approx(4.0 - 0.1 * 5.315072906367325, rel=1e-15)
    assert f[0, 13] == 3.0                                  # gate: untouched
    assert f[1, 13] == pytest.approx(3.5, rel=1e-15)
    for no in nodes:
        np.testing.assert_allclose(a[:, no], f[:, no] / 2.0, rtol=1e-15)
        np.testing.assert_allclose(v[:, no], (nodes[no][1] + 0.01 * f[:, no]) / 2.0, rtol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("C_pf", [0.0, 0.37, 1.0])
def test_picflip_blend_on_affine_grid_fields(dim, C_pf):
    """Grid fields v_n = A x_n + c and a_n = B x_n + d (n2p.jl:14-40).  Linear
    completeness of the basis gives Σ N v_n = A x_p + c, so C_pf = 0 (PIC)
    sets v_p = A x_p + c, C_pf = 1 (FLIP) v_p + dt (B x_p + d), and the blend
    is their convex combination; x_p += dt (A x_p + c) for every C_pf; vmax
    is max |v_p| per axis."""
    ic, cfg, o = _oracle(dim, C_pf=C_pf, musl=False)
    mp, me = ic.mpts, ic.mesh
    L = load()
    assert L.or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    rng = np.random.default_rng(22)
    A, B = rng.normal(size=(dim, dim)), rng.normal(size=(dim, dim))
    c, d = rng.normal(size=dim), rng.normal(size=dim)
    me.v[:] = A @ me.x + c[:, None]
    me.a[:] = B @ me.x + d[:, None]
    mp.s.v[:] = rng.normal(size=mp.s.v.shape)
    v0, x0 = mp.s.v.copy(), mp.x.copy()
    for i in range(dim):
        o.pt.vmax[i] = 0.0
    dt = 0.01
    L.or_n2p(C.byref(o.pt), C.byref(o.m), dt, C.byref(o.ins))
    pic = A @ x0 + c[:, None]
    flip = v0 + dt * (B @ x0 + d[:, None])
    scale = np.abs(pic).max() + np.abs(flip).max()
    np.testing.assert_allclose(mp.s.v, C_pf * flip + (1 - C_pf) * pic, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(mp.x, x0 + dt * pic, rtol=0, atol=1e-12 * np.abs(x0).max())
    for i in range(dim):
        assert o.pt.vmax[i] == np.abs(mp.s.v[i]).max()
