"""Pins of the oracle's non-local regularisation (nonlocal.jl:18-47,
plast.jl:49-60, e2e.jl:20-61).

The reference's kernel runs its three passes in particle order.  Written out,
for a plastic particle p (Δλ of the last return != 0) with neighbourhood N(p)
(particles whose element lies within ceil(ls/h) elements of p's on every
axis) and ω(d) = (d/ls) exp(-(d/ls)²), d from the first two coordinates:
  W_p  = Σ_{q ∈ N(p)} ω(d_pq)                       (each pair counted once,
                                                      by whichever end comes first)
  ϵ2_p = Σ_{q ∈ N(p), not (q plastic and q < p)} ω(d_pq) ϵ1_q / W_p
and ϵ2 = 0 for every other particle.  The tests compare the oracle's loop
with this closed form, computed here from a dense distance matrix."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from cases import slump
from oracle.oracle import Oracle, load


def _setup(dim, ls, frac, seed):
    nel = [40, 10] if dim == 2 else [16, 6, 6]
    ic, cfg = slump(dim, nel=nel)
    cfg.instr["nonloc"] = {"status": True, "ls": ls}
    o = Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)
    L = load()
    assert L.or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    rng = np.random.default_rng(seed)
    mp = ic.mpts
    mp.x[:] += rng.uniform(-0.2, 0.2, mp.x.shape) * ic.mesh.h[:, None]   # positions after the step move
    mp.s.dlam[:] = np.where(rng.random(mp.nmp) < frac, rng.uniform(1e-4, 1e-3, mp.nmp), 0.0)
    mp.s.epII[0] = rng.uniform(0.0, 0.05, mp.nmp)
    mp.s.epII[1] = 123.0          # overwritten
    return ic, o, L


def _closed_form(ic, ls):
    mp, me = ic.mpts, ic.mesh
    d = me.dim
    nel = np.asarray(me.nel[:d], dtype=np.int64)
    e = mp.p2e - 1
    ec = np.zeros((d, mp.nmp), dtype=np.int64)
    ec[d - 1] = e % nel[d - 1]
    r = e // nel[d - 1]
    ec[0] = r % nel[0]
    if d == 3:
        ec[1] = r // nel[0]
    nnel = np.ceil(ls / me.h[:d]).astype(np.int64)
    near = np.ones((mp.nmp, mp.nmp), bool)
    for i in range(d):
        near &= np.abs(ec[i][:, None] - ec[i][None, :]) <= nnel[i]
    dist = np.hypot(mp.x[0][:, None] - mp.x[0][None, :], mp.x[1][:, None] - mp.x[1][None, :])
    om = np.where(near, dist / ls * np.exp(-(dist / ls) ** 2), 0.0)
    plastic = mp.s.dlam != 0.0
    idx = np.arange(mp.nmp)
    excl = plastic[None, :] & (idx[None, :] < idx[:, None])
    W = om.sum(axis=1)
    S = np.where(excl, 0.0, om) @ mp.s.epII[0]
    out = np.where(W > 1e-16, S / np.where(W > 0, W, 1.0), mp.s.epII[0])
    return np.where(plastic, out, 0.0), W


@pytest.mark.parametrize("dim,ls", [(2, 0.5), (2, 3.0), (3, 2.0)])
def test_nonlocal_matches_closed_form(dim, ls):
    ic, o, L = _setup(dim, ls, 0.4, 7)
    want, W = _closed_form(ic, ls)
    assert L.or_nonlocal(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    got = ic.mpts.s.epII[1]
    plastic = ic.mpts.s.dlam != 0
    assert plastic.sum() > 20 and (W[plastic] > 1e-16).any()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)
    assert (got[~plastic] == 0).all()


def test_constant_field_first_particle_full_average():
    """With ϵ1 = c everywhere and every particle plastic, the first particle
    (no earlier plastic neighbour) averages over its whole neighbourhood:
    ϵ2 = c; the last one keeps only itself, whose weight is ω(0) = 0."""
    ic, o, L = _setup(2, 3.0, 1.0, 8)
    mp = ic.mpts
    mp.s.epII[0] = 0.02
    assert L.or_nonlocal(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    assert mp.s.epII[1][0] == pytest.approx(0.02, rel=1e-14)
    assert mp.s.epII[1][-1] == 0.0
    assert (mp.s.epII[1] <= 0.02 * (1 + 1e-14)).all()


def test_isolated_plastic_particle_keeps_its_value():
    """W <= 1e-16 (no neighbour at a distance): ϵ2 = ϵ1 (nonlocal.jl:38-40)."""
    ic, o, L = _setup(2, 0.5, 0.0, 9)
    mp = ic.mpts
    # one plastic particle far from every other: move it to an empty element
    mp.s.dlam[5] = 1e-3
    mp.x[:, 5] = ic.mesh.xB[:, 1] - 0.5 * ic.mesh.h
    assert L.or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    want, W = _closed_form(ic, 0.5)
    assert L.or_nonlocal(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
    if W[5] <= 1e-16:
        assert mp.s.epII[1][5] == mp.s.epII[0][5]
    np.testing.assert_allclose(mp.s.epII[1], want, rtol=1e-12, atol=1e-15)
