"""Pins of the oracle's GIMP domain updates (domain.jl:1-52, chosen by
basis.how per framework in update.jl:7-29) by homogeneous deformations
whose stretches are known in closed form: a pure stretch, a rotation and a
simple shear, whose right stretch U = √(FᵀF) differs from the left stretch
√(FFᵀ)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from cases import slump
from oracle.oracle import Oracle, load

_dp = C.POINTER(C.c_double)
DT = 0.01


def _run(deform, how, dF, F0, steps=1):
    """Grid velocity v = A x with dt A = ΔF - I, so every particle gets ΔF
    exactly (linear completeness) and F = ΔF F0."""
    ic, cfg = slump(2, basis="gimpm", deform=deform, how=how, locking=False)
    o = Oracle(ic.mesh, ic.mpts, ic.cmpr, cfg.instr)
    mp, me = ic.mpts, ic.mesh
    L = load()
    mp.s.F[:] = F0[:, :, None]
    l0 = mp.l0.copy()
    for _ in range(steps):
        assert L.or_shpfun(C.byref(o.pt), C.byref(o.m), C.byref(o.ins)) == 0
        me.v[:] = ((dF - np.eye(2)) / DT) @ me.x
        L.or_update(C.byref(o.pt), C.byref(o.m), DT, C.byref(o.ins))
    assert np.abs(mp.s.dF - dF[:, :, None]).max() < 1e-12
    return l0, mp.l, mp.s.F[:, :, 0]


def _right_stretch(F):
    C_ = F.T @ F
    w, V = np.linalg.eigh(C_)
    return V @ np.diag(np.sqrt(w)) @ V.T


STRETCH = np.diag([1.02, 0.97])
ROT = (lambda t: np.array([[math.cos(t), -math.sin(t)], [math.sin(t), math.cos(t)]]))(0.03)
SHEAR = np.array([[1.0, 0.04], [0.0, 1.0]])


@pytest.mark.parametrize("name,dF", [("stretch", STRETCH), ("rotation", ROT), ("shear", SHEAR)])
def test_finite_domains(name, dF):
    F0 = np.eye(2)
    for how in ("undeformed", "detFij", "Fii", "Uii"):
        l0, l, F = _run("finite", how, dF, F0)
        want = {"undeformed": l0, "detFij": l0 * np.linalg.det(F),
                "Fii": l0 * np.diag(F)[:, None], "Uii": l0 * np.diag(_right_stretch(F))[:, None]}[how]
        assert np.abs(l - want).max() <= 1e-12 * np.abs(want).max(), (name, how)
    # closed forms: U of a rotation is I; U of a simple shear γ has
    # U_11 = 2/√(4+γ²), U_22 = (2+γ²)/√(4+γ²), unlike V = √(FFᵀ)
    l0, l, F = _run("finite", "Uii", dF, F0)
    if name == "rotation":
        assert np.abs(l - l0).max() <= 1e-12 * l0.max()
    if name == "shear":
        g = 0.04
        u = np.array([2.0, 2.0 + g * g]) / math.sqrt(4.0 + g * g)
        assert np.abs(l - l0 * u[:, None]).max() <= 1e-12 * l0.max()
        w, V = np.linalg.eigh(F @ F.T)
        left = np.diag(V @ np.diag(np.sqrt(w)) @ V.T)
        assert np.abs(left - u).max() > 1e-4


@pytest.mark.parametrize("name,dF", [("stretch", STRETCH), ("shear", SHEAR)])
def test_infinitesimal_domains_accumulate(name, dF):
    """detΔFij, ΔFii, ΔUii scale the current ℓ each step: two steps give the
    square of one step's factor."""
    F0 = np.eye(2)
    for how in ("undeformed", "detΔFij", "ΔFii", "ΔUii"):
        l0, l, _ = _run("infinitesimal", how, dF, F0, steps=2)
        f = {"undeformed": np.ones(2), "detΔFij": np.full(2, np.linalg.det(dF)), "ΔFii": np.diag(dF),
             "ΔUii": np.diag(_right_stretch(dF))}[how]
        want = l0 * (f ** 2)[:, None]
        assert np.abs(l - want).max() <= 1e-12 * np.abs(want).max(), (name, how)


def test_framework_checks_how():
    with pytest.raises(ValueError, match="domain update"):
        slump(2, basis="gimpm", deform="finite", how="ΔFii")
    with pytest.raises(ValueError, match="domain update"):
        slump(2, basis="gimpm", deform="infinitesimal", how="Uii")
    with pytest.raises(ValueError, match="domain update"):
        slump(2, basis="gimpm", how=None)
