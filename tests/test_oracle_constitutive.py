"""Pins of the oracle's return mappings (DP.jl:1-99, J2.jl:1-75) by closed
forms and yield-surface residuals."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

from elastoplasm.ic import get_elastic_stiffness
from oracle.oracle import load

_dp = C.POINTER(C.c_double)


def dp(tau, c0, cr, phi, Hp, K, G, e2):
    t = np.array(tau, dtype=np.float64)
    e1, dl = C.c_double(e2), C.c_double()
    load().or_dp_return(len(t), t.ctypes.data_as(_dp), c0, cr, phi, Hp, K, G, e2, C.byref(e1), C.byref(dl))
    return t, e1.value, dl.value


def dp_params(nstr, phi, c):
    if nstr == 3:
        eta = 3 * math.tan(phi) / math.sqrt(9 + 12 * math.tan(phi) ** 2)
        xi = 3 * c / math.sqrt(9 + 12 * math.tan(phi) ** 2)
    else:
        eta = 6 * math.sin(phi) / (math.sqrt(3) * (3 + math.sin(phi)))
        xi = 6 * c * math.cos(phi) / (math.sqrt(3) * (3 + math.sin(phi)))
    return eta, xi


def inv(t):
    t = np.asarray(t)
    if len(t) == 3:
        P = (t[0] + t[1]) / 2
        s = t - np.array([P, P, 0])
        return P, s, math.sqrt(0.5 * (s[0] ** 2 + s[1] ** 2) + s[2] ** 2)
    P = t[:3].mean()
    s = t - np.array([P, P, P, 0, 0, 0])
    return P, s, math.sqrt(0.5 * (s[0] ** 2 + s[1] ** 2 + s[2] ** 2) + s[3] ** 2 + s[4] ** 2 + s[5] ** 2)


@pytest.mark.parametrize("nstr", [3, 6])
def test_dp_shear_return_lands_on_cone(nstr):
    """Shear branch: τ_new = s·(ξ - ηP)/τII + P (ψ = 0 keeps P), Δλ =
    f_s(trial)/G; for P < σm this lands on the cone, f_s(τ_new) = 0.  For
    P >= σm with h > 0 the reference's signed τn = ξ - ηP is negative and
    mirrors the deviator (DP.jl:49-55); the closed form pins that too."""
    rng = np.random.default_rng(5)
    K, G = 1e6 / (3 * 0.4), 1e6 / 2.6
    phi, c0 = math.radians(30), 2e4
    eta, xi = dp_params(nstr, phi, c0)
    hits = 0
    for _ in range(500):
        t = rng.normal(size=nstr) * 6e4
        t[: (2 if nstr == 3 else 3)] -= 5e4
        P0, s0, tII0 = inv(t)
        fs0 = tII0 + eta * P0 - xi
        tn, e1, dl = dp(t, c0, c0, phi, 0.0, K, G, 0.0)
        if fs0 <= 0 and P0 < xi / eta:
            assert np.array_equal(tn, t) and dl == 0.0 and e1 == 0.0
            continue
        P1, s1, tII1 = inv(tn)
        h = tII0 - (xi - eta * xi / eta) - math.sqrt(1 + eta * eta) * (P0 - xi / eta)
        if (fs0 > 0 and P0 < xi / eta) or (h > 0 and P0 >= xi / eta):
            hits += 1
            want = s0 * ((xi - eta * P0) / tII0)
            want[: (2 if nstr == 3 else 3)] += P0
            assert np.allclose(tn, want, rtol=1e-12, atol=1e-9 * xi)
            if P0 < xi / eta:
                assert abs(tII1 + eta * P1 - xi) <= 1e-9 * abs(xi)
            assert dl == pytest.approx(fs0 / G, rel=1e-13)
            assert e1 == pytest.approx(dl * math.sqrt(1 / 3), rel=1e-13)
    assert hits > 50


@pytest.mark.parametrize("nstr", [3, 6])
def test_dp_apex_quirk(nstr):
    """Apex branch as the reference writes it: Pn = σm - P (DP.jl:58), the
    deviator zeroed, Δλ = (P - σm)/K."""
    K, G = 1e6, 5e5
    phi, c0 = math.radians(25), 1e4
    eta, xi = dp_params(nstr, phi, c0)
    sm = xi / eta
    nd = 2 if nstr == 3 else 3
    t = np.zeros(nstr)
    t[:nd] = 3 * sm
    t[nd - 1] += 1.0  # a small deviator
    P0, s0, tII0 = inv(t)
    tn, e1, dl = dp(t, c0, c0, phi, 0.0, K, G, 0.0)
    assert dl == pytest.approx((P0 - sm) / K, rel=1e-14)
    assert np.allclose(tn[:nd], sm - P0, rtol=1e-14)
    assert np.all(tn[nd:] == 0.0)
    assert e1 == pytest.approx(math.sqrt(2) * dl / 3, rel=1e-14)


def test_dp_softening_uses_epII2_and_residual():
    """c = max(c0 + Hp ϵpII[2], cr) (DP.jl:40-43)."""
    K, G = 1e6, 5e5
    phi = math.radians(20)
    t = np.array([-1e4, -3e4, 4e4])
    a = dp(t, 2e4, 4e3, phi, -3e4, K, G, 0.1)    # c = 2e4 - 3e3 = 1.7e4
    b = dp(t, 1.7e4, 4e3, phi, 0.0, K, G, 0.1)
    assert np.array_equal(a[0], b[0])
    c = dp(t, 2e4, 4e3, phi, -3e4, K, G, 10.0)   # clipped at cr
    d = dp(t, 4e3, 4e3, phi, 0.0, K, G, 10.0)
    assert np.array_equal(c[0], d[0])


def j2(tau, c0, cr, Hp, Del, e2):
    t = np.array(tau, dtype=np.float64)
    n = len(t)
    D = np.asfortranarray(Del)
    e1 = C.c_double(e2)
    act, it = C.c_int(), C.c_int()
    load().or_j2_return(n, t.ctypes.data_as(_dp), c0, cr, Hp, D.ctypes.data_as(_dp), e2, C.byref(e1),
                        C.byref(act), C.byref(it))
    return t, e1.value, act.value, it.value


def test_j2_radial_return_closed_form_3d():
    """Deviatoric, shear-free trial stress: Del n = 2G n, so the cutting
    plane is the textbook radial return: |s| = κ(γ), γ = (|s_tr| - κ0)/(2G + H)
    with f = √(2J₂) - κ, κ = c0 + H γ (restated intent of J2.jl:21-50)."""
    E, nu = 1e6, 0.3
    K, G, Del = get_elastic_stiffness(E, nu, 3)
    s = np.array([3e4, -1e4, -2e4, 0.0, 0.0, 0.0])
    P = -5e3
    t = s + np.array([P, P, P, 0, 0, 0])
    kappa0, H = 1e4, 1e4
    tn, e1, act, it = j2(t, kappa0, 1.0, H, Del, 0.0)
    rho_tr = math.sqrt(np.sum(s[:3] ** 2))          # √(2J₂) with zero shear
    gam = (rho_tr - kappa0) / (2 * G + H)
    assert act == 1
    assert e1 == pytest.approx(gam, rel=1e-9)
    want = P + s[:3] * (kappa0 + H * gam) / rho_tr
    assert np.allclose(tn[:3], want, rtol=1e-9)
    assert np.all(tn[3:] == 0.0)


@pytest.mark.parametrize("nstr", [3, 6])
def test_j2_converges_to_yield_surface(nstr):
    d = 2 if nstr == 3 else 3
    K, G, Del = get_elastic_stiffness(1e6, 0.3, d)
    rng = np.random.default_rng(7)
    capped = 0
    for _ in range(100):
        t = rng.normal(size=nstr) * 3e4
        tn, e1, act, it = j2(t, 1e4, 5e3, 1e4, Del, 0.0)
        P, s, _ = inv(tn)
        if nstr == 3:
            J2 = 0.5 * (s[0] ** 2 + s[1] ** 2 + 2 * s[2] ** 2)
        else:
            J2 = 0.5 * (np.sum(s[:3] ** 2) + 2 * np.sum(s[3:] ** 2))
        kappa = max(5e3, 1e4 + 1e4 * e1)
        if act:
            # the loop stops at |f| <= 1e-9 or after 19 corrections, the cap
            # the reference accepts silently (J2.jl:30)
            if it == 20:
                # linear convergence of the Voigt cutting plane: the residual
                # only has to have shrunk
                capped += 1
                P0, s0, _ = inv(t)
                J20 = 0.5 * (s0[0] ** 2 + s0[1] ** 2 + 2 * s0[2] ** 2) if nstr == 3 else \
                    0.5 * (np.sum(s0[:3] ** 2) + 2 * np.sum(s0[3:] ** 2))
                assert abs(math.sqrt(2 * J2) - kappa) < abs(math.sqrt(2 * J20) - 1e4)
            else:
                assert abs(math.sqrt(2 * J2) - kappa) <= 1e-9
        else:
            assert math.sqrt(2 * J2) <= kappa
    assert capped < 100
