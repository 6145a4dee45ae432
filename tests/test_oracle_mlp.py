"""Pins of the surrogate's CPU oracle (oracle/mlp.py): the network's forward
and backward against torch autograd, bf16 rounding against torch.bfloat16,
the draw finalizer against splitmix64's published output, the principal
Drucker-Prager return against the C oracle's Voigt return (DP.jl) and, for
ψ > 0, against the return's closed-form invariants."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest
import torch

from oracle import mlp
from oracle.oracle import _dp, load


def test_splitmix_finalizer_published_value():
    # SplitMix64 seeded with 0: the first output is mix(0x9E3779B97F4A7C15)
    assert int(mlp.splitmix_mix(np.uint64(0x9E3779B97F4A7C15))) == 0xE220A8397B1DCDAF
    u = mlp.u01(44, np.arange(200000, dtype=np.uint64), 3)
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 3e-3 and abs(u.var() - 1 / 12) < 2e-3


def test_bf16_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-30, 30, 10000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -1e-40]])
    ref = torch.from_numpy(x.astype(np.float32)).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mlp.bf16(x), ref)


class _Net(torch.nn.Module):
    def __init__(self, w):
        super().__init__()
        p = mlp.unpack(w)
        self.p = torch.nn.ParameterDict({k: torch.nn.Parameter(torch.tensor(v, dtype=torch.float64))
                                         for k, v in p.items()})

    def forward(self, x):
        p = self.p
        h = torch.nn.functional.silu(x @ p["W1"].T + p["b1"])
        h = torch.nn.functional.silu(h @ p["W2"].T + p["b2"])
        h = torch.nn.functional.silu(h @ p["W3"].T + p["b3"])
        return h @ p["W4"].T + p["b4"]


def test_forward_backward_match_torch_autograd():
    w = mlp.init(7).astype(np.float64) * 1.7
    X, Y, _, _ = mlp.samples(96, 5)
    loss, g = mlp.loss_and_grad(w, X, Y)
    net = _Net(w)
    xt = torch.tensor(X, dtype=torch.float64)
    yt = torch.tensor(Y, dtype=torch.float64)
    lt = ((net(xt) - yt) ** 2).mean()
    lt.backward()
    gt = mlp.pack({k: v.grad.numpy() for k, v in net.p.items()})
    assert abs(loss - lt.item()) <= 1e-14 * abs(loss)
    np.testing.assert_allclose(g, gt, rtol=1e-10, atol=1e-14 * np.abs(gt).max())


def test_gradient_matches_finite_differences():
    w = mlp.init(11).astype(np.float64)
    X, Y, _, _ = mlp.samples(64, 9)
    _, g = mlp.loss_and_grad(w, X, Y)
    rng = np.random.default_rng(0)
    for k in rng.choice(mlp.NPARAM, 24, replace=False):
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (mlp.loss_and_grad(wp, X, Y)[0] - mlp.loss_and_grad(wm, X, Y)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(abs(g[k]), 1e-4)


def _c_dp(tr, c, phi, K, G):
    t = np.array([tr[0], tr[1], tr[2], 0.0, 0.0, 0.0])
    e1, dl = C.c_double(0.0), C.c_double(0.0)
    load().or_dp_return(6, t.ctypes.data_as(_dp), c, c, phi, 0.0, K, G, 0.0, C.byref(e1), C.byref(dl))
    return t[:3], dl.value


def test_principal_return_matches_c_oracle_for_psi_zero():
    X, Y, raw, aux = mlp.samples(3000, 12)
    for i in range(0, 3000, 7):
        tp, dl, br = mlp.dp_return_principal(aux["tr"][i], aux["c"][i], aux["phi"][i], 0.0, aux["K"][i],
                                             aux["G"][i])
        tc, dc = _c_dp(aux["tr"][i], aux["c"][i], aux["phi"][i], aux["K"][i], aux["G"][i])
        np.testing.assert_allclose(tp, tc, rtol=1e-14, atol=1e-14 * np.abs(aux["tr"][i]).max())
        assert abs(dl - dc) <= 1e-14 * max(abs(dc), 1e-300)


def test_psi_return_closed_form_invariants():
    X, Y, raw, aux = mlp.samples(20000, 13)
    assert aux["good"].all()
    tr, tp, dl = aux["tr"], raw[:, :3], raw[:, 3]
    K, G = aux["K"], aux["G"]
    eta, etaB, xi = mlp.material_param(aux["phi"], aux["psi"], aux["c"])
    P, Pn = tr.mean(1), tp.mean(1)
    tII = np.sqrt(0.5 * ((tr - P[:, None]) ** 2).sum(1))
    tIIn = np.sqrt(0.5 * ((tp - Pn[:, None]) ** 2).sum(1))
    scale = np.abs(tr).max(1) + xi
    # on the cone, pressure moved by K ηB Δλ, deviator shrunk by G Δλ, same direction
    assert np.all(np.abs(tIIn + eta * Pn - xi) <= 1e-12 * scale)
    assert np.all(np.abs(Pn - (P - K * etaB * dl)) <= 1e-12 * scale)
    assert np.all(np.abs(tIIn - (tII - G * dl)) <= 1e-11 * scale)
    assert (aux["psi"] > 0.1).sum() > 1000
    # scaled by the overstress, the correction is the closed-form flow direction
    fs = raw[:, 4]
    np.testing.assert_allclose(Y[:, 3], G / (G + K * eta * etaB), rtol=1e-6)
    # the residual check accepts the exact return
    r, shear = mlp.residual(Y.astype(np.float64), aux["e"], aux["E"], aux["nu"], aux["c"], aux["phi"], aux["psi"])
    assert shear.all()
    assert np.all(r < 1e-6)                          # targets are stored in float32


def test_residual_rejects_apex_and_past_apex_returns():
    # single draws of every branch
    X, Y, raw, aux = mlp.samples(20000, 13, regular_only=False)
    br = aux["branch"]
    assert (br == 0).any() and (br == 2).any() and (~aux["good"] & (br == 1)).any()
    eta, etaB, xi = mlp.material_param(aux["phi"], aux["psi"], aux["c"])
    # apex quirk Pn = σm - P (DP.jl:58)
    ap = br == 2
    Pa = aux["tr"][ap].mean(1)
    np.testing.assert_allclose(raw[ap, :3], np.repeat((xi[ap] / eta[ap] - Pa)[:, None], 3, 1), rtol=1e-13)
    # exact results of the other branches, expressed in the network's
    # overstress scaling, fail the check: past the apex (τn < 0, DP.jl:27-29
    # flips the deviator) the flow rule holds and the yield residual is 2|τn|
    fs = aux["fs"]
    with np.errstate(divide="ignore", invalid="ignore"):
        Ye = np.concatenate([(aux["tp"] - aux["tr"]) / fs[:, None], (aux["G"] * aux["dl"] / fs)[:, None]], 1)
    r, shear = mlp.residual(Ye, aux["e"], aux["E"], aux["nu"], aux["c"], aux["phi"], aux["psi"])
    q = (br == 1) & ~aux["good"] & (fs > 0)
    tn = xi[q] - eta[q] * raw[q, :3].mean(1)
    assert (tn < 0).all()
    np.testing.assert_allclose(r[q], 2 * np.abs(tn) / aux["S"][q], rtol=1e-9)
    assert not shear[ap].any()
    assert np.all(r[aux["good"]] < 1e-12)


def test_training_set_features_cover_unit_box():
    X, Y, raw, aux = mlp.samples(50000, 44)
    assert np.all(np.abs(X) <= 1.0 + 1e-6)
    assert np.all(aux["e"][:, 0] >= aux["e"][:, 1]) and np.all(aux["e"][:, 1] >= aux["e"][:, 2])
    assert np.isfinite(Y).all() and np.abs(Y).max() < 10.0
    assert 0.3 < X[:, 7].mean() + 1.0 < 1.7
