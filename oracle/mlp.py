"""CPU restatement of the learned return mapping (numpy, FP64).

TEST INFRASTRUCTURE ONLY: the checker the device network and its training
kernels are compared against.  Only tests/ and bench.py's reporting use it.

- ``samples``: the training-set generator (counter-based draws, trial
  principal log strains and material parameters, the FP64 Drucker-Prager
  return of the 3D formulas in src/home/program/workflow/elastoplast/plast/
  retmap/DP.jl:7-9, 18-22, 45-62 with materialParam's ψ argument).
- ``forward`` / ``backward``: the MLP 8-128-128-128-4 with SiLU and its MSE
  gradient; with ``bf16=True`` the operands are rounded where the device
  rounds them (inputs, weights, activations, SiLU', output-side gradients).
- ``adam``: one Adam step.

Parameter layout (shared with the device): [W1 (128x8), b1, W2 (128x128),
b2, W3 (128x128), b3, W4 (4x128), b4], W row-major [out][in].
"""
from __future__ import annotations

import numpy as np

NI, NH, NO = 8, 128, 4
SHAPES = [("W1", (NH, NI)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)), ("b3", (NH,)),
          ("W4", (NO, NH)), ("b4", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)
EPS_SCALE = 0.2
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def unpack(w):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(w[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in SHAPES])


# ---- bf16 -------------------------------------------------------------------
def bf16(x):
    """Round to bfloat16 (nearest, ties to even) and back, via float32."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)) << np.uint64(16)
    r = r.astype(np.uint32).view(np.float32)
    nan = np.isnan(f)
    if nan.any():
        r = np.where(nan, f, r)
    return r.astype(np.float64)


# ---- counter-based draws ------------------------------------------------------
def splitmix_mix(z):
    """splitmix64's output function (Steele, Lea, Flood 2014)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z ^ (z >> np.uint64(30))
        z = z * np.uint64(0xBF58476D1CE4E5B9)
        z = z ^ (z >> np.uint64(27))
        z = z * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def u01(seed: int, i, k: int):
    i = np.asarray(i, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + \
            (i * np.uint64(8) + np.uint64(k + 1)) * np.uint64(0xD1B54A32D192ED03)
    return (splitmix_mix(z) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


# ---- Drucker-Prager, 3D principal ------------------------------------------------
def material_param(phi, psi, c):
    """materialParam, 3D branch (DP.jl:18-22)."""
    sp, sq = np.sin(phi), np.sin(psi)
    eta = 6.0 * sp / (np.sqrt(3.0) * (3.0 + sp))
    etaB = 6.0 * sq / (np.sqrt(3.0) * (3.0 + sq))
    xi = 6.0 * c * np.cos(phi) / (np.sqrt(3.0) * (3.0 + sp))
    return eta, etaB, xi


def dp_return_principal(tr, c, phi, psi, K, G):
    """finite_DP's closed-form return (DP.jl:40-62) on principal stresses
    tr[..., 3] (shear components zero), Hp = 0, c0 = cr = c.
    Returns (tau', dlam, branch) with branch 0 elastic, 1 shear, 2 apex."""
    tr = np.asarray(tr, dtype=np.float64)
    P = (tr[..., 0] + tr[..., 1] + tr[..., 2]) / 3.0
    t0 = tr - P[..., None]
    tII = np.sqrt(0.5 * (t0[..., 0] ** 2 + t0[..., 1] ** 2 + t0[..., 2] ** 2))
    eta, etaB, xi = material_param(phi, psi, c)
    sm = xi / eta
    tP = xi - eta * (xi / eta)
    fs = tII + eta * P - xi
    h = tII - tP - np.sqrt(1.0 + eta ** 2) * (P - sm)
    out = tr.copy()
    dl = np.zeros_like(P)
    br = np.zeros(P.shape, dtype=np.int64)
    shear = ((fs > 0) & (P < sm)) | ((h > 0) & (P >= sm))
    with np.errstate(divide="ignore", invalid="ignore"):
        d1 = fs / (G + K * eta * etaB)
        Pn = P - K * etaB * d1
        tn = xi - eta * (P - K * etaB * d1)
        r = tn / tII
    s1 = t0 * r[..., None] + Pn[..., None]
    out = np.where(shear[..., None], s1, out)
    dl = np.where(shear, d1, dl)
    br = np.where(shear, 1, br)
    apex = (h <= 0) & (P >= sm)
    d2 = (P - sm) / K
    Pa = sm - P
    with np.errstate(divide="ignore", invalid="ignore"):
        s2 = t0 * (0.0 / tII)[..., None] + Pa[..., None]
    out = np.where(apex[..., None], s2, out)
    dl = np.where(apex, d2, dl)
    br = np.where(apex, 2, br)
    return out, dl, br


def features(e, E, nu, c, phi, psi):
    d2r = np.pi / 180.0
    x = np.empty(np.shape(E) + (8,))
    x[..., 0:3] = np.asarray(e) / EPS_SCALE
    x[..., 3] = (np.log10(E) - 6.5) / 1.5
    x[..., 4] = (nu - 0.325) / 0.125
    x[..., 5] = np.log10(c) - 4.0
    x[..., 6] = (phi - 30.0 * d2r) / (15.0 * d2r)
    x[..., 7] = np.where(phi > 0, 2.0 * psi / np.where(phi > 0, phi, 1.0) - 1.0, -1.0)
    return x


def _draw(n, seed, first):
    return _draw_at(np.arange(first, first + n, dtype=np.uint64), seed)


SAMPLE_TRIES = 128


def samples(n: int, seed: int, first: int = 0, regular_only: bool = True):
    """Inputs X [n][8] (float32), targets Y [n][4] (float32, scaled by the
    trial overstress fs) and raw [n][6] FP64 (τ'1, τ'2, τ'3, Δλ, fs, S) of
    samples first..first+n-1.  With regular_only (the training set) a draw
    that is not a shear return onto the cone is redrawn with the seeds
    seed + a*0x632BE59BD9B4E019, a = 1, 2, ...; otherwise every first draw is
    kept (any branch; fs and Y are then 0 where the draw is not regular)."""
    aux = _draw(n, seed, first)
    if regular_only:
        for a in range(1, SAMPLE_TRIES):
            bad = ~aux["good"]
            if not bad.any():
                break
            with np.errstate(over="ignore"):
                sd = int((np.uint64(seed) + np.uint64(a) * np.uint64(0x632BE59BD9B4E019)) & _M64)
            idx = np.nonzero(bad)[0]
            # redraw only the rejected samples (sample index keeps its value)
            d = _draw_at(idx.astype(np.uint64) + np.uint64(first), sd)
            for k in aux:
                aux[k][idx] = d[k]
    g = aux["good"]
    X = features(aux["e"], aux["E"], aux["nu"], aux["c"], aux["phi"], aux["psi"])
    Y = np.zeros((n, 4))
    with np.errstate(divide="ignore", invalid="ignore"):
        Y[:, :3] = np.where(g[:, None], (aux["tp"] - aux["tr"]) / aux["fs"][:, None], 0.0)
        Y[:, 3] = np.where(g, aux["G"] * aux["dl"] / aux["fs"], 0.0)
    raw = np.stack([aux["tp"][:, 0], aux["tp"][:, 1], aux["tp"][:, 2], aux["dl"], np.where(g, aux["fs"], 0.0),
                    aux["S"]], axis=-1)
    return X.astype(np.float32), Y.astype(np.float32), raw, aux


def _draw_at(idx, seed):
    """One draw (seed) for each sample index in idx, and its FP64 return."""
    i = np.asarray(idx, dtype=np.uint64)
    e = np.stack([0.4 * u01(seed, i, k) - 0.2 for k in range(3)], axis=-1)
    e = -np.sort(-e, axis=-1)
    d2r = np.pi / 180.0
    E = 10.0 ** (5.0 + 3.0 * u01(seed, i, 3))
    nu = 0.2 + 0.25 * u01(seed, i, 4)
    c = 10.0 ** (3.0 + 2.0 * u01(seed, i, 5))
    phi = (15.0 + 30.0 * u01(seed, i, 6)) * d2r
    psi = phi * u01(seed, i, 7)
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    D11, D12 = K + 4.0 / 3.0 * G, K - 2.0 / 3.0 * G
    tr = np.empty_like(e)
    for k in range(3):
        a = [D11 if j == k else D12 for j in range(3)]
        tr[:, k] = a[0] * e[:, 0] + a[1] * e[:, 1] + a[2] * e[:, 2]
    tp, dl, br = dp_return_principal(tr, c, phi, psi, K, G)
    eta, etaB, xi = material_param(phi, psi, c)
    P = tr.mean(-1)
    fs = np.sqrt(0.5 * ((tr - P[:, None]) ** 2).sum(-1)) + eta * P - xi
    S = np.sqrt((tr ** 2).sum(axis=-1)) + xi
    good = (br == 1) & (fs > 0) & (xi - eta * tp.mean(-1) >= 0)
    return dict(e=e, E=E, nu=nu, c=c, phi=phi, psi=psi, K=K, G=G, tr=tr, tp=tp, dl=dl, branch=br, fs=fs, S=S,
                good=good)


# ---- network ------------------------------------------------------------------
def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(w, X, bf16_ops: bool = False):
    """y [n][4] and the cache for backward."""
    p = unpack(w)
    q = (lambda a: bf16(a)) if bf16_ops else (lambda a: np.asarray(a, dtype=np.float64))
    W = {k: (q(v) if k.startswith("W") else v) for k, v in p.items()}
    h0 = q(X)
    z1 = h0 @ W["W1"].T + p["b1"]
    h1 = q(_silu(z1))
    z2 = h1 @ W["W2"].T + p["b2"]
    h2 = q(_silu(z2))
    z3 = h2 @ W["W3"].T + p["b3"]
    h3 = q(_silu(z3))
    y = h3 @ W["W4"].T + p["b4"]
    return y, dict(W=W, h0=h0, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, q=q)


def loss_and_grad(w, X, Y, bf16_ops: bool = False):
    """MSE over the batch and the 4 outputs, and its gradient (flat)."""
    y, c = forward(w, X, bf16_ops)
    n = len(X)
    r = y - np.asarray(Y, dtype=np.float64)
    loss = float((r ** 2).sum() / (4 * n))
    q, W = c["q"], c["W"]
    g4 = q(2.0 * r / (4 * n))
    grads = {"W4": g4.T @ c["h3"], "b4": g4.sum(0)}
    g3 = q((g4 @ W["W4"]) * q(_dsilu(c["z3"])))
    grads.update(W3=g3.T @ c["h2"], b3=g3.sum(0))
    g2 = q((g3 @ W["W3"]) * q(_dsilu(c["z2"])))
    grads.update(W2=g2.T @ c["h1"], b2=g2.sum(0))
    g1 = q((g2 @ W["W2"]) * q(_dsilu(c["z1"])))
    grads.update(W1=g1.T @ c["h0"], b1=g1.sum(0))
    return loss, pack(grads)


def init(seed: int):
    """U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for every weight and bias."""
    k = np.arange(NPARAM, dtype=np.uint64)
    fan = np.where(k < NH * NI + NH, NI, NH).astype(np.float64)
    return ((2.0 * u01(seed, k, 0) - 1.0) / np.sqrt(fan)).astype(np.float32)


def adam(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v


# ---- residual check of one prediction -----------------------------------------------
def residual(y, e, E, nu, c, phi, psi=0.0):
    """Scaled residual max(|flow rule|, |yield|)/S of predicted outputs y[..., 4]
    (scaled by the trial overstress fs, as the targets) for trial principal
    strains e[..., 3], and whether the trial state takes the shear branch
    with fs > 0 (the only states a prediction may stand in for)."""
    K = E / (3.0 * (1.0 - 2.0 * nu))
    G = E / (2.0 * (1.0 + nu))
    D11, D12 = K + 4.0 / 3.0 * G, K - 2.0 / 3.0 * G
    e = np.asarray(e, dtype=np.float64)
    tr = np.stack([(D11 if k == 0 else D12) * e[..., 0] + (D11 if k == 1 else D12) * e[..., 1] +
                   (D11 if k == 2 else D12) * e[..., 2] for k in range(3)], axis=-1)
    P = tr.mean(-1)
    d = tr - P[..., None]
    tII = np.sqrt(0.5 * (d ** 2).sum(-1))
    eta, etaB, xi = material_param(phi, psi, c)
    S = np.sqrt((tr ** 2).sum(-1)) + xi
    sm = xi / eta
    tP = xi - eta * (xi / eta)
    fs = tII + eta * P - xi
    h = tII - tP - np.sqrt(1 + eta ** 2) * (P - sm)
    shear = (((fs > 0) & (P < sm)) | ((h > 0) & (P >= sm))) & (fs > 0)
    tp = tr + fs[..., None] * y[..., :3]
    dl = fs * y[..., 3] / G
    with np.errstate(divide="ignore", invalid="ignore"):
        Rf = np.abs(tp - (tr - (G * dl / tII)[..., None] * d - (K * etaB * dl)[..., None])).max(-1)
    Pp = tp.mean(-1)
    tIIp = np.sqrt(0.5 * ((tp - Pp[..., None]) ** 2).sum(-1))
    Ry = np.abs(tIIp + eta * Pp - xi)
    return np.maximum(Rf, Ry) / S, shear


def near_cone(aux, delta):
    """The same regular shear returns from trial states a fraction delta of
    the way from the returned stress to the original trial stress,
    τtr(δ) = τ' + δ (τtr - τ'): on the return path, so they return to the
    same τ' with overstress δ fs and the same overstress-scaled targets.
    Returns the inputs X and the aux fields of those trial states."""
    tp, tr = aux["tp"], aux["tr"]
    t = tp + delta * (tr - tp)
    K, G = aux["K"], aux["G"]
    # isotropic Del⁻¹ on principal values
    tt = t.sum(1)
    e = (t - ((K - 2.0 * G / 3.0) / (3.0 * K) * tt)[:, None]) / (2.0 * G)[:, None]
    out = dict(aux)
    out["e"] = e
    out["tr"] = t
    out["fs"] = delta * aux["fs"]
    X = features(e, aux["E"], aux["nu"], aux["c"], aux["phi"], aux["psi"])
    return X.astype(np.float32), out
