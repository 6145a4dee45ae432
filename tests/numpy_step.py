"""An independent, slow restatement of one explicit MPM step in numpy, for
tiny inputs only.  It follows the same reference files as the C oracle
(shpfun/, mapsto/, elastoplast/) but is written separately, with library
linear algebra (numpy eigh/det/solve), so that a bookkeeping slip in either
(an index, a sign, a transposed operand) shows up as a disagreement."""
from __future__ import annotations

import math

import numpy as np


def bspline(xi, t):
    a = abs(xi)
    if t == 3:
        if a <= 1:
            return 2 / 3 - xi * xi + a ** 3 / 2, (-2 * xi + 1.5 * xi * a)
        if a <= 2:
            return (2 - a) ** 3 / 6, -np.sign(xi) * (2 - a) ** 2 / 2
        return 0.0, 0.0
    # boundary-modified pieces, bsmpm.jl:1-62
    x2, x3 = xi * xi, xi ** 3
    if t == 1:
        if -2 <= xi <= -1: return x3 / 6 + x2 + 2 * xi + 4 / 3, x2 / 2 + 2 * xi + 2
        if -1 <= xi <= 0: return -x3 / 6 + xi + 1, -x2 / 2 + 1
        if 0 <= xi <= 1: return x3 / 6 - xi + 1, x2 / 2 - 1
        if 1 <= xi <= 2: return -x3 / 6 + x2 - 2 * xi + 4 / 3, -x2 / 2 + 2 * xi - 2
    if t == 2:
        if -1 <= xi <= 0: return -x3 / 3 - x2 + 2 / 3, -x2 - 2 * xi
        if 0 <= xi <= 1: return x3 / 2 - x2 + 2 / 3, 1.5 * x2 - 2 * xi
        if 1 <= xi <= 2: return -x3 / 6 + x2 - 2 * xi + 4 / 3, -x2 / 2 + 2 * xi - 2
    if t == 4:
        if -2 <= xi <= -1: return x3 / 6 + x2 + 2 * xi + 4 / 3, x2 / 2 + 2 * xi + 2
        if -1 <= xi <= 0: return -x3 / 2 - x2 + 2 / 3, -1.5 * x2 - 2 * xi
        if 0 <= xi <= 1: return x3 / 3 - x2 + 2 / 3, x2 - 2 * xi
    return 0.0, 0.0


def node_type(xn, lo, hi, h):
    if xn == lo or xn == hi:
        return 1
    if xn == lo + h:
        return 2
    if xn == hi - h:
        return 4
    return 3


def stencil(mesh, xp):
    """[(node id 0-based, N, dN[dim])] for one particle."""
    d = mesh.dim
    axes = mesh.axes
    el = [min(max(int(math.floor((xp[i] - mesh.x0[i]) / mesh.h[i])), 0), int(mesh.nel[i]) - 1) for i in range(d)]
    w = []
    for i in range(d):
        wi = {}
        for k in range(-1, 3):
            n = el[i] + k
            if 0 <= n < len(axes[i]):
                t = node_type(axes[i][n], mesh.xB[i, 0], mesh.xB[i, 1], mesh.h[i])
                f, df = bspline((xp[i] - axes[i][n]) / mesh.h[i], t)
                wi[n] = (f, df / mesh.h[i])
        w.append(wi)
    out = []
    if d == 2:
        for ix, (fx, dx) in w[0].items():
            for iz, (fz, dz) in w[1].items():
                out.append((iz + len(axes[1]) * ix, fx * fz, np.array([dx * fz, fx * dz])))
    else:
        nz, nx = len(axes[2]), len(axes[0])
        for iy, (fy, dy) in w[1].items():
            for ix, (fx, dx) in w[0].items():
                for iz, (fz, dz) in w[2].items():
                    out.append((iz + nz * (ix + nx * iy), fx * fy * fz,
                                np.array([dx * fy * fz, fx * dy * fz, fx * fy * dz])))
    return out


def voigt(e, d, X):
    return np.array([e[0, 0], e[1, 1], X * e[0, 1]]) if d == 2 else \
        np.array([e[0, 0], e[1, 1], e[2, 2], X * e[1, 2], X * e[0, 2], X * e[0, 1]])


def tensor(v, d, X):
    if d == 2:
        return np.array([[v[0], X * v[2]], [X * v[2], v[1]]])
    return np.array([[v[0], X * v[5], X * v[4]], [X * v[5], v[1], X * v[3]], [X * v[4], X * v[3], v[2]]])


def fsym(A, f):
    lam, V = np.linalg.eigh(0.5 * (A + A.T))
    return (V * f(lam)) @ V.T


def dp(tau, c0, cr, phi, Hp, K, G, e2, d):
    c = max(c0 + Hp * e2, cr)
    nd = d
    if d == 2:
        P = (tau[0] + tau[1]) / 2
        s = tau - np.array([P, P, 0])
        tII = math.sqrt(0.5 * (s[0] ** 2 + s[1] ** 2) + s[2] ** 2)
        eta = 3 * math.tan(phi) / math.sqrt(9 + 12 * math.tan(phi) ** 2)
        xi = 3 * c / math.sqrt(9 + 12 * math.tan(phi) ** 2)
    else:
        P = tau[:3].mean()
        s = tau - np.array([P, P, P, 0, 0, 0])
        tII = math.sqrt(0.5 * np.sum(s[:3] ** 2) + np.sum(s[3:] ** 2))
        eta = 6 * math.sin(phi) / (math.sqrt(3) * (3 + math.sin(phi)))
        xi = 6 * c * math.cos(phi) / (math.sqrt(3) * (3 + math.sin(phi)))
    sm = xi / eta
    fs = tII + eta * P - xi
    h = tII - 0.0 - math.sqrt(1 + eta ** 2) * (P - sm)
    iso = np.zeros_like(tau)
    iso[:nd] = 1.0
    if (fs > 0 and P < sm) or (h > 0 and P >= sm):
        dl = fs / G
        return s * ((xi - eta * P) / tII) + P * iso, e2 + dl * math.sqrt(1 / 3), dl
    if h <= 0 and P >= sm:
        dl = (P - sm) / K
        return (sm - P) * iso, e2 + math.sqrt(2) * dl / 3, dl
    return tau, e2, 0.0


def step(mesh, mp, cm, ops, g, dt, plast=True):
    """finite, bsmpm, std or mUSL, F-bar optional, DP; mutates mp/mesh."""
    d = mesh.dim
    nno = int(mesh.nno[-1])
    st = [stencil(mesh, mp.x[:, p]) for p in range(mp.nmp)]
    ms = mp.s.rho * mp.Om
    sig = mp.s.tau / mp.J
    m, pn, f = np.zeros(nno), np.zeros((d, nno)), np.zeros((d, nno))
    for p in range(mp.nmp):
        S = tensor(sig[:, p], d, 1.0)
        for no, N, dN in st[p]:
            m[no] += N * ms[p]
            pn[:, no] += N * ms[p] * mp.s.v[:, p]
            f[:, no] -= mp.Om[p] * (S @ dN)
            f[-1, no] += N * ms[p] * g[-1]
    a, v = np.zeros((d, nno)), np.zeros((d, nno))
    for no in range(nno):
        if m[no] == 0:
            continue
        for k in range(d):
            if mesh.bcs.status[k, no]:
                continue
            mi = 1 / m[no]
            D = ops["damping"] * np.linalg.norm(f[:, no]) * np.sign(pn[k, no] * mi)
            if abs(pn[k, no] * mi) >= 1e-3:
                f[k, no] -= D
            a[k, no] = f[k, no] * mi
            v[k, no] = (pn[k, no] + dt * f[k, no]) * mi
    C = ops["C_pf"]
    for p in range(mp.nmp):
        vp = sum(N * v[:, no] for no, N, _ in st[p])
        ap = sum(N * a[:, no] for no, N, _ in st[p])
        vnew = C * (mp.s.v[:, p] + dt * ap) + (1 - C) * vp
        mp.s.v[:, p] = vnew
        mp.x[:, p] += dt * vp
    if ops["musl"]:
        pa = np.zeros((d, nno))
        for p in range(mp.nmp):
            for no, N, _ in st[p]:
                pa[:, no] += N * ms[p] * mp.s.v[:, p]
        v = np.zeros((d, nno))
        for no in range(nno):
            if m[no] != 0:
                v[:, no] = np.where(mesh.bcs.status[:, no], 0.0, pa[:, no] / m[no])
    dFs, dJs = [], []
    for p in range(mp.nmp):
        L = sum(np.outer(v[:, no], dN) for no, _, dN in st[p])
        dF = np.eye(d) + dt * L
        mp.s.F[:, :, p] = dF @ mp.s.F[:, :, p]
        mp.J[p] = np.linalg.det(mp.s.F[:, :, p])
        mp.Om[p] = mp.J[p] * mp.Om0[p]
        mp.s.rho[p] = mp.s.rho0[p] / mp.J[p]
        dFs.append(dF)
        dJs.append(np.linalg.det(dF))
    if ops["locking"]:
        dJn = np.zeros(nno)
        for p in range(mp.nmp):
            for no, N, _ in st[p]:
                dJn[no] += N * mp.s.rho[p] * mp.Om[p] * dJs[p]
        dJn = np.where(m != 0, dJn / np.where(m != 0, m, 1), 0.0)
        for p in range(mp.nmp):
            s = sum(N * dJn[no] / dJs[p] for no, N, _ in st[p])
            dFs[p] = dFs[p] * s ** (1 / d)
    for p in range(mp.nmp):
        b = dFs[p] @ mp.s.b[:, :, p] @ dFs[p].T
        eps = 0.5 * fsym(b, np.log)
        tau = cm.Del @ voigt(eps, d, 2.0)
        if plast:
            e2 = mp.s.epII[0, p]
            tau, e1, dl = dp(tau, mp.s.c0[p], mp.s.cr[p], mp.s.phi[p], cm.Hp, cm.Kc, cm.Gc, e2, d)
            mp.s.epII[0, p] = e1
            mp.s.dlam[p] = dl
            b = fsym(tensor(np.linalg.solve(cm.Del, tau), d, 0.5), lambda x: np.exp(2 * x))
        mp.s.b[:, :, p] = b
        mp.s.tau[:, p] = tau
