"""Pins of the oracle's dense linear algebra and of the finite-strain
elastic predictor (elast.jl:45-56) against library routines: numpy eigh /
det / solve and scipy.linalg logm / expm."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import scipy.linalg as sl

from oracle.oracle import load

_dp = C.POINTER(C.c_double)


def eig(A):
    d = A.shape[0]
    a = np.asfortranarray(A, dtype=np.float64)
    lam, V = np.zeros(d), np.zeros((d, d), order="F")
    load().or_sym_eig(d, a.ctypes.data_as(_dp), lam.ctypes.data_as(_dp), V.ctypes.data_as(_dp))
    return lam, V


def matfun(A, f):
    lam, V = eig(A)
    return (V * f(lam)) @ V.T


@pytest.mark.parametrize("d", [2, 3])
def test_sym_eig_random(d):
    rng = np.random.default_rng(1)
    for _ in range(200):
        M = rng.normal(size=(d, d))
        A = M @ M.T + 0.1 * np.eye(d)
        lam, V = eig(A)
        assert np.allclose(V.T @ V, np.eye(d), atol=1e-14)
        assert np.allclose(V @ np.diag(lam) @ V.T, A, atol=1e-13 * np.abs(A).max())
        assert np.allclose(np.sort(lam), np.linalg.eigvalsh(A), rtol=1e-13)


@pytest.mark.parametrize("d", [2, 3])
def test_log_exp_near_identity_and_degenerate(d):
    """b ≈ I (every particle at t=0) and repeated eigenvalues: the matrix
    functions stay accurate (elast.jl:51-52, DP.jl:65-66)."""
    rng = np.random.default_rng(2)
    for scale in (0.0, 1e-12, 1e-8, 1e-4, 1e-1):
        S = rng.normal(size=(d, d))
        S = 0.5 * (S + S.T)
        b = np.eye(d) + scale * S
        L = matfun(b, np.log)
        assert np.allclose(L, sl.logm(b).real, atol=1e-15 + 1e-13 * scale)
        E = matfun(0.5 * L, lambda x: np.exp(2 * x))
        assert np.allclose(E, b, atol=1e-14)
    # exactly repeated eigenvalue with a rotation
    Q, _ = np.linalg.qr(rng.normal(size=(d, d)))
    lam = np.array([1.2, 1.2, 0.7][:d])
    b = Q @ np.diag(lam) @ Q.T
    assert np.allclose(matfun(b, np.log), sl.logm(b).real, atol=1e-14)


@pytest.mark.parametrize("d", [2, 3])
def test_det_and_solve(d):
    rng = np.random.default_rng(3)
    L = load()
    for _ in range(100):
        A = np.asfortranarray(np.eye(d) + 0.3 * rng.normal(size=(d, d)))
        assert L.or_det(d, A.ctypes.data_as(_dp)) == pytest.approx(np.linalg.det(A), rel=1e-13, abs=1e-15)
        b = rng.normal(size=d)
        x = np.zeros(d)
        L.or_lu_solve(d, A.ctypes.data_as(_dp), b.ctypes.data_as(_dp), x.ctypes.data_as(_dp))
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-12, atol=1e-13)
    T = np.asfortranarray(np.triu(1 + rng.random((d, d))))
    assert L.or_det(d, T.ctypes.data_as(_dp)) == np.prod(np.diag(T))


def test_solve_6x6_del():
    from elastoplasm.ic import get_elastic_stiffness
    K, G, Del = get_elastic_stiffness(1e6, 0.3, 3)
    rng = np.random.default_rng(4)
    t = rng.normal(size=6) * 1e4
    x = np.zeros(6)
    load().or_lu_solve(6, np.asfortranarray(Del).ctypes.data_as(_dp), t.ctypes.data_as(_dp), x.ctypes.data_as(_dp))
    assert np.allclose(Del @ x, t, rtol=1e-12)
    # isotropic: K, G relations of get_elastic_stiffness (setup_cmpr.jl:21-48)
    assert Del[0, 0] == pytest.approx(K + 4 * G / 3, rel=1e-14)
    assert Del[0, 1] == pytest.approx(K - 2 * G / 3, rel=1e-14)
    assert Del[3, 3] == pytest.approx(G, rel=1e-14)
