"""Seeded Gaussian random fields for material parameters.

The reference builds heterogeneous cohesion with GRFS_gauss
(src/home/program/utils/setup/mpts/GRF.jl:66-143), a spectral randomisation
method: a sum of Nh harmonics with random phases, amplitudes and wave vectors
drawn from the Gaussian spectrum, unseeded.  Here the harmonics are drawn
from a seeded numpy Generator and the sum is evaluated at every particle on
the device (ep_grf_gauss).  A wave vector is sqrt(2) times a standard normal
vector divided by the correlation length ℓ (the reference's rejection-sampled
radius ∝ k² exp(-k²/2) with a uniform direction, GRF.jl:102-125, is the same
distribution), so the unit-variance field has covariance exp(-(r/ℓ)²).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import hip

_dp = C.POINTER(C.c_double)


def harmonics(seed: int, dim: int, ell, nh: int = 5000):
    """Wave vectors [nh][dim] and amplitudes [nh][2] of one field."""
    rng = np.random.default_rng(seed)
    ell = np.broadcast_to(np.asarray(ell, dtype=np.float64), (dim,))
    waves = np.sqrt(2.0) * rng.standard_normal((nh, dim)) / ell
    ab = rng.standard_normal((nh, 2))
    return np.ascontiguousarray(waves), np.ascontiguousarray(ab)


def gauss_field(x, seed: int, ell, nh: int = 5000) -> np.ndarray:
    """Unit-variance Gaussian field at the points x [dim, n] (column-major,
    as mpts.x), evaluated on the device."""
    x = np.asfortranarray(x, dtype=np.float64)
    dim, n = x.shape
    waves, ab = harmonics(seed, dim, ell, nh)
    g = np.zeros(n)
    L = hip.load()
    L.ep_grf_gauss.argtypes = [C.c_int32, C.c_int64, _dp, C.c_int32, _dp, _dp, _dp]
    hip._check(L.ep_grf_gauss(dim, n, x.ctypes.data_as(_dp), nh, waves.ctypes.data_as(_dp),
                              ab.ctypes.data_as(_dp), g.ctypes.data_as(_dp)))
    return g


def lognormal_field(x, mean: float, cov: float, seed: int, ell, nh: int = 5000) -> np.ndarray:
    """Log-normal field with the given mean and coefficient of variation."""
    s2 = np.log1p(cov * cov)
    mu = np.log(mean) - 0.5 * s2
    return np.exp(mu + np.sqrt(s2) * gauss_field(x, seed, ell, nh))


def set_cohesion_field(ic, mean: float = 20.0e3, cov: float = 0.3, ell: float = 5.0, seed: int = 42,
                       nh: int = 5000):
    """Heterogeneous cohesion c0 = cr (no softening) of a slump's particles."""
    c = lognormal_field(ic.mpts.x, mean, cov, seed, ell, nh)
    ic.mpts.s.c0[:] = c
    ic.mpts.s.cr[:] = c
    return c
