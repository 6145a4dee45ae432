"""Seeded Gaussian random field for cohesion (elastoplasm.grf): the spectrum
the harmonics are drawn from gives the squared-exponential covariance
exp(-(r/ℓ)²) (a property of the spectral method of GRF.jl:66-143), and the
device sum matches a numpy evaluation of the same harmonics."""
from __future__ import annotations

import numpy as np
import pytest

from elastoplasm import grf


def numpy_field(x, waves, ab):
    ph = waves @ x                       # [nh, n]
    return (ab[:, :1] * np.sin(ph) + ab[:, 1:] * np.cos(ph)).sum(0) / np.sqrt(len(waves))


@pytest.mark.parametrize("dim", [2, 3])
def test_spectrum_gives_squared_exponential_covariance(dim):
    ell = 5.0
    waves, ab = grf.harmonics(42, dim, ell, nh=200000)
    for r in (0.0, 2.0, 5.0, 8.0):
        for axis in range(dim):
            d = np.zeros(dim)
            d[axis] = r
            rho = np.cos(waves @ d).mean()          # E[g(x) g(x + r)] over the spectrum
            assert abs(rho - np.exp(-(r / ell) ** 2)) < 0.01, (r, axis, rho)


def test_amplitudes_give_unit_variance():
    waves, ab = grf.harmonics(7, 3, 5.0, nh=5000)
    x = np.random.default_rng(0).uniform(0, 400, size=(3, 4000))
    g = numpy_field(x, waves, ab)
    assert abs(g.mean()) < 0.1 and abs(g.var() - 1.0) < 0.15


def test_lognormal_moments_formula():
    # exp(mu + s g) with s² = log(1 + cov²), mu = log(mean) - s²/2 has the
    # requested mean and coefficient of variation for standard normal g
    mean, cov = 20.0e3, 0.3
    s2 = np.log1p(cov * cov)
    mu = np.log(mean) - 0.5 * s2
    assert np.isclose(np.exp(mu + 0.5 * s2), mean, rtol=1e-14)
    assert np.isclose(np.sqrt(np.expm1(s2)), cov, rtol=1e-14)


@pytest.mark.gpu
def test_device_field_matches_numpy():
    rng = np.random.default_rng(1)
    x = np.asfortranarray(rng.uniform(0, 100, size=(3, 3000)))
    g = grf.gauss_field(x, 42, 5.0, nh=777)
    waves, ab = grf.harmonics(42, 3, 5.0, nh=777)
    ref = numpy_field(x, waves, ab)
    np.testing.assert_allclose(g, ref, rtol=0, atol=1e-12 * np.abs(ref).max() * 10)
    c = grf.lognormal_field(x, 20.0e3, 0.3, 42, 5.0, nh=777)
    assert np.all(c > 0)
