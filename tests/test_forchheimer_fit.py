"""The Forchheimer fit (CPU): it recovers k1 and k2 from points that follow
the relation exactly, is unaffected by the order of the points, reports no
inertial term for Darcy points, and refuses fewer than two points."""
from __future__ import annotations

import math

import numpy as np
import pytest

from ns3d_amd import forchheimer


def test_fit_recovers_coefficients_exactly():
    mu, rho, k1, k2 = 1e-3, 1e3, 2.5e-10, 4.0e-5
    v = np.array([1e-5, 1e-4, 5e-4, 1e-3, 3e-3])
    g = mu / k1 * v + rho / k2 * v * v
    f = forchheimer.fit(g, v, mu, rho)
    assert f["k1"] == pytest.approx(k1, rel=1e-10)
    assert f["k2"] == pytest.approx(k2, rel=1e-8)
    assert f["rel_residual"] < 1e-12
    # the fit does not depend on the order of the points
    perm = np.array([3, 0, 4, 1, 2])
    f2 = forchheimer.fit(g[perm], v[perm], mu, rho)
    assert f2["k1"] == pytest.approx(f["k1"], rel=1e-12) and f2["k2"] == pytest.approx(f["k2"], rel=1e-9)


def test_fit_with_noise_and_darcy_points():
    rng = np.random.default_rng(3)
    mu, rho, k1, k2 = 1.0, 1.0, 0.8, 0.05
    v = np.linspace(0.01, 1.0, 12)
    g = (mu / k1 * v + rho / k2 * v * v) * (1 + 1e-4 * rng.standard_normal(v.size))
    f = forchheimer.fit(g, v, mu, rho)
    assert f["k1"] == pytest.approx(k1, rel=2e-3) and f["k2"] == pytest.approx(k2, rel=2e-3)
    # pure Darcy flow: no inertial coefficient
    d = forchheimer.fit(mu / k1 * v, v, mu, rho)
    assert d["k1"] == pytest.approx(k1, rel=1e-12)
    assert abs(d["b"]) < 1e-9 * d["a"] and (d["k2"] == math.inf or d["k2"] > 1e8)


def test_fit_rejects_bad_input():
    with pytest.raises(ValueError):
        forchheimer.fit([1.0], [1.0], 1.0, 1.0)
    with pytest.raises(ValueError):
        forchheimer.fit([1.0, 2.0], [1.0, float("nan")], 1.0, 1.0)
    with pytest.raises(ValueError):
        forchheimer.fit([1.0, 2.0, 3.0], [1.0, 2.0], 1.0, 1.0)
