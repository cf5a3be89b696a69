"""Density from the level set: heaviside (SuLi-IC/14_level_set.f90:508-594,
t_tens = 0) restated in the C oracle, pinned bit for bit against the
reference's own compiled heaviside, and against what the formula fixes:
hs(0) = 1/2, hs(-l) = 1 - hs(l), the band edges, monotonicity, and rho, mu
as exact blends at hs = 0 and 1."""
import numpy as np
import pytest

import oracle
from suli_amd import cases

needs_ref = pytest.mark.skipif(not oracle.ref_available("ic"), reason="reference not built")

EPS = 1.5 * 5e-3


def level_set(shape, seed):
    rng = np.random.default_rng(seed)
    ls = np.asfortranarray(rng.uniform(-2 * EPS, 2 * EPS, shape))
    ls.flat[:5] = [0.0, EPS, -EPS, np.nextafter(EPS, 1), np.nextafter(-EPS, -1)]
    return ls


@needs_ref
def test_heaviside_matches_reference():
    g = cases.Grid(20, 12, 16, 5e-3, 5e-3, 5e-3)
    ls = level_set((g.nx, g.ny, g.nz), 31)
    got = oracle.heaviside(ls, EPS, 1.2, 1000.0, 1.8e-5, 1.0e-3)
    exp = oracle.ref_heaviside(g, ls, 1.5, 5e-3, 1.2, 1000.0, 1.8e-5, 1.0e-3)
    for a, b in zip(got, exp):
        np.testing.assert_array_equal(a, b)


def test_heaviside_properties():
    x = np.linspace(-2 * EPS, 2 * EPS, 4001)
    hs, rho, mu = oracle.heaviside(x, EPS, 1.2, 1000.0, 1.8e-5, 1.0e-3)
    assert hs[2000] == 0.5 and rho[2000] == 1.2 * 0.5 + 1000.0 * 0.5
    np.testing.assert_allclose(hs + hs[::-1], 1.0, atol=4e-16)
    assert np.all(np.diff(hs) >= -2e-16)   # monotone up to rounding
    assert np.all(hs[x < -EPS] == 0) and np.all(hs[x > EPS] == 1)
    assert np.all(rho[x < -EPS] == 1.2) and np.all(rho[x > EPS] == 1000.0)
    assert np.all(mu[x > EPS] == 1.0e-3)
    # continuous at the band edges: the smooth branch reaches 0 and 1 there
    edge = oracle.heaviside(np.array([-EPS, EPS]), EPS, 1.2, 1000.0, 0.0, 0.0)[0]
    np.testing.assert_allclose(edge, [0.0, 1.0], atol=1e-16)
    np.testing.assert_array_equal(oracle.heaviside_rho(x, EPS, 1.2, 1000.0), rho)
