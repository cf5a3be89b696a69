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


# ---- t_tens = 1: the surface-tension gradients hsx, hsy, hsz ---------------
# (SuLi-IC/14_level_set.f90:519-550,563-567 over der_weno, :338-370, and
# the WENO5 derivatives of 8_derivadas.f90:675-1095).  Pinned bit for bit
# against the reference's compiled heaviside with t_tens = 1, and by what
# WENO5 fixes: a linear field's derivative is exact away from the
# extrapolated ends, and a smooth field's converges at fifth order.

def wavy_level_set(g, seed):
    """ls = z - h(x, y): a wavy surface crossing mid-depth, with noise of a
    few percent of a cell so that the WENO weights vary."""
    rng = np.random.default_rng(seed)
    x = (np.arange(g.nx) + 0.5) * g.dx
    y = (np.arange(g.ny) + 0.5) * g.dy
    z = (np.arange(g.nz) + 0.5) * g.dz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    Lx, Ly = g.nx * g.dx, g.ny * g.dy
    h = 0.5 * g.nz * g.dz + 2 * g.dz * np.sin(2 * np.pi * X / Lx + 0.3) * np.cos(2 * np.pi * Y / Ly)
    return np.asfortranarray(Z - h + 0.03 * g.dz * rng.standard_normal(X.shape))


@needs_ref
@pytest.mark.parametrize("ccx0,ccy0", [(1, 1), (0, 1), (1, 0), (0, 0)])
def test_heaviside_tens_matches_reference(ccx0, ccy0):
    g = cases.Grid(14, 10, 12, 5e-3, 4e-3, 3e-3, ccx0=ccx0, ccy0=ccy0)
    ls = wavy_level_set(g, 50 + ccx0 + 2 * ccy0)
    alpha1, dx1 = 1.5, max(g.dx, g.dy, g.dz)
    eps = alpha1 * dx1
    got = oracle.heaviside_tens(g, ls, eps, 1.2, 1000.0, 1.8e-5, 1.0e-3)
    exp = oracle.ref_heaviside_tens(g, ls, alpha1, dx1, 1.2, 1000.0, 1.8e-5, 1.0e-3)
    band = (np.abs(ls) <= eps).sum()
    assert band > g.ncell // 10
    for a, b in zip(got, exp):
        np.testing.assert_array_equal(a, b)


def test_ls_gradient_linear_is_exact():
    g = cases.Grid(16, 14, 12, 0.1, 0.2, 0.05)
    x = np.arange(1, g.nx + 1) * g.dx
    y = np.arange(1, g.ny + 1) * g.dy
    z = np.arange(1, g.nz + 1) * g.dz
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    a, b, c = 0.7, -1.3, 2.1
    gx, gy, gz = oracle.ls_gradient(g, np.asfortranarray(a * X + b * Y + c * Z))
    # the stencils reach three cells: beyond them the extrapolated ends
    np.testing.assert_allclose(gx[3:-3], a, rtol=1e-11)
    np.testing.assert_allclose(gy[:, 3:-3], b, rtol=1e-11)
    np.testing.assert_allclose(gz[:, :, 3:-3], c, rtol=1e-11)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_ls_gradient_fifth_order(axis):
    errs = []
    for n in (40, 80):
        shape = [6, 6, 6]
        shape[axis] = n
        h = 1.0 / n
        g = cases.Grid(*shape, h, h, h)
        t = (np.arange(1, n + 1) - 0.5) * h
        f, df = np.sin(2 * np.pi * t + 0.4), 2 * np.pi * np.cos(2 * np.pi * t + 0.4)
        ls = np.asfortranarray(np.broadcast_to(f.reshape([-1 if d == axis else 1 for d in range(3)]), shape))
        gr = oracle.ls_gradient(g, ls)[axis]
        gr = np.moveaxis(gr, axis, 0)[:, 0, 0]
        errs.append(np.max(np.abs(gr - df)[n // 4: 3 * n // 4]))
    assert errs[0] / errs[1] > 2 ** 4.5, errs
