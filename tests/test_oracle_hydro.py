"""Hydrostatic mode (t_press = 0): pressh (SuLi-IC/9_grad.f90:667-702) and
posdin's hydrostatic branch (SuLi-IC/10_pos.f90:24-37 and :48-93), restated in the C
oracle and pinned bit for bit against the reference's own compiled code, on
uniform and stretched (SuLi-ICdz) grids; plus what the formulas fix: a
constant density gives p(k) = gz rho (H - z_bottom(k)) + O(dz) exactly as
Simpson's rule, and the reconstructed w makes every column divergence-free.
"""
import numpy as np
import pytest

import oracle
from suli_amd import cases

needs_ref = pytest.mark.skipif(not oracle.ref_available("ic"), reason="reference not built")


def hydro_case(nx, ny, nz, seed, stretched=False):
    c = cases.random_case(nx, ny, nz, seed=seed, stretched=stretched, surface=0.55)
    c.grid.t_press = 0
    rng = np.random.default_rng(seed)
    c.fields.prd1[...] = rng.uniform(-1, 1, c.fields.prd1.shape)
    g = c.grid
    # a Heaviside-like field: 1 in water, smooth band, exact 0 above
    hs = np.clip((c.fields.rho - 1.2) / (1000.0 - 1.2), 0.0, 1.0)
    hs[hs < 1e-3] = 0.0
    return c, np.asfortranarray(hs)


@needs_ref
@pytest.mark.parametrize("stretched", [False, True])
def test_pressh_matches_reference(stretched):
    c, _ = hydro_case(12, 9, 10, seed=11, stretched=stretched)
    got = oracle.pressh(c.grid, c.fields.rho, 9.81, c.fields.prd1)
    exp = oracle.ref_pressh(c.grid, c.fields.rho, 9.81, c.fields.prd1)
    np.testing.assert_array_equal(got, exp)


@needs_ref
@pytest.mark.parametrize("stretched", [False, True])
def test_posdin_hydro_matches_reference(stretched):
    c, hs = hydro_case(12, 9, 10, seed=12, stretched=stretched)
    got = oracle.posdin_hydro(c.grid, c.fields, hs, c.fields.prd1)
    exp = oracle.ref_posdin_hydro(c.grid, c.fields, hs, c.fields.prd1)
    for a, b in zip(got, exp):
        np.testing.assert_array_equal(a, b)


def test_pressh_constant_density_closed_form():
    g = cases.Grid(6, 5, 9, 0.01, 0.01, 0.02, t_press=0)
    rho = np.full((g.nx, g.ny, g.nz), 998.0, order="F")
    p = oracle.pressh(g, rho, 9.81, np.zeros((g.nx + 2, g.ny + 2, g.nz + 2), order="F"))
    # Simpson on a constant integrates exactly: p(k) = gz rho dz (nz - k + 1) for
    # the cells reached from the top in steps of two, and with the trapezoid
    # start at nz-1 the same value
    k = np.arange(1, g.nz + 1)
    expect = 9.81 * 998.0 * g.dz * (g.nz - k + 1)
    np.testing.assert_allclose(p[1:-1, 1:-1, 1:-1], np.broadcast_to(expect, (g.nx, g.ny, g.nz)),
                               rtol=1e-13)


def test_posdin_hydro_columns_divergence_free():
    c, hs = hydro_case(10, 8, 7, seed=13)
    hs[...] = 1.0   # all water: nothing is zeroed afterwards
    g = c.grid
    u, v, w, prd = oracle.posdin_hydro(g, c.fields, hs, c.fields.prd1)
    div = ((u[2:g.nx + 2, 1:-1, 1:-1] - u[1:g.nx + 1, 1:-1, 1:-1]) / g.dx +
           (v[1:-1, 2:g.ny + 2, 1:-1] - v[1:-1, 1:g.ny + 1, 1:-1]) / g.dy +
           (w[1:-1, 1:-1, 2:g.nz + 2] - w[1:-1, 1:-1, 1:g.nz + 1]) / g.dz)
    assert np.abs(div).max() <= 1e-12 * (np.abs(u).max() / g.dx)
    assert np.all(w[:, :, 1] == 0.0)
    np.testing.assert_array_equal(prd, c.fields.prd1)


def test_posdin_hydro_horizontal_correction_is_posdins():
    """Both branches share the u, v correction: with hs = 1 everywhere the
    hydrostatic branch's u, v equal the non-hydrostatic posdin's."""
    c, hs = hydro_case(9, 7, 6, seed=14)
    hs[...] = 1.0
    u, v, _, _ = oracle.posdin_hydro(c.grid, c.fields, hs, c.fields.prd1)
    u1, v1, _, _ = oracle.posdin(c.grid, c.fields, c.fields.prd1)
    np.testing.assert_array_equal(u, u1)
    np.testing.assert_array_equal(v, v1)
    assert not np.array_equal(u, c.fields.u)
