"""Density from the level set on the GPU (suli_ps_heaviside_dev) against
the C oracle's heaviside, itself bit-exact with the reference's compiled
heaviside (tests/test_oracle_levelset.py).

The kernel evaluates the reference's expression in its order without FMA
contraction; the one operation outside IEEE basic arithmetic is sin(),
whose device implementation may differ from the host libm's in the last
place.  Tolerance: 2 ulp of 1 on hs (4.5e-16 absolute), carried into the
blends rho and mu as (phase 2 - phase 1) * 4.5e-16 plus 1e-15 relative;
most values are expected bit-exact."""
import numpy as np
import pytest

import oracle
from suli_amd import cases

pytestmark = pytest.mark.gpu

EPS = 1.5 * 5e-3


@pytest.fixture(scope="module")
def solver():
    from suli_amd.solver import PressureSolver
    s = PressureSolver(verbose=False)
    yield s
    s.close()


def run_dev(solver, g, ls, with_hs=True, with_mu=True):
    shape = (g.nx, g.ny, g.nz)
    p_ls = solver.dev_array(ls)
    p_rho = solver.dev_array(np.zeros(shape))
    p_hs = solver.dev_array(np.zeros(shape)) if with_hs else None
    p_mu = solver.dev_array(np.zeros(shape)) if with_mu else None
    try:
        solver.heaviside_dev(g, p_ls, EPS, 1000.0, 1.8e-5, 1.0e-3, p_rho, p_hs, p_mu)
        return (solver.dev_download(p_hs, shape) if with_hs else None,
                solver.dev_download(p_rho, shape),
                solver.dev_download(p_mu, shape) if with_mu else None)
    finally:
        for p in (p_ls, p_rho, p_hs, p_mu):
            if p is not None:
                solver.dev_free(p)


def test_heaviside_dev_vs_oracle(solver):
    g = cases.Grid(40, 24, 30, 5e-3, 5e-3, 5e-3, rho_f1=1.2)
    rng = np.random.default_rng(41)
    ls = np.asfortranarray(rng.uniform(-2 * EPS, 2 * EPS, (g.nx, g.ny, g.nz)))
    ls.flat[:5] = [0.0, EPS, -EPS, np.nextafter(EPS, 1), np.nextafter(-EPS, -1)]
    hs, rho, mu = run_dev(solver, g, ls)
    ehs, erho, emu = oracle.heaviside(ls, EPS, 1.2, 1000.0, 1.8e-5, 1.0e-3)
    assert np.abs(hs - ehs).max() <= 4.5e-16
    np.testing.assert_allclose(rho, erho, rtol=1e-15, atol=(1000.0 - 1.2) * 4.5e-16)
    np.testing.assert_allclose(mu, emu, rtol=1e-15, atol=(1.0e-3 - 1.8e-5) * 4.5e-16)
    # outside the band everything is exact
    out = np.abs(ls) > EPS
    assert np.array_equal(hs[out], ehs[out]) and np.array_equal(rho[out], erho[out])
    exact = np.mean(hs == ehs)
    print(f"heaviside bit-exact fraction {exact:.4f}")
    assert exact > 0.9
    # hs and mu are optional outputs
    _, rho2, _ = run_dev(solver, g, ls, with_hs=False, with_mu=False)
    np.testing.assert_array_equal(rho2, rho)


def test_graddin_from_level_set_full_size(solver):
    """The two-phase tank at 256x128x96 fed by its level set: heaviside_dev
    writes rho into the solve's device array, then graddin_dev; the result
    matches the solve of the host-built density to the solver tolerance."""
    c = cases.config2(seed=3)
    g = c.grid
    height = 0.25 * (g.nz * g.dx) / 0.48
    ls = cases.sloshing_level_set(g, height, 0.02 * (g.nz * g.dx) / 0.48)
    print("graddin from level set 256x128x96", flush=True)
    solver.configure(precision="mixed", stop_rule="relative", tol=1e-9, verbose=False)
    try:
        df = solver.upload(g, c.fields)
        try:
            st = solver.graddin_dev(g, df)
            assert st.status == 0
            p_host = df.download("prd1")
        finally:
            df.free()
        f = c.fields.copy()
        f.rho[...] = 0.0
        df = solver.upload(g, f)
        p_ls = solver.dev_array(ls)
        try:
            solver.heaviside_dev(g, p_ls, EPS, 1000.0, 0.0, 0.0, df.ptrs["rho"])
            np.testing.assert_allclose(df.download("rho"), c.fields.rho, rtol=1e-15,
                                       atol=(1000.0 - 1.2) * 4.5e-16)
            st = solver.graddin_dev(g, df)
            assert st.status == 0
            p_dev = df.download("prd1")
        finally:
            solver.dev_free(p_ls)
            df.free()
    finally:
        solver.configure(verbose=False)
    inner = (slice(1, -1),) * 3
    scale = np.abs(p_host[inner]).max()
    assert np.abs(p_dev[inner] - p_host[inner]).max() <= 1e-7 * scale


@pytest.mark.parametrize("ccx0,ccy0", [(1, 1), (0, 1), (1, 0), (0, 0)])
def test_heaviside_tens_dev_vs_oracle(solver, ccx0, ccy0):
    """t_tens = 1: the surface-tension gradients hsx, hsy, hsz against the
    oracle, itself bit-exact with the compiled reference heaviside
    (tests/test_oracle_levelset.py).  The WENO5 derivatives are IEEE basic
    arithmetic; hsx = 0.5 g (1 + cos(pi ls/eps))/eps also goes through
    cos, whose device result may differ in the last place: each gradient
    is within 0.5 |g| (3e-16)/eps + 4e-16 |hsx| of the oracle, outside the
    band exactly 0, and most values are bit-exact."""
    from tests.test_oracle_levelset import wavy_level_set
    g = cases.Grid(40, 24, 30, 5e-3, 4e-3, 3e-3, ccx0=ccx0, ccy0=ccy0, rho_f1=1.2)
    eps = 1.5 * max(g.dx, g.dy, g.dz)
    ls = wavy_level_set(g, 60 + ccx0 + 2 * ccy0)
    shape = (g.nx, g.ny, g.nz)
    exp = oracle.heaviside_tens(g, ls, eps, 1.2, 1000.0, 1.8e-5, 1.0e-3)
    grads = oracle.ls_gradient(g, ls)
    p_ls = solver.dev_array(ls)
    outs = [solver.dev_array(np.zeros(shape)) for _ in range(6)]
    try:
        solver.heaviside_tens_dev(g, p_ls, eps, 1000.0, 1.8e-5, 1.0e-3, outs[1], outs[3], outs[4], outs[5],
                                  hs_ptr=outs[0], mu_ptr=outs[2])
        got = [solver.dev_download(p, shape) for p in outs]
    finally:
        for p in [p_ls] + outs:
            solver.dev_free(p)
    hs, rho, mu = got[:3]
    assert np.abs(hs - exp[0]).max() <= 4.5e-16
    np.testing.assert_allclose(rho, exp[1], rtol=1e-15, atol=(1000.0 - 1.2) * 4.5e-16)
    band = np.abs(ls) <= eps
    assert band.mean() > 0.05
    for k in range(3):
        a, b, gr = got[3 + k], exp[3 + k], grads[k]
        assert np.all(a[~band] == 0.0)
        bound = 0.5 * np.abs(gr) * 3e-16 / eps + 4e-16 * np.abs(b)
        assert np.all(np.abs(a - b) <= bound), np.max(np.abs(a - b) / np.maximum(bound, 1e-300))
        exact = np.mean(a[band] == b[band])
        print(f"hs{'xyz'[k]} bit-exact fraction in the band {exact:.4f}")
        assert exact > 0.5
