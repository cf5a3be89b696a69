"""Hydrostatic mode (t_press = 0) on the GPU: suli_ps_pressh[_dev] and
suli_ps_posdin_hydro[_dev] against the C oracle, which is itself pinned bit
for bit against the compiled reference (tests/test_oracle_hydro.py).

Both are straight sweeps of FP64 arithmetic in the reference's operation
order (the kernels are compiled without FMA contraction), so the bar is
bit-exact equality, at small and ragged sizes and at 256x128x96."""
import dataclasses

import numpy as np
import pytest

import oracle
from suli_amd import cases

pytestmark = pytest.mark.gpu

GZ = 9.80665


@pytest.fixture(scope="module")
def solver():
    from suli_amd.solver import PressureSolver
    s = PressureSolver(verbose=False)
    yield s
    s.close()


def hydro_case(nx, ny, nz, seed, stretched=False, periodic_x=False):
    c = cases.random_case(nx, ny, nz, seed=seed, stretched=stretched, surface=0.55,
                          periodic_x=periodic_x)
    c.grid.t_press = 0
    rng = np.random.default_rng(seed)
    c.fields.prd1[...] = rng.uniform(-1, 1, c.fields.prd1.shape)
    hs = np.clip((c.fields.rho - 1.2) / (1000.0 - 1.2), 0.0, 1.0)
    hs[hs < 1e-3] = 0.0
    c.fields.hs = np.asfortranarray(hs)
    return c


SIZES = [
    ("small", dict(nx=12, ny=9, nz=10, seed=21)),
    ("nz2", dict(nx=7, ny=5, nz=2, seed=22)),
    ("nz3_odd", dict(nx=13, ny=3, nz=3, seed=23)),
    ("stretched", dict(nx=20, ny=12, nz=17, seed=24, stretched=True)),
    ("periodic_x", dict(nx=33, ny=10, nz=9, seed=25, periodic_x=True)),
    # two 64-cell tiles along x (the second partial), a partial row of
    # tiles along y, columns of 19 planes (two blocks of eight and three)
    ("ragged_tiles", dict(nx=70, ny=33, nz=19, seed=26)),
    ("ragged_tiles_stretched", dict(nx=70, ny=33, nz=21, seed=27, stretched=True)),
]


@pytest.mark.parametrize("name,kw", SIZES, ids=[s[0] for s in SIZES])
def test_pressh_dev_bitwise(solver, name, kw):
    c = hydro_case(**kw)
    print(f"pressh {name}", flush=True)
    exp = oracle.pressh(c.grid, c.fields.rho, GZ, c.fields.prd1)
    df = solver.upload(c.grid, c.fields)
    try:
        solver.pressh_dev(c.grid, df, GZ)
        got = df.download("prd1")
    finally:
        df.free()
    np.testing.assert_array_equal(got, exp)
    # the halo is not written
    halo = np.ones(got.shape, bool)
    halo[1:-1, 1:-1, 1:-1] = False
    np.testing.assert_array_equal(got[halo], c.fields.prd1[halo])


@pytest.mark.parametrize("name,kw", SIZES, ids=[s[0] for s in SIZES])
def test_posdin_hydro_dev_bitwise(solver, name, kw):
    c = hydro_case(**kw)
    print(f"posdin_hydro {name}", flush=True)
    u, v, w, _ = oracle.posdin_hydro(c.grid, c.fields, c.fields.hs, c.fields.prd1)
    df = solver.upload(c.grid, c.fields)
    try:
        solver.posdin_hydro_dev(c.grid, df)
        np.testing.assert_array_equal(df.download("u"), u)
        np.testing.assert_array_equal(df.download("v"), v)
        np.testing.assert_array_equal(df.download("w"), w)
    finally:
        df.free()


def test_hydro_host_entry_points(solver):
    c = hydro_case(16, 11, 9, seed=26, stretched=True)
    exp_p = oracle.pressh(c.grid, c.fields.rho, GZ, c.fields.prd1)
    f = c.fields.copy()
    solver.pressh(c.grid, f, GZ)
    np.testing.assert_array_equal(f.prd1, exp_p)
    u, v, w, prd_exp = oracle.posdin_hydro(c.grid, f, f.hs, f.prd1)
    prd = solver.posdin_hydro(c.grid, f)
    np.testing.assert_array_equal(f.u, u)
    np.testing.assert_array_equal(f.v, v)
    np.testing.assert_array_equal(f.w, w)
    np.testing.assert_array_equal(prd, prd_exp)


def test_hydro_step_full_size(solver):
    """pressh then posdin at 256x128x96 (the two-phase tank's size), bit
    for bit; plus what holds whatever the size: a constant-density column
    integrates exactly, and every all-water column is divergence-free."""
    c = hydro_case(256, 128, 96, seed=27)
    g = c.grid
    print("hydro full 256x128x96", flush=True)
    p_exp = oracle.pressh(g, c.fields.rho, GZ, c.fields.prd1)
    f = c.fields.copy()
    f.prd1[...] = p_exp
    u, v, w, _ = oracle.posdin_hydro(g, f, f.hs, p_exp)
    df = solver.upload(g, c.fields)
    try:
        solver.pressh_dev(g, df, GZ)
        np.testing.assert_array_equal(df.download("prd1"), p_exp)
        solver.posdin_hydro_dev(g, df)
        ug, vg, wg = df.download("u"), df.download("v"), df.download("w")
    finally:
        df.free()
    np.testing.assert_array_equal(ug, u)
    np.testing.assert_array_equal(vg, v)
    np.testing.assert_array_equal(wg, w)
    # constant density: p(k) = gz rho dz (nz - k + 1)
    rho = np.full((g.nx, g.ny, g.nz), 998.0, order="F")
    f2 = c.fields.copy()
    f2.rho = rho
    df = solver.upload(g, f2)
    try:
        solver.pressh_dev(g, df, GZ)
        p = df.download("prd1")
    finally:
        df.free()
    k = np.arange(1, g.nz + 1)
    np.testing.assert_allclose(p[1:-1, 1:-1, 1:-1],
                               np.broadcast_to(GZ * 998.0 * g.dz * (g.nz - k + 1), (g.nx, g.ny, g.nz)),
                               rtol=1e-12)
    # a cell none of whose faces was zeroed (face hs >= 1e-4 all round) is
    # divergence-free: w was built from continuity
    hx, hy, hz = oracle.interp_faces(g, c.fields.hs)
    keep = ((hx[:-1] >= 1e-4) & (hx[1:] >= 1e-4) & (hy[:, :-1] >= 1e-4) & (hy[:, 1:] >= 1e-4) &
            (hz[:, :, :-1] >= 1e-4) & (hz[:, :, 1:] >= 1e-4))
    assert keep.sum() > g.ncell // 4
    div = ((ug[2:g.nx + 2, 1:-1, 1:-1] - ug[1:g.nx + 1, 1:-1, 1:-1]) / g.dx +
           (vg[1:-1, 2:g.ny + 2, 1:-1] - vg[1:-1, 1:g.ny + 1, 1:-1]) / g.dy +
           (wg[1:-1, 1:-1, 2:g.nz + 2] - wg[1:-1, 1:-1, 1:g.nz + 1]) / g.dz)
    assert np.abs(div[keep]).max() <= 1e-9 * np.abs(ug).max() / g.dx


def test_hydro_argument_errors(solver):
    from suli_amd.solver import SuliPsError
    c = hydro_case(8, 6, 5, seed=28)
    g = dataclasses.replace(c.grid, nz=1)
    with pytest.raises(SuliPsError, match="dimension"):
        solver.pressh(g, c.fields, GZ)
    f = c.fields.copy()
    f.hs = None
    with pytest.raises(SuliPsError, match="hs"):
        solver.posdin_hydro(c.grid, f)
