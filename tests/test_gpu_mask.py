"""Solid-cell mask on the GPU (suli_ps_graddin_masked_dev and friends)
against the CPU oracle's masked restatement (itself pinned by brute force in
tests/test_oracle_mask.py).

Tolerances: pressure relative max-difference <= 1e-8 with both sides solved
to 1e-12 (FP64) / 1e-11 (mixed); posdin bit for bit from the same prd1.
"""
import numpy as np
import pytest

import oracle
from suli_amd import cases
from tests import helpers

pytestmark = pytest.mark.gpu

PRESSURE_RTOL = 1e-8


@pytest.fixture(scope="module")
def solver():
    from suli_amd.solver import PressureSolver
    s = PressureSolver(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
    yield s
    s.close()


def masked(nx, ny, nz, seed, **kw):
    c = cases.random_case(nx, ny, nz, seed=seed, **kw)
    c.fields.solid = cases.cylinder_cells(c.grid, 0.25 * ny * c.grid.dy)
    return c


MASKED = [
    ("walls", lambda: masked(24, 20, 12, 51)),
    ("px", lambda: masked(32, 16, 10, 52, periodic_x=True)),
    ("py_surface", lambda: masked(20, 24, 14, 53, periodic_y=True, surface=0.4)),
    ("cfg3_small", lambda: cases.config3(seed=54, nx=48, ny=24, nz=16)),
]


def solve_dev(solver, c):
    df = solver.upload(c.grid, c.fields)
    try:
        st = solver.graddin_dev(c.grid, df)
        return df.download("prd1"), st
    finally:
        df.free()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-11)])
@pytest.mark.parametrize("name,make", MASKED, ids=[m[0] for m in MASKED])
def test_masked_vs_oracle(solver, name, make, precision, tol):
    c = make()
    print(f"masked {name} {precision}", flush=True)
    p_or, st_or, rc = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-12,
                                     project_null=True)
    assert rc == 0
    solver.configure(precision=precision, stop_rule="relative", tol=tol, verbose=False)
    prd1, st = solve_dev(solver, c)
    assert st.status == 0, st.as_dict()
    mask = helpers.face_mask(prd1.shape)
    assert helpers.rel_diff(prd1, p_or, mask) <= PRESSURE_RTOL
    solid = c.fields.solid.astype(bool)
    assert np.all(prd1[1:-1, 1:-1, 1:-1][solid] == 0.0)


@pytest.mark.parametrize("name,make", MASKED[:2] + MASKED[3:], ids=["walls", "px", "cfg3_small"])
def test_masked_posdin_exact(solver, name, make):
    c = make()
    p_or, _, _ = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-10,
                                project_null=True)
    f = c.fields.copy()
    f.prd1[...] = p_or
    u, v, w, _ = oracle.posdin(c.grid, f, p_or)
    df = solver.upload(c.grid, f)
    try:
        solver.posdin_dev(c.grid, df)
        np.testing.assert_array_equal(df.download("u"), u)
        np.testing.assert_array_equal(df.download("v"), v)
        np.testing.assert_array_equal(df.download("w"), w)
    finally:
        df.free()


def test_masked_batch_bitwise(solver):
    cs = [m[1]() for m in MASKED] + [cases.random_case(30, 18, 12, seed=55)]
    solver.configure(precision="mixed", stop_rule="relative", tol=1e-11, verbose=False)
    dfs = [solver.upload(c.grid, c.fields) for c in cs]
    try:
        sts = solver.graddin_batch_dev([(c.grid, df) for c, df in zip(cs, dfs)])
        got = [df.download("prd1") for df in dfs]
    finally:
        for df in dfs:
            df.free()
    for i, c in enumerate(cs):
        alone, st1 = solve_dev(solver, c)
        assert sts[i].status == 0 and sts[i].iters == st1.iters
        assert np.array_equal(got[i], alone), f"problem {i}"
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)


def test_cfg3_full_size_fp64_and_mixed(solver):
    """384x192x128 stretched grid with the masked cylinder: both solvers
    reach 1e-8 on the FP64 true residual and agree within that tolerance."""
    c = cases.config3(seed=1)
    out = {}
    for precision in ("fp64", "mixed"):
        print(f"cfg3 full {precision}", flush=True)
        solver.configure(precision=precision, stop_rule="relative", tol=1e-8, verbose=False)
        prd1, st = solve_dev(solver, c)
        assert st.status == 0 and st.rel_true <= 1.05e-8, st.as_dict()
        out[precision] = prd1
    mask = helpers.face_mask(out["fp64"].shape)
    assert helpers.rel_diff(out["mixed"], out["fp64"], mask) <= 1e-5
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
