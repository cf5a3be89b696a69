"""Constant-density operator (IPA, t_press = 2: every face uses rho_f1,
SuLi-IC/9_grad.f90:109-116).  Every face of a plane then carries the same
scaled coefficient (pinned on the oracle in tests/test_oracle.py), so the
iteration kernels read per-plane tables instead of the three face arrays:
the SpMV moves 2 words per cell, the FP64 step 9 instead of 12, the mixed
step 8 instead of 11.

The tables hold exactly the values of the arrays, so a solve through them
must equal, bit for bit and iteration for iteration, the same solve through
the face arrays; the arrays are forced by an all-fluid solid mask, which
changes nothing else.  Against the oracle: pressure relative
max-difference <= 1e-8."""
import numpy as np
import pytest

import oracle
from suli_amd import cases
from suli_amd.solver import TILES_PIPELINED
from tests import helpers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    from suli_amd.solver import PressureSolver
    s = PressureSolver(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
    yield s
    s.close()


def ipa(c):
    c.grid.t_press = 2
    return c


def stretched_ipa():
    c = cases.random_case(36, 20, 24, seed=84, stretched=True)
    cases.make_stretched_compatible(c)
    return ipa(c)


CASES = [
    ("walls", lambda: ipa(cases.random_case(40, 24, 20, seed=81))),
    ("periodic_x", lambda: ipa(cases.random_case(48, 20, 16, seed=82, periodic_x=True))),
    ("periodic_y", lambda: ipa(cases.random_case(32, 24, 18, seed=83, periodic_y=True))),
    ("stretched", stretched_ipa),
    ("ibm_rhs", lambda: ipa(cases.random_case(40, 20, 16, seed=85, ibm=True))),
    ("dirichlet_lid", lambda: ipa(cases.config1(seed=86, nx=32, ny=24, nz=16, top_dirichlet=True))),
    # two 64-cell tiles along x (the second partial), a partial row of tiles along y
    ("ragged_tiles", lambda: ipa(cases.random_case(70, 33, 19, seed=88, periodic_x=True))),
]


def make(name):
    return dict(CASES)[name]()


def solve(solver, c, solid=None):
    f = c.fields.copy()
    f.solid = solid
    df = solver.upload(c.grid, f)
    try:
        st = solver.graddin_dev(c.grid, df)
        return df.download("prd1"), st
    finally:
        df.free()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-11)])
@pytest.mark.parametrize("name", [n for n, _ in CASES])
def test_tables_equal_face_arrays(solver, name, precision, tol):
    c = make(name)
    print(f"ipa {name} {precision}", flush=True)
    solver.configure(precision=precision, stop_rule="relative", tol=tol, verbose=False)
    p_tab, st_tab = solve(solver, c)
    p_arr, st_arr = solve(solver, c, solid=np.zeros((c.grid.nx, c.grid.ny, c.grid.nz), np.uint8, order="F"))
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
    assert st_tab.status == 0 and st_arr.status == 0
    assert st_tab.iters == st_arr.iters
    np.testing.assert_array_equal(p_tab, p_arr)
    p_or, _, rc = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-12, project_null=True)
    assert rc == 0
    assert helpers.rel_diff(p_tab, p_or, helpers.face_mask(p_tab.shape)) <= 1e-8


def test_tables_with_wide_lanes_and_spmv(solver):
    """Two and four cells per lane read the tables too; the timing SpMV
    runs on them."""
    c = make("walls")
    out = []
    for vx in (2, 4, 2 | TILES_PIPELINED, 4 | TILES_PIPELINED):
        solver.set_tiles(c.grid, "mixed", 6, vx)
        solver.configure(precision="mixed", stop_rule="relative", tol=1e-11, verbose=False, autotune=True)
        p, st = solve(solver, c)
        assert st.status == 0
        out.append(p)
        assert solver.time_spmv("mixed", reps=3) > 0
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
    _, st = solve(solver, c)   # the timing hook runs on the last solve's operator
    assert st.status == 0 and solver.time_spmv("fp64", reps=3) > 0
    p_or, _, _ = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-12, project_null=True)
    for p in out:
        assert helpers.rel_diff(p, p_or, helpers.face_mask(p.shape)) <= 1e-8


@pytest.mark.parametrize("name", [n for n, _ in CASES if n != "dirichlet_lid"])
def test_exp_press_vs_oracle(solver, name):
    """exp_press (SuLi-IC/11_diversos.f90:455-495) on the device against the
    oracle, itself bit-exact with the compiled reference exp_press
    (tests/test_oracle_exp_press.py): every operation is IEEE basic
    arithmetic in the reference's order, so bit for bit; host and device
    entry points."""
    c = make(name)
    g = c.grid
    rng = np.random.default_rng(90)
    box = (g.nx + 2, g.ny + 2, g.nz + 2)
    prd1 = np.asfortranarray(rng.uniform(-50.0, 50.0, box))
    prd0 = np.asfortranarray(rng.uniform(-1e4, 1e4, box))
    rho = c.fields.rho
    e0, epx, epy, epz = oracle.exp_press(g, rho, prd1, prd0)
    p0 = prd0.copy(order="F")
    px, py, pz = solver.exp_press(g, rho, prd1, p0)
    for a, b in ((p0, e0), (px, epx), (py, epy), (pz, epz)):
        np.testing.assert_array_equal(a, b)
    ptr = {k: solver.dev_array(a) for k, a in dict(rho=rho, prd1=prd1, prd0=prd0, px=np.zeros_like(px),
                                                    py=np.zeros_like(py), pz=np.zeros_like(pz)).items()}
    if g.stretched:
        ptr["dz"], ptr["dzz"] = solver.dev_array(g.dzv), solver.dev_array(g.dzzv)
    try:
        solver.exp_press_dev(g, ptr["rho"], ptr["prd1"], ptr["prd0"], ptr["px"], ptr["py"], ptr["pz"],
                             ptr.get("dz"), ptr.get("dzz"))
        np.testing.assert_array_equal(solver.dev_download(ptr["prd0"], box), e0)
        np.testing.assert_array_equal(solver.dev_download(ptr["px"], px.shape), epx)
        np.testing.assert_array_equal(solver.dev_download(ptr["pz"], pz.shape), epz)
    finally:
        for p in ptr.values():
            solver.dev_free(p)
