"""SuLi-IT's non-hydrostatic pressure on the GPU (suli_ps_it_graddin[_dev],
suli_ps_it_posdin[_dev]) against the C oracle, which is bit-exact with the
compiled SuLi-IT graddin at its 50x50x25 size and pinned by brute force at
other sizes (tests/test_oracle_it.py).

Tolerances: pressure relative max-difference <= 1e-8 with both sides at a
1e-12 relative residual (FP64) or the GPU at 1e-11 (mixed); under SuLi-IT's
own stop rule (|r.r| <= 1e-11 on the previous iteration's r.r) the FP64
iteration count equals the oracle's within one (the dot products are summed
in another order); posdin bit for bit from the same prd1."""
import numpy as np
import pytest

import oracle
from suli_amd import cases
from tests import helpers

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    from suli_amd.solver import PressureSolver
    s = PressureSolver(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
    yield s
    s.close()


def ref_size(seed):
    pi = np.pi
    return cases.it_case(50, 50, 25, seed=seed, dx=2.0 * pi / 50.0, dy=2.0 * pi / 50.0, dt=0.001 * pi)


def periodic_fix(g, f):
    if g.ccx0 == 0:
        f.dzx[1, :, :] = f.dzx[g.nx + 1, :, :]
        f.u[1, :, :] = f.u[g.nx + 1, :, :]
    if g.ccy0 == 0:
        f.dzy[:, 1, :] = f.dzy[:, g.ny + 1, :]
        f.v[:, 1, :] = f.v[:, g.ny + 1, :]
    return g, f


CASES = [
    ("ref_size", lambda: ref_size(31)),
    ("ragged", lambda: cases.it_case(13, 7, 5, seed=32)),
    ("periodic_x", lambda: periodic_fix(*cases.it_case(40, 24, 12, seed=33, periodic_x=True))),
    ("periodic_y", lambda: periodic_fix(*cases.it_case(24, 36, 10, seed=34, periodic_y=True))),
]


def solve_dev(solver, g, f):
    d = solver.it_upload(f)
    try:
        st = solver.it_graddin_dev(g, d)
        return solver.dev_download(d["prd1"], f.prd1.shape), st
    finally:
        solver.it_free(d)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-11)])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_it_vs_oracle(solver, name, make, precision, tol):
    g, f = make()
    print(f"it {name} {precision}", flush=True)
    p_or, st_or, rc = oracle.it_graddin(g, f, stop_rule=oracle.STOP_RELATIVE, tol=1e-12)
    assert rc == 0
    solver.configure(precision=precision, stop_rule="relative", tol=tol, verbose=False)
    p, st = solve_dev(solver, g, f)
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
    assert st.status == 0, st.as_dict()
    assert helpers.rel_diff(p, p_or, helpers.face_mask(p.shape)) <= 1e-8


def test_it_reference_stop_rule(solver):
    g, f = ref_size(35)
    f.prd1[...] = np.random.default_rng(35).uniform(-1, 1, f.prd1.shape)   # warm start
    p_or, st_or, rc = oracle.it_graddin(g, f)
    assert rc == 0
    solver.configure(precision="fp64", stop_rule="reference", verbose=False)
    f2 = f.copy()
    st = solver.it_graddin(g, f2)                  # host entry point
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
    assert st.status == 0
    assert abs(st.iters - st_or.iters) <= 1, (st.iters, st_or.iters)
    assert helpers.rel_diff(f2.prd1, p_or, helpers.face_mask(p_or.shape)) <= 1e-8


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_it_posdin_bitwise(solver, name, make):
    g, f = make()
    p, _, _ = oracle.it_graddin(g, f, stop_rule=oracle.STOP_RELATIVE, tol=1e-10)
    f.prd1[...] = p
    u, v, w, eta1, prd = oracle.it_posdin(g, f, p)
    d = solver.it_upload(f)
    try:
        solver.it_posdin_dev(g, d)
        np.testing.assert_array_equal(solver.dev_download(d["u"], u.shape), u)
        np.testing.assert_array_equal(solver.dev_download(d["v"], v.shape), v)
        np.testing.assert_array_equal(solver.dev_download(d["w"], w.shape), w)
        np.testing.assert_array_equal(solver.dev_download(d["eta1"], eta1.shape), eta1)
    finally:
        solver.it_free(d)
    # host entry point
    f2 = f.copy()
    prd2 = solver.it_posdin(g, f2)
    np.testing.assert_array_equal(f2.u, u)
    np.testing.assert_array_equal(f2.eta1, eta1)
    np.testing.assert_array_equal(prd2, prd)


def test_it_argument_errors(solver):
    from suli_amd.solver import SuliPsError
    g, f = cases.it_case(8, 6, 4, seed=36)
    g.tetah = 0.0
    with pytest.raises(SuliPsError, match="tetah"):
        solver.it_graddin(g, f.copy())


def test_it_posdin_without_eta(solver):
    """eta1 = NULL: the caller updates the surface itself (as SuLi-IT does
    before dz_it); only the velocities change."""
    from suli_amd import _lib
    import ctypes as C
    g, f = cases.it_case(16, 12, 8, seed=37)
    p, _, _ = oracle.it_graddin(g, f, stop_rule=oracle.STOP_RELATIVE, tol=1e-10)
    f.prd1[...] = p
    u, v, w, _, _ = oracle.it_posdin(g, f, p)
    f2 = f.copy()
    cg = solver._it_cgrid(g)
    solver._check(solver.lib.suli_ps_it_posdin(
        solver.ctx, C.byref(cg), g.tetah, g.gz, C.c_void_p(f2.u.ctypes.data), C.c_void_p(f2.v.ctypes.data),
        C.c_void_p(f2.w.ctypes.data), C.c_void_p(f2.dzz.ctypes.data), None, C.c_void_p(f2.prd1.ctypes.data),
        None))
    np.testing.assert_array_equal(f2.u, u)
    np.testing.assert_array_equal(f2.w, w)
    np.testing.assert_array_equal(f2.eta1, f.eta1)


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_it_posdin_layers_bitwise(solver, name, make):
    """The whole SuLi-IT posdin on the device (surface, top layer, dz_it,
    correction) against the oracle, itself bit-exact with the compiled
    SuLi-IT posdin (tests/test_oracle_it.py): bit for bit, through the
    device and the host entry points."""
    g, f = make()
    dz, eta0, prd1 = helpers.it_layer_state(g, f, 40)
    f.prd1[...] = prd1
    exp = oracle.it_posdin_dz(g, f, prd1, dz, eta0)
    names = ("u", "v", "w", "dz", "dzx", "dzy", "dzz", "eta1", "prd")
    d = solver.it_upload(f)
    p_dz, p_e0 = solver.dev_array(dz), solver.dev_array(eta0)
    try:
        solver.it_posdin_layers_dev(g, d, p_dz, p_e0)
        got = {k: solver.dev_download(d[k], getattr(f, k).shape) for k in ("u", "v", "w", "dzx", "dzy", "dzz",
                                                                            "eta1")}
        got["dz"] = solver.dev_download(p_dz, dz.shape)
        for k, e in zip(names, exp):
            if k in got:
                np.testing.assert_array_equal(got[k], e, err_msg=k)
    finally:
        solver.it_free(d)
        solver.dev_free(p_dz)
        solver.dev_free(p_e0)
    f2, dz2 = f.copy(), dz.copy(order="F")
    prd = solver.it_posdin_layers(g, f2, dz2, eta0)
    host = dict(u=f2.u, v=f2.v, w=f2.w, dz=dz2, dzx=f2.dzx, dzy=f2.dzy, dzz=f2.dzz, eta1=f2.eta1, prd=prd)
    for k, e in zip(names, exp):
        np.testing.assert_array_equal(host[k], e, err_msg=k)
