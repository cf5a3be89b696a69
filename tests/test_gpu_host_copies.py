"""The host-pointer entry points' copies move every byte.

Above a couple of MiB per call, the library stages host arrays through
pinned chunks of its own with several host threads; below, it copies
directly.  graddin then posdin on host arrays must give, bit for bit, what
the same calls give on device-resident copies of the same arrays: at a size
that copies directly, at one whose uploads just pass the threshold (fewer
chunks than worker threads), and at one whose arrays end part-way through a
chunk, with and without the stretched-grid and immersed-boundary arrays, and
with posdin uploading again or reusing graddin's copies.  posdin's
prd = prd1 on the host arrays is checked as well.  The hydrostatic step,
exp_press and SuLi-IT's solve and posdin get the same check at a size whose
copies are staged.
"""
import numpy as np
import pytest

from suli_amd import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def solver():
    from suli_amd.solver import PressureSolver
    s = PressureSolver(precision="mixed", stop_rule="relative", tol=1e-8, verbose=False)
    yield s
    s.close()


# (nx, ny, nz): about 0.2 MB per array (direct copies); 0.5 MB per array,
# 2.5 MB per graddin upload (staged, one chunk per array); 5 MB per array
# (two chunks each, the second partly filled; staged downloads too)
SIZES = [(16, 12, 10), (48, 40, 32), (100, 70, 90)]


@pytest.mark.parametrize("reuse", [False, True])
@pytest.mark.parametrize("extras", [False, True])
@pytest.mark.parametrize("size", SIZES)
def test_host_arrays_match_device_arrays(solver, size, extras, reuse):
    nx, ny, nz = size
    print(f"host copies {nx}x{ny}x{nz} extras={extras} reuse={reuse}", flush=True)
    solver.configure(precision="mixed", stop_rule="relative", tol=1e-8, verbose=False, reuse_staged=reuse)
    c = cases.random_case(nx, ny, nz, seed=3, stretched=extras, ibm=extras)
    g = c.grid

    df = solver.upload(g, c.fields)
    try:
        st_dev = solver.graddin_dev(g, df)
        solver.posdin_dev(g, df)
        dev = {k: df.download(k) for k in ("prd1", "u", "v", "w")}
    finally:
        df.free()

    f = c.fields.copy()
    st_host = solver.graddin(g, f)
    assert st_host.status == 0 and st_dev.status == 0
    assert st_host.iters == st_dev.iters
    np.testing.assert_array_equal(f.prd1, dev["prd1"])
    prd = solver.posdin(g, f)
    for k in ("u", "v", "w"):
        np.testing.assert_array_equal(getattr(f, k), dev[k])
    np.testing.assert_array_equal(prd, f.prd1)


# 100 x 70 x 40 cells: 2.3 MB per cell array, so every call below stages
# its copies through the pinned chunks
ROWS_SHAPE = (100, 70, 40)


def test_hydrostatic_step_host_matches_device(solver):
    """pressh, then posdin's hydrostatic branch, on host arrays against the
    same calls on device-resident copies, bit for bit."""
    nx, ny, nz = ROWS_SHAPE
    c = cases.random_case(nx, ny, nz, seed=5, surface=0.55)
    g = c.grid
    g.t_press = 0
    f = c.fields.copy()
    f.hs = np.asfortranarray(np.clip((f.rho - 1.2) / (1000.0 - 1.2), 0.0, 1.0))
    fd = f.copy()
    fd.hs = f.hs.copy(order="F")
    df = solver.upload(g, fd)
    try:
        solver.pressh_dev(g, df, 9.80665)
        solver.posdin_hydro_dev(g, df)
        dev = {k: df.download(k) for k in ("prd1", "u", "v", "w")}
    finally:
        df.free()
    solver.pressh(g, f, 9.80665)
    prd = solver.posdin_hydro(g, f)
    for k in ("prd1", "u", "v", "w"):
        np.testing.assert_array_equal(getattr(f, k), dev[k])
    np.testing.assert_array_equal(prd, f.prd1)


def test_exp_press_host_matches_device(solver):
    nx, ny, nz = ROWS_SHAPE
    c = cases.random_case(nx, ny, nz, seed=6)
    g = c.grid
    rng = np.random.default_rng(7)
    box = (nx + 2, ny + 2, nz + 2)
    prd1 = np.asfortranarray(rng.uniform(-50.0, 50.0, box))
    prd0 = np.asfortranarray(rng.uniform(-1e4, 1e4, box))
    rho = c.fields.rho
    shapes = ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))
    ptr = {k: solver.dev_array(a) for k, a in dict(rho=rho, prd1=prd1, prd0=prd0).items()}
    pxyz = [solver.dev_array(np.zeros(s)) for s in shapes]
    try:
        solver.exp_press_dev(g, ptr["rho"], ptr["prd1"], ptr["prd0"], *pxyz)
        dev = [solver.dev_download(ptr["prd0"], box)] + [solver.dev_download(p, s) for p, s in zip(pxyz, shapes)]
    finally:
        for p in list(ptr.values()) + pxyz:
            solver.dev_free(p)
    p0 = prd0.copy(order="F")
    host = [p0] + list(solver.exp_press(g, rho, prd1, p0))
    for a, b in zip(host, dev):
        np.testing.assert_array_equal(a, b)


def test_it_step_host_matches_device(solver):
    """SuLi-IT's solve and posdin on host arrays against device-resident
    copies, bit for bit (the same kernels on the same bytes)."""
    nx, ny, nz = ROWS_SHAPE
    g, f = cases.it_case(nx, ny, nz, seed=8)
    solver.configure(precision="mixed", stop_rule="relative", tol=1e-8, verbose=False)
    d = solver.it_upload(f)
    try:
        st_dev = solver.it_graddin_dev(g, d)
        solver.it_posdin_dev(g, d)
        dev = {k: solver.dev_download(d[k], getattr(f, k).shape) for k in ("prd1", "u", "v", "w", "eta1")}
    finally:
        solver.it_free(d)
    st_host = solver.it_graddin(g, f)
    assert st_host.status == 0 and st_host.iters == st_dev.iters
    solver.it_posdin(g, f)
    for k in ("prd1", "u", "v", "w", "eta1"):
        np.testing.assert_array_equal(getattr(f, k), dev[k])
