"""The host-pointer entry points' copies move every byte.

Above a couple of MiB per call, the library stages host arrays through
pinned chunks of its own with several host threads; below, it copies
directly.  graddin then posdin on host arrays must give, bit for bit, what
the same calls give on device-resident copies of the same arrays: at a size
that copies directly, at one whose uploads just pass the threshold (fewer
chunks than worker threads), and at one whose arrays end part-way through a
chunk, with and without the stretched-grid and immersed-boundary arrays, and
with posdin uploading again or reusing graddin's copies.  posdin's
prd = prd1 on the host arrays is checked as well.
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
