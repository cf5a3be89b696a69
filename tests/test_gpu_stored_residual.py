"""The mixed solver's two ways to the residual of the previous iteration
(suli_ps_opts.stored_residual): recovered from the directions,
r_{n-2} = p_{n-1} - beta_{n-1} p_{n-2} (the default, which stores no
residual), or stored by every FP32 step and read back by the next.  Both
run CG in FP32 between the same FP64 residual replacements; they differ in
FP32 rounding only, and the recovered form's error scales with |p| rather
than |r|.

On the hard cases (the stretched masked cylinder with rho 1000/1.2, an
IBM-masked right-hand side, the IPA tables at rho 1000, a lid-driven
Dirichlet case) and on rows of two segments, with either walk of the step:
both modes converge to the FP64 true-residual tolerance, match the oracle
solved to 1e-12 within 1e-8 (pressure relative max-difference), and take
iteration counts within 10 % of each other.  Measured: equal counts or 1.5 %
apart on the four hard cases, 5.3 % on the long thin random case with rows
of two segments (264x14x12: 1824 iterations recovered, 1728 stored).  Their
results are not bitwise equal, which shows the stored path ran."""
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
    s = PressureSolver(precision="mixed", stop_rule="relative", tol=1e-11, verbose=False)
    yield s
    s.close()


def ipa(c):
    c.grid.t_press = 2
    c.grid.rho_f1 = 1000.0
    return c


CASES = [
    ("cylinder", lambda: cases.config3(seed=61, nx=64, ny=32, nz=24)),
    ("ibm_rhs", lambda: cases.random_case(48, 32, 24, seed=62, ibm=True)),
    ("ipa", lambda: ipa(cases.random_case(48, 32, 24, seed=63))),
    ("dirichlet_lid", lambda: cases.config1(seed=64, nx=40, ny=32, nz=20, top_dirichlet=True)),
    ("rows_of_two_segments", lambda: cases.random_case(264, 14, 12, seed=65)),
]


@pytest.mark.parametrize("walk", [0, TILES_PIPELINED], ids=["planewise", "pipelined"])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_stored_and_recovered_residual(solver, name, make, walk):
    """Four cells per lane with either walk of the step: rows of one
    segment, and of two (264 cells: the cells beyond a segment's ends are
    loaded)."""
    c = make()
    print(f"stored residual {name} walk {walk}", flush=True)
    solver.set_tiles(c.grid, "mixed", 6, 4 | walk)
    p_or, _, rc = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-12,
                                 project_null=True)
    assert rc == 0
    mask = helpers.face_mask(p_or.shape)
    out = {}
    for stored in (False, True):
        solver.configure(precision="mixed", stop_rule="relative", tol=1e-11, verbose=False,
                         stored_residual=stored, autotune=True)
        df = solver.upload(c.grid, c.fields)
        try:
            st = solver.graddin_dev(c.grid, df)
            out[stored] = (df.download("prd1"), st)
        finally:
            df.free()
        p, st = out[stored]
        assert st.status == 0 and st.rel_true <= 1e-11, (stored, st.as_dict())
        assert helpers.rel_diff(p, p_or, mask) <= 1e-8, stored
    (p0, s0), (p1, s1) = out[False], out[True]
    print(f"  iterations recovered {s0.iters} stored {s1.iters}")
    assert abs(s1.iters - s0.iters) <= 0.10 * s0.iters
    assert np.any(p0 != p1)
    solver.configure(precision="mixed", stop_rule="relative", tol=1e-11, verbose=False)
