"""A restart record feeds the GPU pressure step: a two-phase state written
in SuLi's restart layout (tests/restart_helpers.py), read back by
suli_ps_restart_read, solved on the GPU and compared with the oracle's
solve of the original arrays (pressure relative max-difference <= 1e-8,
both at 1e-12 relative residual); the hydrostatic step from the record's
hs, bit for bit."""
import numpy as np
import pytest

import oracle
from suli_amd import cases
from suli_amd.restart import fields_from_restart
from tests import helpers
from tests.restart_helpers import write_record

pytestmark = pytest.mark.gpu


def test_solve_from_restart_record(tmp_path):
    from suli_amd.solver import PressureSolver
    c = cases.random_case(40, 24, 20, seed=71, surface=0.5)
    g = c.grid
    hs = np.asfortranarray(np.clip((c.fields.rho - 1.2) / (1000.0 - 1.2), 0.0, 1.0))
    path = str(tmp_path / "arquivorestart")
    write_record(path, dict(u=c.fields.u, v=c.fields.v, w=c.fields.w, prd1=c.fields.prd1, rho=c.fields.rho,
                            hs=hs), g.nx, g.ny, g.nz)
    f = fields_from_restart(path, g, with_hs=True)
    for name in ("u", "v", "w", "rho", "prd1"):
        np.testing.assert_array_equal(getattr(f, name), getattr(c.fields, name))
    print("restart solve 40x24x20", flush=True)
    p_or, _, rc = oracle.graddin(g, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-12, project_null=True)
    assert rc == 0
    with PressureSolver(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False) as s:
        df = s.upload(g, f)
        try:
            st = s.graddin_dev(g, df)
            assert st.status == 0
            prd1 = df.download("prd1")
        finally:
            df.free()
        assert helpers.rel_diff(prd1, p_or, helpers.face_mask(prd1.shape)) <= 1e-8
        # the hydrostatic step of the same state from the record's hs
        gh = cases.Grid(**{**g.__dict__, "t_press": 0})
        exp_p = oracle.pressh(gh, f.rho, 9.80665, f.prd1)
        u, v, w, _ = oracle.posdin_hydro(gh, f, f.hs, exp_p)
        f.prd1[...] = exp_p
        df = s.upload(gh, f)
        try:
            s.posdin_hydro_dev(gh, df)
            np.testing.assert_array_equal(df.download("u"), u)
            np.testing.assert_array_equal(df.download("w"), w)
        finally:
            df.free()
