"""Serving loop (suli_ps_server_*) through the C ABI.

* Requests of uneven sizes and boundary types, submitted while others are
  iterating, each give the same bits as a solo solve with the same options
  (FP64 and mixed precision).
* A budget that holds only one or two requests at a time still serves them
  all, never holding more device memory than allowed.
* A request larger than the budget is refused at submit.
"""
import time

import numpy as np
import pytest

from suli_amd import cases

pytestmark = pytest.mark.gpu


def request_cases(seed=5, count=8):
    shapes = [(24, 20, 16), (70, 33, 19), (9, 7, 5), (64, 4, 8), (40, 24, 20), (33, 65, 12),
              (48, 48, 24), (17, 29, 31)]
    kinds = [dict(), dict(periodic_x=True), dict(periodic_y=True, surface=True), dict(ibm=True),
             dict(periodic_x=True, periodic_y=True), dict(stretched=True), dict(surface=True),
             dict(ibm=True, periodic_x=True)]
    out = []
    for i in range(count):
        c = cases.random_case(*shapes[i % len(shapes)], seed=seed + i, **kinds[i % len(kinds)])
        if kinds[i % len(kinds)].get("stretched"):
            cases.make_stretched_compatible(c)
        out.append(c)
    return out


def solo(precision, tol, c):
    from suli_amd.solver import PressureSolver
    with PressureSolver(precision=precision, stop_rule="relative", tol=tol, verbose=False) as s:
        f = c.fields.copy()
        st = s.graddin(c.grid, f)
        return f.prd1, st


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-10)])
def test_served_results_equal_solo(precision, tol):
    from suli_amd.server import PressureServer
    cs = request_cases()
    expect = [solo(precision, tol, c) for c in cs]
    work = [cases.Case(c.name, c.grid, c.fields.copy(), c.seed) for c in cs]
    with PressureServer(precision=precision, stop_rule="relative", tol=tol) as srv:
        tickets = []
        for i, c in enumerate(work):
            print(f"submit {i} {c.grid.nx}x{c.grid.ny}x{c.grid.nz}", flush=True)
            tickets.append(srv.submit(c.grid, c.fields))
            time.sleep(0.002 * (i % 3))        # arrivals while others iterate
        sts = [srv.wait(t) for t in tickets]
        info = srv.info()
    assert info["completed"] == len(cs)
    for i, (c, st) in enumerate(zip(work, sts)):
        p_solo, st_solo = expect[i]
        assert st.status == 0 and st.iters == st_solo.iters, (i, st.as_dict(), st_solo.as_dict())
        assert np.array_equal(c.fields.prd1, p_solo), f"request {i} differs from its solo solve"


def test_small_budget_serves_everything():
    from suli_amd.server import PressureServer
    cs = request_cases(seed=21, count=6)
    expect = [solo("mixed", 1e-10, c) for c in cs]
    work = [cases.Case(c.name, c.grid, c.fields.copy(), c.seed) for c in cs]
    # room for the largest request and a little more
    big = max(c.grid.ncell for c in cs)
    budget = 60 * 8 * big
    with PressureServer(precision="mixed", tol=1e-10, hbm_budget=budget) as srv:
        tickets = [srv.submit(c.grid, c.fields) for c in work]
        sts = [srv.wait(t) for t in tickets]
        info = srv.info()
    assert info["peak_resident_bytes"] <= budget
    for i, c in enumerate(work):
        assert sts[i].status == 0
        assert np.array_equal(c.fields.prd1, expect[i][0]), f"request {i}"


def test_request_larger_than_budget_is_refused():
    from suli_amd.server import PressureServer
    from suli_amd.solver import SuliPsError
    c = request_cases(count=1)[0]
    with PressureServer(precision="fp64", tol=1e-10, hbm_budget=1 << 16) as srv:
        with pytest.raises(SuliPsError, match="budget"):
            srv.submit(c.grid, c.fields)


def test_poisson_stream_reports_latency():
    from suli_amd.server import PressureServer, poisson_stream
    cs = request_cases(seed=40, count=8)
    with PressureServer(precision="mixed", tol=1e-8) as srv:
        rep = poisson_stream(srv, cs, rate=200.0, seed=3)
    assert rep["requests"] == 8 and rep["not_converged"] == 0
    assert 0 < rep["p50_ms"] <= rep["p99_ms"]
    for c, st in zip(cs, rep["stats"]):
        assert st.rel_true <= 1.5e-8
