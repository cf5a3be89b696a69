"""Batched graddin (suli_ps_graddin_batch_dev) through the C ABI.

* Every problem of a batch of uneven sizes and boundary types gets the same
  bits as when it is solved alone (the kernels' arithmetic depends only on
  the problem's own tiling), in FP64 and mixed precision.
* Each batch result agrees with the CPU oracle within 1e-8 relative.
* Problems that stop early (zero right-hand side, iteration cap) leave the
  others undisturbed.
"""
import numpy as np
import pytest

import oracle
from suli_amd import cases
from suli_amd.solver import TILES_PIPELINED
from tests import helpers

pytestmark = pytest.mark.gpu

PRESSURE_RTOL = 1e-8


@pytest.fixture(scope="module")
def solver():
    from suli_amd.solver import PressureSolver
    s = PressureSolver(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
    yield s
    s.close()


def mixed_cases(seed=7):
    shapes = [(24, 20, 16), (70, 33, 19), (9, 7, 5), (64, 4, 8), (40, 24, 20), (33, 65, 12)]
    kinds = [dict(), dict(periodic_x=True), dict(periodic_y=True, surface=True),
             dict(ibm=True), dict(periodic_x=True, periodic_y=True), dict(stretched=True)]
    out = []
    for i, (shape, kw) in enumerate(zip(shapes, kinds)):
        c = cases.random_case(*shape, seed=seed + i, **kw)
        if kw.get("stretched"):
            cases.make_stretched_compatible(c)
        out.append(c)
    return out


def solve_alone(solver, c):
    df = solver.upload(c.grid, c.fields)
    try:
        st = solver.graddin_dev(c.grid, df)
        return df.download("prd1"), st
    finally:
        df.free()


def solve_batch(solver, cs):
    dfs = [solver.upload(c.grid, c.fields) for c in cs]
    try:
        sts = solver.graddin_batch_dev([(c.grid, df) for c, df in zip(cs, dfs)])
        return [df.download("prd1") for df in dfs], sts
    finally:
        for df in dfs:
            df.free()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-11)])
def test_batch_bitwise_equals_alone(solver, precision, tol):
    solver.configure(precision=precision, stop_rule="relative", tol=tol, verbose=False)
    cs = mixed_cases()
    print(f"batch of {len(cs)} {precision}", flush=True)
    got, sts = solve_batch(solver, cs)
    for i, c in enumerate(cs):
        alone, st1 = solve_alone(solver, c)
        assert sts[i].status == 0 and st1.status == 0, (i, sts[i].as_dict(), st1.as_dict())
        assert sts[i].iters == st1.iters, (i, sts[i].iters, st1.iters)
        assert np.array_equal(got[i], alone), f"problem {i} differs from its solo solve"


def test_batch_matches_oracle(solver):
    solver.configure(precision="mixed", stop_rule="relative", tol=1e-11, verbose=False)
    cs = mixed_cases(seed=31)
    got, sts = solve_batch(solver, cs)
    for i, c in enumerate(cs):
        p_or, _, rc = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-12,
                                     project_null=True)
        assert rc == 0
        mask = helpers.face_mask(got[i].shape)
        d = helpers.rel_diff(got[i], p_or, mask)
        assert d <= PRESSURE_RTOL, (i, d, sts[i].as_dict())


def test_batch_with_early_stoppers(solver):
    """A zero right-hand side (0 iterations) and a capped problem in the
    same batch as ordinary ones."""
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, max_iter=15, verbose=False)
    cs = mixed_cases(seed=3)[:3]
    z = cases.random_case(20, 12, 10, seed=99)
    z.fields.u[:] = 0.0
    z.fields.v[:] = 0.0
    z.fields.w[:] = 0.0
    z.fields.prd1[:] = 0.0
    batch = [cs[0], z, cs[1], cs[2]]
    got, sts = solve_batch(solver, batch)
    assert sts[1].status == 0 and sts[1].iters == 0
    assert np.all(got[1] == 0.0)
    for i in (0, 2, 3):
        assert sts[i].status == 1 and sts[i].iters == 15
        alone, st1 = solve_alone(solver, batch[i])
        assert np.array_equal(got[i], alone)
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)


def test_batch_of_one_and_empty(solver):
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
    c = mixed_cases()[0]
    got, sts = solve_batch(solver, [c])
    alone, _ = solve_alone(solver, c)
    assert np.array_equal(got[0], alone)
    assert solver.graddin_batch_dev([]) == []


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-11)])
def test_graph_replay_equals_direct_launches(solver, precision, tol):
    """Captured-graph replay and host launches run the same kernels on the
    same steps: identical bits, batch and single."""
    cs = mixed_cases(seed=11)[:4]
    solver.configure(precision=precision, stop_rule="relative", tol=tol, verbose=False, graphs=True)
    got_g, sts_g = solve_batch(solver, cs)
    solver.configure(precision=precision, stop_rule="relative", tol=tol, verbose=False, graphs=False)
    got_d, sts_d = solve_batch(solver, cs)
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)
    for i in range(len(cs)):
        assert sts_g[i].iters == sts_d[i].iters
        assert np.array_equal(got_g[i], got_d[i]), f"problem {i}"


def test_autotune_picks_a_depth_and_keeps_bit_identity(solver):
    """The tuner returns one of its candidate depths with a positive time,
    caches it, and tuned solves stay bit-identical alone and in a batch."""
    cs = mixed_cases(seed=17)[:3]
    kz1, vx1, us1 = solver.autotune(cs[1].grid, "mixed")
    kz2, vx2, us2 = solver.autotune(cs[1].grid, "mixed")
    assert (kz1, vx1, us1) == (kz2, vx2, us2)
    assert 2 <= kz1 <= max(48, cs[1].grid.nz) and us1 > 0
    assert vx1 in (1, 2, 4, 2 | TILES_PIPELINED, 4 | TILES_PIPELINED)
    assert solver.autotune(cs[1].grid, "fp64")[1] == 1
    solver.configure(precision="mixed", stop_rule="relative", tol=1e-11, verbose=False,
                     autotune=True)
    got, sts = solve_batch(solver, cs)
    for i, c in enumerate(cs):
        alone, st1 = solve_alone(solver, c)
        assert sts[i].status == 0 and sts[i].iters == st1.iters
        assert np.array_equal(got[i], alone), f"problem {i}"
        p_or, _, rc = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-12,
                                     project_null=True)
        assert helpers.rel_diff(got[i], p_or, helpers.face_mask(got[i].shape)) <= PRESSURE_RTOL
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)


@pytest.mark.parametrize("kz,vx", [(3, 2), (8, 2), (3, 4), (8, 4), (3, 2 | TILES_PIPELINED),
                                   (8, 2 | TILES_PIPELINED), (3, 4 | TILES_PIPELINED), (8, 4 | TILES_PIPELINED)])
def test_wide_lanes_vs_oracle(solver, kz, vx):
    """FP32 tiles with two or four x cells per lane (nx a multiple; walls,
    x-periodic, y-periodic, stretched, Dirichlet lid), with the step's plane
    walk issued plane by plane or pipelined: same tolerance against the
    oracle, and batch results still equal solo ones."""
    nxp = 70 if vx & ~TILES_PIPELINED == 2 else 72
    cs = [cases.random_case(24, 20, 16, seed=61), cases.random_case(nxp, 33, 19, seed=62, periodic_x=True),
          cases.random_case(64, 4, 8, seed=63, periodic_y=True), cases.random_case(40, 24, 20, seed=64,
                                                                                    stretched=True),
          cases.config1(seed=65, nx=32, ny=12, nz=10, top_dirichlet=True)]
    cases.make_stretched_compatible(cs[3])
    for c in cs:
        solver.set_tiles(c.grid, "mixed", kz, vx)
        assert solver.autotune(c.grid, "mixed")[:2] == (kz, vx)
    solver.configure(precision="mixed", stop_rule="relative", tol=1e-11, verbose=False, autotune=True)
    got, sts = solve_batch(solver, cs)
    for i, c in enumerate(cs):
        p_or, _, rc = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-12,
                                     project_null=True)
        assert rc == 0 and sts[i].status == 0
        assert helpers.rel_diff(got[i], p_or, helpers.face_mask(got[i].shape)) <= PRESSURE_RTOL, i
        alone, st1 = solve_alone(solver, c)
        assert np.array_equal(got[i], alone), f"problem {i}"
    solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False)


@pytest.mark.parametrize("graphs", [True, False], ids=["graphs", "direct"])
def test_every_lane_kind_in_one_batch(solver, graphs):
    """Five problems, each on another kind of the mixed step (one cell per
    lane; two and four, plane by plane and pipelined), in one batch: every
    step launches all five kernels (launched directly, on four streams at
    once).  Each result equals its solo solve bit for bit and agrees with
    the oracle."""
    cs = [cases.random_case(40, 20, 14, seed=71), cases.random_case(64, 18, 12, seed=72, periodic_x=True),
          cases.random_case(48, 22, 16, seed=73, surface=True), cases.random_case(72, 16, 10, seed=74),
          cases.random_case(264, 12, 9, seed=75, periodic_y=True)]
    kinds = [1, 2, 2 | TILES_PIPELINED, 4, 4 | TILES_PIPELINED]
    for c, vx in zip(cs, kinds):
        solver.set_tiles(c.grid, "mixed", 4, vx)
        assert solver.autotune(c.grid, "mixed")[:2] == (4, vx)
    solver.configure(precision="mixed", stop_rule="relative", tol=1e-11, verbose=False, autotune=True,
                     graphs=graphs)
    print(f"batch of every lane kind, graphs {graphs}", flush=True)
    try:
        got, sts = solve_batch(solver, cs)
        for i, c in enumerate(cs):
            assert sts[i].status == 0, (i, sts[i].as_dict())
            p_or, _, rc = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-12,
                                         project_null=True)
            assert rc == 0
            assert helpers.rel_diff(got[i], p_or, helpers.face_mask(got[i].shape)) <= PRESSURE_RTOL, i
            alone, _ = solve_alone(solver, c)
            assert np.array_equal(got[i], alone), f"problem {i}"
    finally:
        solver.configure(precision="fp64", stop_rule="relative", tol=1e-12, verbose=False, graphs=True)
