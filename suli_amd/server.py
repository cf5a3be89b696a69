"""Serving loop on the MI355X library (``suli_ps_server_*`` in include/suli_ps.h).

``PressureServer.submit(grid, fields)`` queues one graddin solve
(SuLi-IC/9_grad.f90:10-663 on the given arrays) and returns a ticket;
``wait(ticket)`` returns its statistics once ``fields.prd1`` holds the
result.  Requests are admitted by an HBM budget and solved together with the
others resident on the device; each result is bit-identical to a solo solve
with the same options.  :func:`poisson_stream` replays a Poisson arrival
process against a server and reports latency percentiles and throughput.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from typing import List, Optional, Sequence

import numpy as np

from . import _lib
from .cases import Case, Fields, Grid
from .solver import PRECISIONS, STOP_RULES, SuliPsError, _c_grid, _host_ptr


class PressureServer:
    def __init__(self, precision: str = "mixed", stop_rule: str = "relative", tol: float = 1e-8,
                 max_iter: int = 10000, hbm_budget: int = 0, max_active: int = 0,
                 graphs: bool = True, autotune: bool = False):
        self.lib = _lib.load()
        o = _lib.ServerOpts()
        o.solve.precision = PRECISIONS[precision]
        o.solve.stop_rule = STOP_RULES[stop_rule]
        o.solve.tol = tol
        o.solve.max_iter = max_iter
        o.solve.verbose = 0
        o.solve.no_graphs = 0 if graphs else 1
        o.solve.autotune = 1 if autotune else 0
        o.hbm_budget = hbm_budget
        o.max_active = max_active
        self.srv = C.c_void_p(self.lib.suli_ps_server_create(C.byref(o)))
        if not self.srv:
            raise SuliPsError("suli_ps_server_create failed (no HIP device?)")
        self._keep = {}   # ticket -> arrays that must stay alive until wait

    def _err(self, rc: int) -> SuliPsError:
        msg = self.lib.suli_ps_server_last_error(self.srv)
        return SuliPsError(f"suli_ps server error {rc}: {msg.decode() if msg else ''}")

    def submit(self, grid: Grid, fields: Fields) -> int:
        """Queue a solve; fields.prd1 (Fortran-ordered float64) receives the
        result.  The arrays must not change until wait() returns."""
        cg = _c_grid(grid)
        arrs = (fields.u, fields.v, fields.w, fields.rho, fields.obs_lsx, fields.obs_lsy,
                fields.obs_lsz, grid.dzv, grid.dzzv, fields.prd1)
        t = self.lib.suli_ps_server_submit(self.srv, C.byref(cg), *[_host_ptr(a) for a in arrs])
        if t < 0:
            raise self._err(t)
        self._keep[t] = (cg, arrs)
        return t

    def wait(self, ticket: int, timeout: float = -1.0) -> Optional[_lib.Stats]:
        """Statistics of a finished request (None on timeout)."""
        st = _lib.Stats()
        rc = self.lib.suli_ps_server_wait(self.srv, ticket, timeout, C.byref(st))
        if rc == _lib.SULI_PS_EAGAIN:
            return None
        self._keep.pop(ticket, None)
        if rc < 0:
            raise self._err(rc)
        return st

    def info(self) -> dict:
        inf = _lib.ServerInfo()
        rc = self.lib.suli_ps_server_get_info(self.srv, C.byref(inf))
        if rc < 0:
            raise self._err(rc)
        return inf.as_dict()

    def close(self) -> None:
        if self.srv:
            self.lib.suli_ps_server_destroy(self.srv)
            self.srv = C.c_void_p()
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def poisson_stream(server: PressureServer, cases: Sequence[Case], rate: float,
                   seed: int = 0) -> dict:
    """Submit ``cases`` at the arrival times of a Poisson process of ``rate``
    requests/s, wait for all, and report latency (arrival to result on the
    host) percentiles and throughput.  Each case's fields.prd1 receives its
    result."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=len(cases))
    arrivals = np.cumsum(gaps) - gaps[0]
    tickets: List[int] = [None] * len(cases)
    t_arr = [0.0] * len(cases)
    t_done = [0.0] * len(cases)
    stats = [None] * len(cases)
    start = time.perf_counter()
    err = []

    def waiter():
        try:
            for i in range(len(cases)):
                while tickets[i] is None:
                    time.sleep(1e-4)
                stats[i] = server.wait(tickets[i])
                t_done[i] = time.perf_counter()
        except Exception as exc:   # reported by the caller
            err.append(exc)

    th = threading.Thread(target=waiter)
    th.start()
    for i, c in enumerate(cases):
        now = time.perf_counter() - start
        if arrivals[i] > now:
            time.sleep(arrivals[i] - now)
        t_arr[i] = time.perf_counter()
        tickets[i] = server.submit(c.grid, c.fields)
    th.join()
    if err:
        raise err[0]
    lat = np.array([(d - a) * 1e3 for a, d in zip(t_arr, t_done)])
    span = max(t_done) - t_arr[0]
    cells = sum(c.grid.ncell for c in cases)
    return {"requests": len(cases), "rate_per_s": rate, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()),
            "solves_per_s": len(cases) / span, "mcells_per_s": cells / span / 1e6,
            "iters_mean": float(np.mean([s.iters for s in stats])),
            "not_converged": int(sum(1 for s in stats if s.status != 0)),
            "info": server.info(), "stats": stats}
