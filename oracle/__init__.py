"""ctypes bindings for the CPU checker.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by
bench.py's cpu_baseline leg, never by the product package (suli_amd).

* ``liboracle.so``: the C restatement of SuLi's graddin/posdin
  (oracle/suli_oracle.c, following SuLi-IC/9_grad.f90:10-663 and
  SuLi-IC/10_pos.f90:1-99).
* ``_ref/libsuli_ref_ic.so`` / ``_ref/libsuli_ref_icdz.so``: the reference's
  own Fortran graddin/posdin compiled with amdflang (oracle/Makefile), called
  through oracle/ref/ref_bridge.F90.  Present only where they were built.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
STOP_REFERENCE = 0
STOP_RELATIVE = 1

_dp = C.POINTER(C.c_double)


class OracleGrid(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
        ("dzv", _dp), ("dzzv", _dp),
        ("dt", C.c_double),
        ("ccx0", C.c_int), ("ccy0", C.c_int),
        ("t_press", C.c_int), ("rho_f1", C.c_double),
        ("ibm_mask", C.c_int),
        ("obsx", _dp), ("obsy", _dp), ("obsz", _dp),
        ("omp_t", C.c_int), ("stop_rule", C.c_int), ("tol", C.c_double),
        ("max_iter", C.c_int), ("threads", C.c_int), ("project_null", C.c_int),
        ("solid", C.c_void_p), ("top_dirichlet", C.c_int),
    ]


class OracleStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("rr0", C.c_double),
                ("bb", C.c_double), ("rr", C.c_double), ("thresh", C.c_double)]


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(_dp)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (run `make -C oracle`)")
        _lib = C.CDLL(path)
        _lib.oracle_graddin.restype = C.c_int
        _lib.oracle_graddin.argtypes = [C.POINTER(OracleGrid), _dp, _dp, _dp, _dp, _dp,
                                        C.POINTER(OracleStats)]
        _lib.oracle_posdin.restype = None
        _lib.oracle_posdin.argtypes = [C.POINTER(OracleGrid), _dp, _dp, _dp, _dp, _dp, _dp]
        _lib.oracle_interp_faces.argtypes = [C.POINTER(OracleGrid), _dp, _dp, _dp, _dp]
        _lib.oracle_scaled_operator.argtypes = [C.POINTER(OracleGrid)] + [_dp] * 10
        _lib.oracle_heaviside_rho.argtypes = [C.c_int, _dp, C.c_double, C.c_double,
                                              C.c_double, _dp]
        _lib.oracle_it_graddin.restype = C.c_int
        _lib.oracle_it_graddin.argtypes = [C.POINTER(OracleGrid), C.c_double, C.c_double] + [_dp] * 9 + \
            [C.POINTER(OracleStats)]
        _lib.oracle_it_posdin.restype = None
        _lib.oracle_it_posdin.argtypes = [C.POINTER(OracleGrid), C.c_double, C.c_double] + [_dp] * 7
        _lib.oracle_heaviside.restype = None
        _lib.oracle_heaviside.argtypes = [C.c_int, _dp] + [C.c_double] * 5 + [_dp] * 3
        _lib.oracle_pressh.restype = None
        _lib.oracle_pressh.argtypes = [C.POINTER(OracleGrid), _dp, C.c_double, _dp]
        _lib.oracle_posdin_hydro.restype = None
        _lib.oracle_posdin_hydro.argtypes = [C.POINTER(OracleGrid), _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        _lib.oracle_exp_press.restype = None
        _lib.oracle_exp_press.argtypes = [C.POINTER(OracleGrid)] + [_dp] * 6
        _lib.oracle_ls_gradient.restype = None
        _lib.oracle_ls_gradient.argtypes = [C.POINTER(OracleGrid)] + [_dp] * 4
        _lib.oracle_heaviside_tens.restype = None
        _lib.oracle_heaviside_tens.argtypes = [C.POINTER(OracleGrid), _dp] + [C.c_double] * 5 + [_dp] * 6
        _lib.oracle_it_posdin_dz.restype = None
        _lib.oracle_it_posdin_dz.argtypes = [C.POINTER(OracleGrid), C.c_double, C.c_double] + [_dp] * 11
    return _lib


def make_grid(g, fields=None, *, stop_rule=STOP_REFERENCE, tol=1e-8, max_iter=10000,
              omp_t=0, threads=1, project_null=False) -> OracleGrid:
    og = OracleGrid()
    og.nx, og.ny, og.nz = g.nx, g.ny, g.nz
    og.dx, og.dy, og.dz = g.dx, g.dy, g.dz
    og.dzv = _ptr(g.dzv)
    og.dzzv = _ptr(g.dzzv)
    og.dt = g.dt
    og.ccx0, og.ccy0 = g.ccx0, g.ccy0
    og.t_press, og.rho_f1 = g.t_press, g.rho_f1
    og.ibm_mask = 1 if g.ibm_mask else 0
    if g.ibm_mask:
        og.obsx, og.obsy, og.obsz = _ptr(fields.obs_lsx), _ptr(fields.obs_lsy), _ptr(fields.obs_lsz)
    og.omp_t, og.stop_rule, og.tol = omp_t, stop_rule, tol
    og.max_iter, og.threads = max_iter, threads
    og.project_null = 1 if project_null else 0
    og.top_dirichlet = 1 if getattr(g, "top_bc", 0) == 1 else 0
    solid = getattr(fields, "solid", None) if fields is not None else None
    if solid is not None:
        assert solid.dtype == np.uint8 and solid.flags.f_contiguous
        og.solid = solid.ctypes.data
    return og


def graddin(case_grid, fields, **kw):
    """Run the C restatement of graddin. Returns (prd1, stats, status)."""
    prd1 = fields.prd1.copy(order="F")
    og = make_grid(case_grid, fields, **kw)
    st = OracleStats()
    rc = lib().oracle_graddin(C.byref(og), _ptr(fields.u), _ptr(fields.v), _ptr(fields.w),
                              _ptr(fields.rho), _ptr(prd1), C.byref(st))
    if rc < 0:
        raise ValueError(f"oracle_graddin failed ({rc})")
    return prd1, st, rc


def posdin(case_grid, fields, prd1):
    """Run the C restatement of posdin. Returns (u, v, w, prd)."""
    u, v, w = (fields.u.copy(order="F"), fields.v.copy(order="F"), fields.w.copy(order="F"))
    prd = np.zeros_like(prd1, order="F")
    og = make_grid(case_grid, fields)
    lib().oracle_posdin(C.byref(og), _ptr(u), _ptr(v), _ptr(w), _ptr(fields.rho),
                        _ptr(np.asfortranarray(prd1)), _ptr(prd))
    return u, v, w, prd


def interp_faces(case_grid, rho):
    g = case_grid
    rx = np.zeros((g.nx + 1, g.ny, g.nz), order="F")
    ry = np.zeros((g.nx, g.ny + 1, g.nz), order="F")
    rz = np.zeros((g.nx, g.ny, g.nz + 1), order="F")
    og = make_grid(g)
    lib().oracle_interp_faces(C.byref(og), _ptr(rho), _ptr(rx), _ptr(ry), _ptr(rz))
    return rx, ry, rz


def scaled_operator(case_grid, fields):
    """(shx, shy, shz, d, q, b): Jacobi-scaled face coefficients, diagonal
    with ghosts, q = -div u and b = q/sqrt(d), as graddin builds them."""
    g = case_grid
    shx = np.zeros((g.nx + 1, g.ny, g.nz), order="F")
    shy = np.zeros((g.nx, g.ny + 1, g.nz), order="F")
    shz = np.zeros((g.nx, g.ny, g.nz + 1), order="F")
    d = np.zeros((g.nx + 2, g.ny + 2, g.nz + 2), order="F")
    q = np.zeros((g.nx, g.ny, g.nz), order="F")
    b = np.zeros((g.nx, g.ny, g.nz), order="F")
    og = make_grid(g, fields)
    lib().oracle_scaled_operator(C.byref(og), _ptr(fields.u), _ptr(fields.v), _ptr(fields.w),
                                 _ptr(fields.rho), _ptr(shx), _ptr(shy), _ptr(shz), _ptr(d),
                                 _ptr(q), _ptr(b))
    return shx, shy, shz, d, q, b


def pressh(case_grid, rho, gz, prd1):
    """C restatement of pressh (hydrostatic pressure): returns a copy of
    prd1 with its interior written."""
    out = np.asfortranarray(prd1).copy(order="F")
    og = make_grid(case_grid)
    lib().oracle_pressh(C.byref(og), _ptr(np.asfortranarray(rho, dtype=np.float64)), gz, _ptr(out))
    return out


def posdin_hydro(case_grid, fields, hs, prd1):
    """C restatement of posdin's hydrostatic branch. Returns (u, v, w, prd)."""
    u, v, w = (fields.u.copy(order="F"), fields.v.copy(order="F"), fields.w.copy(order="F"))
    prd = np.zeros_like(prd1, order="F")
    og = make_grid(case_grid)
    lib().oracle_posdin_hydro(C.byref(og), _ptr(u), _ptr(v), _ptr(w), _ptr(fields.rho),
                              _ptr(np.asfortranarray(hs, dtype=np.float64)),
                              _ptr(np.asfortranarray(prd1)), _ptr(prd))
    return u, v, w, prd


def _it_grid(g, stop_rule, tol, max_iter):
    og = OracleGrid()
    og.nx, og.ny, og.nz = g.nx, g.ny, g.nz
    og.dx, og.dy, og.dz, og.dt = g.dx, g.dy, 0.0, g.dt
    og.ccx0, og.ccy0 = g.ccx0, g.ccy0
    og.stop_rule, og.tol, og.max_iter = stop_rule, tol, max_iter
    return og


def it_graddin(g, f, *, stop_rule=STOP_REFERENCE, tol=1e-8, max_iter=10000):
    """C restatement of SuLi-IT's graddin (g: cases.ITGrid, f: cases.ITFields).
    Returns (prd1, stats, rc)."""
    og = _it_grid(g, stop_rule, tol, max_iter)
    prd1 = f.prd1.copy(order="F")
    st = OracleStats()
    rc = lib().oracle_it_graddin(C.byref(og), g.tetah, g.gz, _ptr(f.u), _ptr(f.v), _ptr(f.w), _ptr(f.dzx),
                                 _ptr(f.dzy), _ptr(f.dzz), _ptr(f.eta1), _ptr(f.deltah), _ptr(prd1), C.byref(st))
    return prd1, st, rc


def it_posdin(g, f, prd1):
    """C restatement of SuLi-IT's posdin without dz_it. Returns (u, v, w,
    eta1, prd)."""
    og = _it_grid(g, STOP_REFERENCE, 0.0, 0)
    u, v, w, eta1 = (f.u.copy(order="F"), f.v.copy(order="F"), f.w.copy(order="F"), f.eta1.copy(order="F"))
    prd = np.zeros_like(prd1, order="F")
    lib().oracle_it_posdin(C.byref(og), g.tetah, g.gz, _ptr(u), _ptr(v), _ptr(w), _ptr(f.dzz), _ptr(eta1),
                           _ptr(np.asfortranarray(prd1)), _ptr(prd))
    return u, v, w, eta1, prd


def it_posdin_dz(g, f, prd1, dz, eta0):
    """C restatement of SuLi-IT's posdin with its layer update (eta1 and
    dz(:,:,nz) from prd1, then dz_it). Returns (u, v, w, dz, dzx, dzy, dzz,
    eta1, prd)."""
    og = _it_grid(g, STOP_REFERENCE, 0.0, 0)
    u, v, w, eta1 = (f.u.copy(order="F"), f.v.copy(order="F"), f.w.copy(order="F"), f.eta1.copy(order="F"))
    dz, dzx, dzy, dzz = (np.asfortranarray(dz).copy(order="F"), f.dzx.copy(order="F"), f.dzy.copy(order="F"),
                         f.dzz.copy(order="F"))
    prd = np.zeros_like(prd1, order="F")
    lib().oracle_it_posdin_dz(C.byref(og), g.tetah, g.gz, _ptr(u), _ptr(v), _ptr(w), _ptr(dz), _ptr(dzx),
                              _ptr(dzy), _ptr(dzz), _ptr(np.asfortranarray(eta0)), _ptr(eta1),
                              _ptr(np.asfortranarray(prd1)), _ptr(prd))
    return u, v, w, dz, dzx, dzy, dzz, eta1, prd


def _unmasked(g):
    """The grid without its IBM mask, for the functions that never read it."""
    import dataclasses
    return dataclasses.replace(g, obst_t=0, ibm_t=0)


def exp_press(case_grid, rho, prd1, prd0):
    """C restatement of exp_press (IPA). Returns (prd0 after the increment,
    px, py, pz)."""
    g = case_grid
    out0 = np.asfortranarray(prd0).copy(order="F")
    px = np.zeros((g.nx + 1, g.ny, g.nz), order="F")
    py = np.zeros((g.nx, g.ny + 1, g.nz), order="F")
    pz = np.zeros((g.nx, g.ny, g.nz + 1), order="F")
    og = make_grid(_unmasked(g))
    lib().oracle_exp_press(C.byref(og), _ptr(np.asfortranarray(rho, dtype=np.float64)),
                           _ptr(np.asfortranarray(prd1)), _ptr(out0), _ptr(px), _ptr(py), _ptr(pz))
    return out0, px, py, pz


def ls_gradient(case_grid, ls):
    """The level set's WENO5 derivatives as heaviside's t_tens = 1 branch
    picks them. Returns (gx, gy, gz), each shaped like ls."""
    ls = np.asfortranarray(ls, dtype=np.float64)
    gx, gy, gz = (np.empty_like(ls, order="F") for _ in range(3))
    og = make_grid(_unmasked(case_grid))
    lib().oracle_ls_gradient(C.byref(og), _ptr(ls), _ptr(gx), _ptr(gy), _ptr(gz))
    return gx, gy, gz


def heaviside_tens(case_grid, ls, eps, rho1, rho2, mu1, mu2):
    """C restatement of heaviside with t_tens = 1. Returns (hs, rho, mu,
    hsx, hsy, hsz)."""
    ls = np.asfortranarray(ls, dtype=np.float64)
    out = [np.empty_like(ls, order="F") for _ in range(6)]
    og = make_grid(_unmasked(case_grid))
    lib().oracle_heaviside_tens(C.byref(og), _ptr(ls), eps, rho1, rho2, mu1, mu2, *[_ptr(a) for a in out])
    return tuple(out)


def heaviside(ls, eps, rho1, rho2, mu1, mu2):
    """C restatement of heaviside (t_tens = 0). Returns (hs, rho, mu), each
    shaped like ls (Fortran order)."""
    ls = np.asfortranarray(ls, dtype=np.float64)
    hs, rho, mu = (np.empty_like(ls, order="F") for _ in range(3))
    lib().oracle_heaviside(ls.size, _ptr(ls), eps, rho1, rho2, mu1, mu2, _ptr(hs), _ptr(rho), _ptr(mu))
    return hs, rho, mu


def heaviside_rho(ls, eps, rho1, rho2):
    ls = np.ascontiguousarray(ls, dtype=np.float64).ravel(order="K")
    out = np.empty_like(ls)
    lib().oracle_heaviside_rho(ls.size, _ptr(ls), eps, rho1, rho2, _ptr(out))
    return out


# --------------------------------------------------------------------------
# The reference itself (amdflang build), when present.

_ref_libs = {}


def ref_available(variant: str = "ic") -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libsuli_ref_{variant}.so"))


def ref_lib(variant: str = "ic"):
    if variant not in _ref_libs:
        path = os.path.join(HERE, "_ref", f"libsuli_ref_{variant}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (needs the reference sources)")
        L = C.CDLL(path)
        if variant == "it":
            L.suli_ref_it_graddin.argtypes = [_dp] * 9
            L.suli_ref_it_posdin.argtypes = [_dp] * 11
            _ref_libs[variant] = L
            return L
        L.suli_ref_setup.restype = None
        L.suli_ref_setup.argtypes = ([C.c_int] * 3 + [C.c_double] * 4 + [C.c_int] * 6 +
                                     [C.c_double, _dp, _dp])
        L.suli_ref_graddin.argtypes = [_dp] * 4 + [C.c_int] + [_dp] * 4
        L.suli_ref_posdin.argtypes = [_dp] * 6
        L.suli_ref_pressh.argtypes = [_dp, C.c_double, _dp]
        L.suli_ref_posdin_hydro.argtypes = [_dp] * 7
        L.suli_ref_exp_press.argtypes = [_dp] * 6
        if variant == "ic":
            L.suli_ref_heaviside.argtypes = [_dp] + [C.c_double] * 6 + [_dp] * 3
            L.suli_ref_heaviside_tens.argtypes = [_dp] + [C.c_double] * 6 + [_dp] * 6
        _ref_libs[variant] = L
    return _ref_libs[variant]


def _big_stack_call(nbytes, fn, *args):
    """The reference's graddin keeps ~15 doubles per cell in automatic
    arrays (SuLi-IC/9_grad.f90:25-39), which amdflang places on the stack:
    run it on a thread whose stack holds them."""
    import threading
    out = {}

    def run():
        try:
            out["r"] = fn(*args)
        except BaseException as e:  # pragma: no cover - re-raised below
            out["e"] = e

    old = threading.stack_size()
    threading.stack_size(max(int(nbytes), 64 << 20))
    try:
        t = threading.Thread(target=run)
        t.start()
        t.join()
    finally:
        threading.stack_size(old)
    if "e" in out:
        raise out["e"]
    return out.get("r")


def _stack_bytes(g):
    nbox = (g.nx + 2) * (g.ny + 2) * (g.nz + 2)
    return 24 * 8 * nbox + (256 << 20)


def _ref_setup(L, g, omp_t):
    nz2 = g.nz + 2
    dz = g.dzv if g.dzv is not None else np.full(nz2, g.dz)
    dzz = g.dzzv if g.dzzv is not None else np.full(nz2, g.dz)
    L.suli_ref_setup(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.dt, g.ccx0, g.ccy0, g.t_press,
                     g.obst_t, g.ibm_t, omp_t, g.rho_f1, _ptr(np.ascontiguousarray(dz)),
                     _ptr(np.ascontiguousarray(dzz)))


def ref_graddin(case_grid, fields, omp_t=0):
    """Run the reference's compiled graddin. Returns prd1 (full halo box)."""
    g = case_grid
    L = ref_lib("icdz" if g.stretched else "ic")
    _ref_setup(L, g, omp_t)
    prd1 = fields.prd1.copy(order="F")
    have_obs = 1 if fields.obs_lsx is not None else 0
    dummy = np.zeros(1)
    _big_stack_call(_stack_bytes(g), L.suli_ref_graddin, _ptr(fields.u), _ptr(fields.v),
                    _ptr(fields.w), _ptr(fields.rho), have_obs,
                    _ptr(fields.obs_lsx if have_obs else dummy),
                    _ptr(fields.obs_lsy if have_obs else dummy),
                    _ptr(fields.obs_lsz if have_obs else dummy), _ptr(prd1))
    return prd1


def ref_graddin_timed(case_grid, fields, reps, omp_t=0):
    """The reference's compiled graddin run reps times on one thread (the
    same big-stack thread each time, so that the OpenMP runtime sees one
    root thread whose affinity nothing pinned before it started).  Returns
    (prd1 of the last run, wall seconds of each run)."""
    import time
    g = case_grid
    L = ref_lib("icdz" if g.stretched else "ic")
    _ref_setup(L, g, omp_t)
    have_obs = 1 if fields.obs_lsx is not None else 0
    dummy = np.zeros(1)
    prd1 = fields.prd1.copy(order="F")

    def runs():
        times = []
        for _ in range(reps):
            prd1[...] = fields.prd1
            t = time.perf_counter()
            L.suli_ref_graddin(_ptr(fields.u), _ptr(fields.v), _ptr(fields.w), _ptr(fields.rho), have_obs,
                               _ptr(fields.obs_lsx if have_obs else dummy),
                               _ptr(fields.obs_lsy if have_obs else dummy),
                               _ptr(fields.obs_lsz if have_obs else dummy), _ptr(prd1))
            times.append(time.perf_counter() - t)
        return times

    return prd1, _big_stack_call(_stack_bytes(g), runs)


def ref_posdin(case_grid, fields, prd1):
    g = case_grid
    L = ref_lib("icdz" if g.stretched else "ic")
    _ref_setup(L, g, 0)
    u, v, w = (fields.u.copy(order="F"), fields.v.copy(order="F"), fields.w.copy(order="F"))
    prd = np.zeros_like(prd1, order="F")
    _big_stack_call(_stack_bytes(g), L.suli_ref_posdin, _ptr(u), _ptr(v), _ptr(w),
                    _ptr(fields.rho), _ptr(np.asfortranarray(prd1)), _ptr(prd))
    return u, v, w, prd


def ref_pressh(case_grid, rho, gz, prd1):
    g = case_grid
    L = ref_lib("icdz" if g.stretched else "ic")
    _ref_setup(L, g, 0)
    out = np.asfortranarray(prd1).copy(order="F")
    _big_stack_call(_stack_bytes(g), L.suli_ref_pressh, _ptr(np.asfortranarray(rho, dtype=np.float64)),
                    gz, _ptr(out))
    return out


def ref_it_available() -> bool:
    return ref_available("it")


def ref_it_info():
    """(nx, ny, nz, dx, dy, dt, tetah, gz) that SuLi-IT's modules fix at
    compile time."""
    L = ref_lib("it")
    dims = (C.c_int * 3)()
    par = (C.c_double * 5)()
    L.suli_ref_it_info(dims, par)
    return tuple(dims) + tuple(par)


def ref_it_graddin(f):
    """The reference's SuLi-IT graddin on fields of its compiled size.
    Returns prd1."""
    L = ref_lib("it")
    prd1 = f.prd1.copy(order="F")
    L.suli_ref_it_graddin(_ptr(f.u), _ptr(f.v), _ptr(f.w), _ptr(f.dzx), _ptr(f.dzy), _ptr(f.dzz),
                          _ptr(f.eta1), _ptr(f.deltah), _ptr(prd1))
    return prd1


def ref_heaviside(case_grid, ls, alpha1, dx1, rho1, rho2, mu1, mu2):
    """The reference's heaviside (SuLi-IC build, t_tens = 0). Returns
    (hs, rho, mu)."""
    g = case_grid
    L = ref_lib("ic")
    _ref_setup(L, g, 0)
    ls = np.asfortranarray(ls, dtype=np.float64)
    hs, rho, mu = (np.empty_like(ls, order="F") for _ in range(3))
    L.suli_ref_heaviside(_ptr(ls), alpha1, dx1, rho1, rho2, mu1, mu2, _ptr(hs), _ptr(rho), _ptr(mu))
    return hs, rho, mu


def ref_it_posdin(f, prd1, dz, eta0):
    """The reference's SuLi-IT posdin (with dz_it) on fields of its compiled
    size. Returns (u, v, w, dz, dzx, dzy, dzz, eta1, prd)."""
    L = ref_lib("it")
    u, v, w, eta1 = (f.u.copy(order="F"), f.v.copy(order="F"), f.w.copy(order="F"), f.eta1.copy(order="F"))
    dz, dzx, dzy, dzz = (np.asfortranarray(dz).copy(order="F"), f.dzx.copy(order="F"), f.dzy.copy(order="F"),
                         f.dzz.copy(order="F"))
    prd = np.zeros_like(prd1, order="F")
    L.suli_ref_it_posdin(_ptr(u), _ptr(v), _ptr(w), _ptr(dz), _ptr(dzx), _ptr(dzy), _ptr(dzz),
                         _ptr(np.asfortranarray(eta0)), _ptr(eta1), _ptr(np.asfortranarray(prd1)), _ptr(prd))
    return u, v, w, dz, dzx, dzy, dzz, eta1, prd


def ref_exp_press(case_grid, rho, prd1, prd0):
    """The reference's exp_press (SuLi-IC, or SuLi-ICdz on a stretched
    grid). Returns (prd0 after the increment, px, py, pz)."""
    g = case_grid
    L = ref_lib("icdz" if g.stretched else "ic")
    _ref_setup(L, g, 0)
    out0 = np.asfortranarray(prd0).copy(order="F")
    px = np.zeros((g.nx + 1, g.ny, g.nz), order="F")
    py = np.zeros((g.nx, g.ny + 1, g.nz), order="F")
    pz = np.zeros((g.nx, g.ny, g.nz + 1), order="F")
    _big_stack_call(_stack_bytes(g), L.suli_ref_exp_press, _ptr(np.asfortranarray(rho, dtype=np.float64)),
                    _ptr(np.asfortranarray(prd1)), _ptr(out0), _ptr(px), _ptr(py), _ptr(pz))
    return out0, px, py, pz


def ref_heaviside_tens(case_grid, ls, alpha1, dx1, rho1, rho2, mu1, mu2):
    """The reference's heaviside with t_tens = 1 (SuLi-IC build): returns
    (hs, rho, mu, hsx, hsy, hsz)."""
    g = case_grid
    L = ref_lib("ic")
    _ref_setup(L, g, 0)
    ls = np.asfortranarray(ls, dtype=np.float64)
    out = [np.empty_like(ls, order="F") for _ in range(6)]
    _big_stack_call(_stack_bytes(g), L.suli_ref_heaviside_tens, _ptr(ls), alpha1, dx1, rho1, rho2, mu1, mu2,
                    *[_ptr(a) for a in out])
    return tuple(out)


def ref_posdin_hydro(case_grid, fields, hs, prd1):
    """The reference's posdin with t_press = 0 (its hydrostatic branch)."""
    import dataclasses
    g = dataclasses.replace(case_grid, t_press=0)
    L = ref_lib("icdz" if g.stretched else "ic")
    _ref_setup(L, g, 0)
    u, v, w = (fields.u.copy(order="F"), fields.v.copy(order="F"), fields.w.copy(order="F"))
    prd = np.zeros_like(prd1, order="F")
    _big_stack_call(_stack_bytes(g), L.suli_ref_posdin_hydro, _ptr(u), _ptr(v), _ptr(w),
                    _ptr(np.asfortranarray(fields.rho)), _ptr(np.asfortranarray(hs, dtype=np.float64)), _ptr(np.asfortranarray(prd1)),
                    _ptr(prd))
    return u, v, w, prd
