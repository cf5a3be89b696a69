"""Python mirror of SuLi's pressure step on the MI355X library.

``PressureSolver.graddin(grid, fields)`` and ``PressureSolver.posdin(grid,
fields)`` do what SuLi's argument-less ``graddin()`` and ``posdin()`` do
(SuLi-IC/9_grad.f90:10-663, SuLi-IC/10_pos.f90:1-99) to the module arrays,
here passed explicitly as a :class:`suli_amd.cases.Grid` and
:class:`suli_amd.cases.Fields` in the reference layouts.  Every call goes
through the C ABI of libsuli_ps.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from .cases import Fields, Grid


class SuliPsError(RuntimeError):
    pass


def _c_grid(g: Grid) -> _lib.Grid:
    cg = _lib.Grid()
    cg.nx, cg.ny, cg.nz = g.nx, g.ny, g.nz
    cg.ccx0, cg.ccy0 = g.ccx0, g.ccy0
    cg.t_press, cg.obst_t, cg.ibm_t = g.t_press, g.obst_t, g.ibm_t
    cg.dx, cg.dy, cg.dz, cg.dt = g.dx, g.dy, g.dz, g.dt
    cg.rho_f1 = g.rho_f1
    cg.top_bc = getattr(g, "top_bc", 0)
    return cg


def _host_ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    if a.dtype != np.float64 or not (a.flags.f_contiguous or a.ndim == 1):
        raise ValueError("arrays must be float64 in Fortran order")
    return C.c_void_p(a.ctypes.data)


PRECISIONS = {"fp64": _lib.SULI_PS_FP64, "mixed": _lib.SULI_PS_MIXED}
# flag in a tile shape's vx: the FP32 step's pipelined plane walk (two or
# four cells per lane; SULI_PS_TILES_PIPELINED)
TILES_PIPELINED = _lib.SULI_PS_TILES_PIPELINED


def tiles_dict(kz: int, vx: int) -> dict:
    """A tile shape as (kz, vx) codes it, spelt out."""
    return {"kz": kz, "vx": vx & ~TILES_PIPELINED, "pipelined": bool(vx & TILES_PIPELINED)}
STOP_RULES = {"reference": _lib.SULI_PS_STOP_REFERENCE, "relative": _lib.SULI_PS_STOP_RELATIVE}


class DeviceFields:
    """Device copies of the pressure-step arrays (reference layouts)."""

    NAMES = ("u", "v", "w", "rho", "prd1", "obs_lsx", "obs_lsy", "obs_lsz", "dz", "dzz", "solid",
             "hs")

    def __init__(self, solver: "PressureSolver", grid: Grid, fields: Fields):
        self.solver = solver
        self.shapes = {}
        self.ptrs = {}
        host = dict(u=fields.u, v=fields.v, w=fields.w, rho=fields.rho, prd1=fields.prd1,
                    obs_lsx=fields.obs_lsx, obs_lsy=fields.obs_lsy, obs_lsz=fields.obs_lsz,
                    dz=grid.dzv, dzz=grid.dzzv, solid=getattr(fields, "solid", None),
                    hs=getattr(fields, "hs", None))
        self.dtypes = {}
        for name, arr in host.items():
            if arr is None:
                self.ptrs[name] = None
                continue
            dt = np.uint8 if name == "solid" else np.float64
            arr = np.asfortranarray(arr, dtype=dt)
            self.dtypes[name] = dt
            p = solver.lib.suli_ps_dev_alloc(solver.ctx, arr.nbytes)
            if not p:
                self.free()
                raise SuliPsError(f"device allocation of {name} ({arr.nbytes} B) failed")
            self.ptrs[name] = C.c_void_p(p)
            self.shapes[name] = arr.shape
            solver._check(solver.lib.suli_ps_memcpy_h2d(solver.ctx, self.ptrs[name],
                                                         C.c_void_p(arr.ctypes.data), arr.nbytes))

    def download(self, name: str) -> np.ndarray:
        out = np.zeros(self.shapes[name], dtype=self.dtypes.get(name, np.float64), order="F")
        self.solver._check(self.solver.lib.suli_ps_memcpy_d2h(
            self.solver.ctx, C.c_void_p(out.ctypes.data), self.ptrs[name], out.nbytes))
        return out

    def upload(self, name: str, arr: np.ndarray) -> None:
        arr = np.asfortranarray(arr, dtype=self.dtypes.get(name, np.float64))
        assert arr.shape == self.shapes[name]
        self.solver._check(self.solver.lib.suli_ps_memcpy_h2d(
            self.solver.ctx, self.ptrs[name], C.c_void_p(arr.ctypes.data), arr.nbytes))

    def zero(self, name: str) -> None:
        """Asynchronously zero a device array (e.g. prd1 for a cold start)."""
        nbytes = int(np.prod(self.shapes[name])) * np.dtype(self.dtypes.get(name, np.float64)).itemsize
        self.solver._check(self.solver.lib.suli_ps_dev_zero(self.solver.ctx, self.ptrs[name], nbytes))

    def free(self) -> None:
        for name, p in list(self.ptrs.items()):
            if p is not None:
                self.solver.lib.suli_ps_dev_free(self.solver.ctx, p)
            self.ptrs[name] = None


class PressureSolver:
    def __init__(self, precision: str = "fp64", stop_rule: str = "reference", tol: float = 1e-8,
                 max_iter: int = 10000, chunk: int = 0, verbose: bool = True, timing: bool = False,
                 keep_null: bool = False, graphs: bool = True, autotune: bool = False,
                 reuse_staged: bool = False, stored_residual: bool = False):
        self.lib = _lib.load()
        self.opts = _lib.Opts()
        self._set(precision, stop_rule, tol, max_iter, chunk, verbose, timing, keep_null, graphs,
                  autotune, reuse_staged, stored_residual)
        self.ctx = C.c_void_p(self.lib.suli_ps_create(C.byref(self.opts)))
        if not self.ctx:
            raise SuliPsError("suli_ps_create failed (no HIP device?)")

    def _set(self, precision, stop_rule, tol, max_iter, chunk, verbose, timing, keep_null,
             graphs=True, autotune=False, reuse_staged=False, stored_residual=False):
        self.opts.precision = PRECISIONS[precision]
        self.opts.stop_rule = STOP_RULES[stop_rule]
        self.opts.tol = tol
        self.opts.max_iter = max_iter
        self.opts.chunk = chunk
        self.opts.verbose = 1 if verbose else 0
        self.opts.timing = 1 if timing else 0
        self.opts.keep_null = 1 if keep_null else 0
        self.opts.no_graphs = 0 if graphs else 1
        self.opts.autotune = 1 if autotune else 0
        self.opts.reuse_staged = 1 if reuse_staged else 0
        self.opts.stored_residual = 1 if stored_residual else 0

    def configure(self, precision: str = "fp64", stop_rule: str = "reference", tol: float = 1e-8,
                  max_iter: int = 10000, chunk: int = 0, verbose: bool = True,
                  timing: bool = False, keep_null: bool = False, graphs: bool = True,
                  autotune: bool = False, reuse_staged: bool = False, stored_residual: bool = False) -> None:
        """Solver options (include/suli_ps.h, suli_ps_opts)."""
        self._set(precision, stop_rule, tol, max_iter, chunk, verbose, timing, keep_null, graphs,
                  autotune, reuse_staged, stored_residual)
        self._check(self.lib.suli_ps_set_opts(self.ctx, C.byref(self.opts)))

    def _check(self, rc: int) -> int:
        if rc < 0:
            msg = self.lib.suli_ps_last_error(self.ctx)
            raise SuliPsError(f"suli_ps error {rc}: {msg.decode() if msg else ''}")
        return rc

    # -- host arrays --------------------------------------------------------
    def graddin(self, grid: Grid, fields: Fields) -> _lib.Stats:
        """Solve for fields.prd1 in place (whole halo box)."""
        st = _lib.Stats()
        cg = _c_grid(grid)
        self._check(self.lib.suli_ps_graddin(
            self.ctx, C.byref(cg), _host_ptr(fields.u), _host_ptr(fields.v), _host_ptr(fields.w),
            _host_ptr(fields.rho), _host_ptr(fields.obs_lsx), _host_ptr(fields.obs_lsy),
            _host_ptr(fields.obs_lsz), _host_ptr(grid.dzv), _host_ptr(grid.dzzv),
            _host_ptr(fields.prd1), C.byref(st)))
        return st

    def posdin(self, grid: Grid, fields: Fields) -> np.ndarray:
        """Correct fields.u/v/w in place; returns prd (= prd1)."""
        prd = np.zeros_like(fields.prd1, order="F")
        cg = _c_grid(grid)
        self._check(self.lib.suli_ps_posdin(
            self.ctx, C.byref(cg), _host_ptr(fields.u), _host_ptr(fields.v), _host_ptr(fields.w),
            _host_ptr(fields.rho), _host_ptr(fields.prd1), _host_ptr(grid.dzv),
            _host_ptr(grid.dzzv), _host_ptr(prd)))
        return prd

    # -- device-resident arrays ---------------------------------------------
    def upload(self, grid: Grid, fields: Fields) -> DeviceFields:
        return DeviceFields(self, grid, fields)

    def graddin_dev(self, grid: Grid, df: DeviceFields) -> _lib.Stats:
        st = _lib.Stats()
        cg = _c_grid(grid)
        p = df.ptrs
        if p.get("solid") is not None:
            self._check(self.lib.suli_ps_graddin_masked_dev(
                self.ctx, C.byref(cg), p["u"], p["v"], p["w"], p["rho"], p["obs_lsx"], p["obs_lsy"],
                p["obs_lsz"], p["dz"], p["dzz"], p["solid"], p["prd1"], C.byref(st)))
        else:
            self._check(self.lib.suli_ps_graddin_dev(
                self.ctx, C.byref(cg), p["u"], p["v"], p["w"], p["rho"], p["obs_lsx"], p["obs_lsy"],
                p["obs_lsz"], p["dz"], p["dzz"], p["prd1"], C.byref(st)))
        return st

    def graddin_noalloc_dev(self, grid: Grid, df: DeviceFields) -> _lib.Stats:
        """graddin_dev whose solver state lives in the library's fixed device
        array (suli_ps_graddin_noalloc_dev): no allocation during the call,
        problems up to NOALLOC_MAX_BYTES of state, no solid mask."""
        st = _lib.Stats()
        cg = _c_grid(grid)
        p = df.ptrs
        if p.get("solid") is not None:
            raise ValueError("the no-allocation solve takes no solid-cell mask")
        self._check(self.lib.suli_ps_graddin_noalloc_dev(
            self.ctx, C.byref(cg), p["u"], p["v"], p["w"], p["rho"], p["obs_lsx"], p["obs_lsy"],
            p["obs_lsz"], p["dz"], p["dzz"], p["prd1"], C.byref(st)))
        return st

    def graddin_batch_dev(self, items) -> list:
        """Solve many problems together; ``items`` is a sequence of
        ``(grid, DeviceFields)``.  Returns one Stats per problem."""
        n = len(items)
        arr = (_lib.Problem * n)()
        for i, (grid, df) in enumerate(items):
            arr[i].grid = _c_grid(grid)
            for name in ("u", "v", "w", "rho", "obs_lsx", "obs_lsy", "obs_lsz", "dz", "dzz", "prd1",
                         "solid"):
                setattr(arr[i], name, df.ptrs.get(name))
        self._check(self.lib.suli_ps_graddin_batch_dev(self.ctx, arr, n))
        out = []
        for i in range(n):
            st = _lib.Stats()
            C.pointer(st)[0] = arr[i].stats
            out.append(st)
        return out

    def posdin_dev(self, grid: Grid, df: DeviceFields) -> None:
        cg = _c_grid(grid)
        p = df.ptrs
        if p.get("solid") is not None:
            self._check(self.lib.suli_ps_posdin_masked_dev(
                self.ctx, C.byref(cg), p["u"], p["v"], p["w"], p["rho"], p["prd1"], p["dz"], p["dzz"],
                p["solid"], None))
        else:
            self._check(self.lib.suli_ps_posdin_dev(
                self.ctx, C.byref(cg), p["u"], p["v"], p["w"], p["rho"], p["prd1"], p["dz"], p["dzz"],
                None))

    def pressh(self, grid: Grid, fields: Fields, gz: float) -> np.ndarray:
        """Hydrostatic prd1 (t_press = 0), host arrays; prd1 updated in place."""
        cg = _c_grid(grid)
        self._check(self.lib.suli_ps_pressh(self.ctx, C.byref(cg), _host_ptr(fields.rho),
                                            _host_ptr(grid.dzv), gz, _host_ptr(fields.prd1)))
        return fields.prd1

    def posdin_hydro(self, grid: Grid, fields: Fields) -> np.ndarray:
        """posdin's hydrostatic branch on host arrays (fields.hs required);
        u, v, w updated in place, returns prd."""
        cg = _c_grid(grid)
        prd = np.zeros_like(fields.prd1, order="F")
        self._check(self.lib.suli_ps_posdin_hydro(
            self.ctx, C.byref(cg), _host_ptr(fields.u), _host_ptr(fields.v), _host_ptr(fields.w),
            _host_ptr(fields.rho), _host_ptr(fields.hs), _host_ptr(fields.prd1), _host_ptr(grid.dzv),
            _host_ptr(prd)))
        return prd

    def pressh_dev(self, grid: Grid, df: DeviceFields, gz: float) -> None:
        cg = _c_grid(grid)
        p = df.ptrs
        self._check(self.lib.suli_ps_pressh_dev(self.ctx, C.byref(cg), p["rho"], p["dz"], gz, p["prd1"]))

    def posdin_hydro_dev(self, grid: Grid, df: DeviceFields) -> None:
        cg = _c_grid(grid)
        p = df.ptrs
        self._check(self.lib.suli_ps_posdin_hydro_dev(
            self.ctx, C.byref(cg), p["u"], p["v"], p["w"], p["rho"], p["hs"], p["prd1"], p["dz"], None))

    def heaviside_dev(self, grid: Grid, ls_ptr, eps: float, rho_f2: float, mu_f1: float, mu_f2: float,
                      rho_ptr, hs_ptr=None, mu_ptr=None) -> None:
        """rho (and hs, mu) from a device-resident level set (raw device
        pointers, e.g. DeviceFields.ptrs entries or suli_ps_dev_alloc)."""
        cg = _c_grid(grid)
        self._check(self.lib.suli_ps_heaviside_dev(self.ctx, C.byref(cg), ls_ptr, eps, rho_f2, mu_f1, mu_f2,
                                                   hs_ptr, rho_ptr, mu_ptr))

    def heaviside_tens_dev(self, grid: Grid, ls_ptr, eps: float, rho_f2: float, mu_f1: float, mu_f2: float,
                           rho_ptr, hsx_ptr, hsy_ptr, hsz_ptr, hs_ptr=None, mu_ptr=None) -> None:
        """heaviside with t_tens = 1: rho (hs, mu) and the surface-tension
        gradients hsx, hsy, hsz from a device-resident level set."""
        cg = _c_grid(grid)
        self._check(self.lib.suli_ps_heaviside_tens_dev(self.ctx, C.byref(cg), ls_ptr, eps, rho_f2, mu_f1, mu_f2,
                                                        hs_ptr, rho_ptr, mu_ptr, hsx_ptr, hsy_ptr, hsz_ptr))

    def exp_press(self, grid: Grid, rho: np.ndarray, prd1: np.ndarray, prd0: np.ndarray):
        """exp_press (IPA) on host arrays: prd0 incremented in place; returns
        (px, py, pz)."""
        cg = _c_grid(grid)
        px = np.zeros((grid.nx + 1, grid.ny, grid.nz), order="F")
        py = np.zeros((grid.nx, grid.ny + 1, grid.nz), order="F")
        pz = np.zeros((grid.nx, grid.ny, grid.nz + 1), order="F")
        self._check(self.lib.suli_ps_exp_press(self.ctx, C.byref(cg), _host_ptr(rho), _host_ptr(prd1),
                                               _host_ptr(prd0), _host_ptr(grid.dzv), _host_ptr(grid.dzzv),
                                               _host_ptr(px), _host_ptr(py), _host_ptr(pz)))
        return px, py, pz

    def exp_press_dev(self, grid: Grid, rho_ptr, prd1_ptr, prd0_ptr, px_ptr, py_ptr, pz_ptr, dz_ptr=None,
                      dzz_ptr=None) -> None:
        cg = _c_grid(grid)
        self._check(self.lib.suli_ps_exp_press_dev(self.ctx, C.byref(cg), rho_ptr, prd1_ptr, prd0_ptr, dz_ptr,
                                                   dzz_ptr, px_ptr, py_ptr, pz_ptr))

    def dev_array(self, arr: np.ndarray):
        """Copy a float64 array to a new device allocation; returns the pointer
        (free with dev_free)."""
        arr = np.asfortranarray(arr, dtype=np.float64)
        p = self.lib.suli_ps_dev_alloc(self.ctx, arr.nbytes)
        if not p:
            raise SuliPsError(f"device allocation of {arr.nbytes} B failed")
        self._check(self.lib.suli_ps_memcpy_h2d(self.ctx, C.c_void_p(p), C.c_void_p(arr.ctypes.data),
                                                arr.nbytes))
        return C.c_void_p(p)

    def dev_download(self, ptr, shape) -> np.ndarray:
        out = np.zeros(shape, dtype=np.float64, order="F")
        self._check(self.lib.suli_ps_memcpy_d2h(self.ctx, C.c_void_p(out.ctypes.data), ptr, out.nbytes))
        return out

    def dev_free(self, ptr) -> None:
        self.lib.suli_ps_dev_free(self.ctx, ptr)

    def dev_zero(self, ptr, nbytes: int) -> None:
        """Zero nbytes of a device allocation (on the context's stream)."""
        self._check(self.lib.suli_ps_dev_zero(self.ctx, ptr, nbytes))

    # -- SuLi-IT -------------------------------------------------------------
    @staticmethod
    def _it_cgrid(g) -> _lib.Grid:
        cg = _lib.Grid()
        cg.nx, cg.ny, cg.nz = g.nx, g.ny, g.nz
        cg.ccx0, cg.ccy0 = g.ccx0, g.ccy0
        cg.t_press = 1
        cg.dx, cg.dy, cg.dz, cg.dt = g.dx, g.dy, 0.0, g.dt
        return cg

    def it_graddin(self, g, f) -> _lib.Stats:
        """SuLi-IT's pressure solve on host arrays (cases.ITGrid, ITFields);
        f.prd1 updated in place."""
        st = _lib.Stats()
        cg = self._it_cgrid(g)
        self._check(self.lib.suli_ps_it_graddin(
            self.ctx, C.byref(cg), g.tetah, g.gz, _host_ptr(f.u), _host_ptr(f.v), _host_ptr(f.w),
            _host_ptr(f.dzx), _host_ptr(f.dzy), _host_ptr(f.dzz), _host_ptr(f.eta1), _host_ptr(f.deltah),
            _host_ptr(f.prd1), C.byref(st)))
        return st

    def it_posdin(self, g, f) -> np.ndarray:
        """SuLi-IT's velocity correction and eta1 update on host arrays;
        returns prd."""
        cg = self._it_cgrid(g)
        prd = np.zeros_like(f.prd1, order="F")
        self._check(self.lib.suli_ps_it_posdin(
            self.ctx, C.byref(cg), g.tetah, g.gz, _host_ptr(f.u), _host_ptr(f.v), _host_ptr(f.w),
            _host_ptr(f.dzz), _host_ptr(f.eta1), _host_ptr(f.prd1), _host_ptr(prd)))
        return prd

    def it_posdin_layers(self, g, f, dz: np.ndarray, eta0: np.ndarray) -> np.ndarray:
        """SuLi-IT's whole posdin (surface, layer thicknesses, correction) on
        host arrays: f.u, v, w, dzx, dzy, dzz, eta1 and dz updated in place;
        returns prd."""
        cg = self._it_cgrid(g)
        prd = np.zeros_like(f.prd1, order="F")
        self._check(self.lib.suli_ps_it_posdin_layers(
            self.ctx, C.byref(cg), g.tetah, g.gz, _host_ptr(f.u), _host_ptr(f.v), _host_ptr(f.w), _host_ptr(dz),
            _host_ptr(f.dzx), _host_ptr(f.dzy), _host_ptr(f.dzz), _host_ptr(eta0), _host_ptr(f.eta1),
            _host_ptr(f.prd1), _host_ptr(prd)))
        return prd

    def it_posdin_layers_dev(self, g, d: dict, dz_ptr, eta0_ptr) -> None:
        cg = self._it_cgrid(g)
        self._check(self.lib.suli_ps_it_posdin_layers_dev(
            self.ctx, C.byref(cg), g.tetah, g.gz, d["u"], d["v"], d["w"], dz_ptr, d["dzx"], d["dzy"], d["dzz"],
            eta0_ptr, d["eta1"], d["prd1"], None))

    def it_upload(self, f) -> dict:
        """Device copies of ITFields (name -> pointer); free with it_free."""
        return {name: self.dev_array(getattr(f, name))
                for name in ("u", "v", "w", "dzx", "dzy", "dzz", "eta1", "deltah", "prd1")}

    def it_free(self, d: dict) -> None:
        for p in d.values():
            self.dev_free(p)

    def it_graddin_dev(self, g, d: dict) -> _lib.Stats:
        st = _lib.Stats()
        cg = self._it_cgrid(g)
        self._check(self.lib.suli_ps_it_graddin_dev(
            self.ctx, C.byref(cg), g.tetah, g.gz, d["u"], d["v"], d["w"], d["dzx"], d["dzy"], d["dzz"],
            d["eta1"], d["deltah"], d["prd1"], C.byref(st)))
        return st

    def it_posdin_dev(self, g, d: dict) -> None:
        cg = self._it_cgrid(g)
        self._check(self.lib.suli_ps_it_posdin_dev(
            self.ctx, C.byref(cg), g.tetah, g.gz, d["u"], d["v"], d["w"], d["dzz"], d["eta1"], d["prd1"], None))

    def time_spmv(self, precision: str = "fp64", reps: int = 50) -> float:
        ms = C.c_double()
        self._check(self.lib.suli_ps_time_spmv(self.ctx, PRECISIONS[precision], reps, C.byref(ms)))
        return ms.value

    def autotune(self, grid: Grid, precision: str = "mixed"):
        """(kz, vx, us per CG step) chosen for this grid size (tuned on first use);
        vx carries TILES_PIPELINED for the pipelined FP32 step."""
        kz = C.c_int32()
        vx = C.c_int32()
        us = C.c_double()
        cg = _c_grid(grid)
        self._check(self.lib.suli_ps_autotune(self.ctx, C.byref(cg), PRECISIONS[precision],
                                              C.byref(kz), C.byref(vx), C.byref(us)))
        return kz.value, vx.value, us.value

    def set_tiles(self, grid: Grid, precision: str, kz: int, vx: int) -> None:
        """Fix the tile shape used for this grid size (with autotune on); vx 1,
        2 or 4, or 2 or 4 | TILES_PIPELINED."""
        cg = _c_grid(grid)
        self._check(self.lib.suli_ps_set_tiles(self.ctx, C.byref(cg), PRECISIONS[precision], kz, vx))

    def copy_bench(self, nbytes: int, reps: int = 20) -> float:
        """Mean time (ms) of the library's plain device copy of nbytes."""
        ms = C.c_double()
        self._check(self.lib.suli_ps_copy_bench(self.ctx, nbytes, reps, C.byref(ms)))
        return ms.value

    def read_bench(self, nbytes: int, reps: int = 20) -> float:
        """Mean time (ms) of the library's plain device read of nbytes."""
        ms = C.c_double()
        self._check(self.lib.suli_ps_read_bench(self.ctx, nbytes, reps, C.byref(ms)))
        return ms.value

    def synchronize(self) -> None:
        self._check(self.lib.suli_ps_synchronize(self.ctx))

    def close(self) -> None:
        if self.ctx:
            self.lib.suli_ps_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
