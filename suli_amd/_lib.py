"""ctypes binding of libsuli_ps.so (the C ABI in include/suli_ps.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsuli_ps.so")

SULI_PS_FP64 = 0
SULI_PS_MIXED = 1
SULI_PS_STOP_REFERENCE = 0
SULI_PS_STOP_RELATIVE = 1
SULI_PS_OK = 0
SULI_PS_NOT_CONVERGED = 1
SULI_PS_EAGAIN = -4
SULI_PS_TILES_PIPELINED = 8

_dp = C.POINTER(C.c_double)

# Every symbol include/suli_ps.h declares.
EXPORTS = (
    "suli_ps_abi_version", "suli_ps_create", "suli_ps_destroy", "suli_ps_set_opts",
    "suli_ps_last_error", "suli_ps_graddin", "suli_ps_posdin", "suli_ps_graddin_dev",
    "suli_ps_posdin_dev", "suli_ps_dev_alloc", "suli_ps_dev_free", "suli_ps_memcpy_h2d",
    "suli_ps_memcpy_d2h", "suli_ps_dev_zero", "suli_ps_synchronize", "suli_ps_time_spmv",
    "suli_ps_graddin_batch_dev", "suli_ps_server_create", "suli_ps_server_destroy",
    "suli_ps_server_last_error", "suli_ps_server_submit", "suli_ps_server_wait",
    "suli_ps_server_get_info", "suli_ps_copy_bench", "suli_ps_read_bench", "suli_ps_autotune",
    "suli_ps_graddin_masked_dev", "suli_ps_posdin_masked_dev", "suli_ps_set_tiles",
    "suli_ps_pressh", "suli_ps_pressh_dev", "suli_ps_posdin_hydro", "suli_ps_posdin_hydro_dev",
    "suli_ps_heaviside_dev", "suli_ps_restart_read", "suli_ps_it_graddin", "suli_ps_it_graddin_dev",
    "suli_ps_it_posdin", "suli_ps_it_posdin_dev", "suli_ps_graddin_noalloc_dev",
    "suli_ps_heaviside_tens_dev", "suli_ps_exp_press", "suli_ps_exp_press_dev",
    "suli_ps_it_posdin_layers", "suli_ps_it_posdin_layers_dev",
)

# SULI_PS_NOALLOC_MAX_BYTES: the fixed device array of the no-allocation solve
NOALLOC_MAX_BYTES = 64 << 20


class Grid(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("ccx0", C.c_int32), ("ccy0", C.c_int32),
        ("t_press", C.c_int32), ("obst_t", C.c_int32), ("ibm_t", C.c_int32),
        ("top_bc", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("rho_f1", C.c_double),
    ]


class Restart(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in
                ("u", "v", "w", "prd0", "prd1", "prd", "rho", "ls_mu", "ls", "hs")]


class Opts(C.Structure):
    _fields_ = [
        ("precision", C.c_int32), ("stop_rule", C.c_int32), ("tol", C.c_double),
        ("max_iter", C.c_int32), ("chunk", C.c_int32), ("verbose", C.c_int32),
        ("timing", C.c_int32), ("keep_null", C.c_int32), ("no_graphs", C.c_int32),
        ("autotune", C.c_int32), ("reuse_staged", C.c_int32), ("stored_residual", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("iters", C.c_int32), ("refreshes", C.c_int32),
        ("spmv_timed", C.c_int32),
        ("rr0", C.c_double), ("bb", C.c_double), ("rr", C.c_double), ("thresh", C.c_double),
        ("rel_true", C.c_double), ("solve_ms", C.c_double), ("spmv_ms", C.c_double),
        ("graph_replays", C.c_int32), ("reserved", C.c_int32),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Problem(C.Structure):
    _fields_ = [("grid", Grid)] + [(n, C.c_void_p) for n in (
        "u", "v", "w", "rho", "obs_lsx", "obs_lsy", "obs_lsz", "dz", "dzz", "prd1", "solid")] + [
        ("stats", Stats)]


class ServerOpts(C.Structure):
    _fields_ = [("solve", Opts), ("hbm_budget", C.c_uint64), ("max_active", C.c_int32),
                ("reserved", C.c_int32)]


class ServerInfo(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("queued", C.c_int32),
                ("resident", C.c_int32), ("resident_bytes", C.c_uint64),
                ("peak_resident_bytes", C.c_uint64), ("budget", C.c_uint64), ("chunks", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m suli_amd.build` "
                           "(there is no CPU fallback)")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.suli_ps_abi_version.restype = C.c_int
    L.suli_ps_create.restype = vp
    L.suli_ps_create.argtypes = [C.POINTER(Opts)]
    L.suli_ps_destroy.argtypes = [vp]
    L.suli_ps_set_opts.argtypes = [vp, C.POINTER(Opts)]
    L.suli_ps_last_error.restype = C.c_char_p
    L.suli_ps_last_error.argtypes = [vp]
    gargs = [vp, C.POINTER(Grid)] + [vp] * 10 + [C.POINTER(Stats)]
    L.suli_ps_graddin.argtypes = gargs
    L.suli_ps_graddin_dev.argtypes = gargs
    L.suli_ps_graddin_noalloc_dev.argtypes = gargs
    pargs = [vp, C.POINTER(Grid)] + [vp] * 8
    L.suli_ps_posdin.argtypes = pargs
    L.suli_ps_posdin_dev.argtypes = pargs
    L.suli_ps_graddin_masked_dev.argtypes = [vp, C.POINTER(Grid)] + [vp] * 11 + [C.POINTER(Stats)]
    L.suli_ps_posdin_masked_dev.argtypes = [vp, C.POINTER(Grid)] + [vp] * 9
    L.suli_ps_pressh.argtypes = [vp, C.POINTER(Grid), vp, vp, C.c_double, vp]
    L.suli_ps_pressh_dev.argtypes = [vp, C.POINTER(Grid), vp, vp, C.c_double, vp]
    L.suli_ps_posdin_hydro.argtypes = [vp, C.POINTER(Grid)] + [vp] * 8
    L.suli_ps_posdin_hydro_dev.argtypes = [vp, C.POINTER(Grid)] + [vp] * 8
    L.suli_ps_restart_read.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.POINTER(Restart),
                                       C.c_char_p, C.c_size_t]
    it_g = [vp, C.POINTER(Grid), C.c_double, C.c_double] + [vp] * 9 + [C.POINTER(Stats)]
    L.suli_ps_it_graddin.argtypes = it_g
    L.suli_ps_it_graddin_dev.argtypes = it_g
    it_p = [vp, C.POINTER(Grid), C.c_double, C.c_double] + [vp] * 7
    L.suli_ps_it_posdin.argtypes = it_p
    L.suli_ps_it_posdin_dev.argtypes = it_p
    L.suli_ps_heaviside_dev.argtypes = [vp, C.POINTER(Grid), vp] + [C.c_double] * 4 + [vp] * 3
    L.suli_ps_heaviside_tens_dev.argtypes = [vp, C.POINTER(Grid), vp] + [C.c_double] * 4 + [vp] * 6
    L.suli_ps_exp_press.argtypes = [vp, C.POINTER(Grid)] + [vp] * 8
    L.suli_ps_exp_press_dev.argtypes = [vp, C.POINTER(Grid)] + [vp] * 8
    it_l = [vp, C.POINTER(Grid), C.c_double, C.c_double] + [vp] * 11
    L.suli_ps_it_posdin_layers.argtypes = it_l
    L.suli_ps_it_posdin_layers_dev.argtypes = it_l
    L.suli_ps_dev_alloc.restype = vp
    L.suli_ps_dev_alloc.argtypes = [vp, C.c_size_t]
    L.suli_ps_dev_free.argtypes = [vp, vp]
    L.suli_ps_memcpy_h2d.argtypes = [vp, vp, vp, C.c_size_t]
    L.suli_ps_memcpy_d2h.argtypes = [vp, vp, vp, C.c_size_t]
    L.suli_ps_dev_zero.argtypes = [vp, vp, C.c_size_t]
    L.suli_ps_synchronize.argtypes = [vp]
    L.suli_ps_time_spmv.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    L.suli_ps_graddin_batch_dev.argtypes = [vp, C.POINTER(Problem), C.c_int32]
    L.suli_ps_copy_bench.argtypes = [vp, C.c_size_t, C.c_int32, C.POINTER(C.c_double)]
    L.suli_ps_read_bench.argtypes = [vp, C.c_size_t, C.c_int32, C.POINTER(C.c_double)]
    L.suli_ps_autotune.argtypes = [vp, C.POINTER(Grid), C.c_int32, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.POINTER(C.c_double)]
    L.suli_ps_set_tiles.argtypes = [vp, C.POINTER(Grid), C.c_int32, C.c_int32, C.c_int32]
    L.suli_ps_server_create.restype = vp
    L.suli_ps_server_create.argtypes = [C.POINTER(ServerOpts)]
    L.suli_ps_server_destroy.argtypes = [vp]
    L.suli_ps_server_last_error.restype = C.c_char_p
    L.suli_ps_server_last_error.argtypes = [vp]
    L.suli_ps_server_submit.restype = C.c_int64
    L.suli_ps_server_submit.argtypes = [vp, C.POINTER(Grid)] + [vp] * 10
    L.suli_ps_server_wait.argtypes = [vp, C.c_int64, C.c_double, C.POINTER(Stats)]
    L.suli_ps_server_get_info.argtypes = [vp, C.POINTER(ServerInfo)]
    _lib = L
    return L
