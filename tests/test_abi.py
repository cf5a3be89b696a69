"""The C ABI library loads and exports every function include/suli_ps.h
declares (no device calls: these run without a GPU)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from suli_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "suli_ps.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(suli_ps_[a-z0-9_]+)\s*\(", text)))


def test_header_lists_python_exports():
    assert header_functions() == sorted(_lib.EXPORTS)


def test_header_constants_match_python():
    """The integer constants the binding repeats have the header's values."""
    text = open(HEADER).read()
    defines = dict(re.findall(r"#define\s+(SULI_PS_[A-Z0-9_]+)\s+(-?\d+)\b", text))
    enums = dict(re.findall(r"\b(SULI_PS_[A-Z0-9_]+)\s*=\s*(-?\d+)", text))
    values = {**enums, **defines}
    for name in ("SULI_PS_FP64", "SULI_PS_MIXED", "SULI_PS_STOP_REFERENCE", "SULI_PS_STOP_RELATIVE",
                 "SULI_PS_OK", "SULI_PS_NOT_CONVERGED", "SULI_PS_EAGAIN", "SULI_PS_TILES_PIPELINED"):
        assert int(values[name]) == getattr(_lib, name), name


def test_library_exports_every_header_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libsuli_ps.so is not built (python -m suli_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing
    assert lib.suli_ps_abi_version() == 2


C_LAYOUT = r"""
#include <stddef.h>
#include <stdio.h>
#include "suli_ps.h"
#define S(t) printf(#t " %zu\n", sizeof(t))
#define O(t, f) printf(#t "." #f " %zu\n", offsetof(t, f))
int main(void)
{
    S(suli_ps_grid); O(suli_ps_grid, dx); O(suli_ps_grid, rho_f1);
    S(suli_ps_opts); O(suli_ps_opts, tol); O(suli_ps_opts, autotune); O(suli_ps_opts, reuse_staged);
    S(suli_ps_stats); O(suli_ps_stats, rr0); O(suli_ps_stats, spmv_ms); O(suli_ps_stats, graph_replays);
    S(suli_ps_problem); O(suli_ps_problem, u); O(suli_ps_problem, solid); O(suli_ps_problem, stats);
    S(suli_ps_server_opts); O(suli_ps_server_opts, hbm_budget); O(suli_ps_server_opts, max_active);
    S(suli_ps_server_info); O(suli_ps_server_info, chunks);
    S(suli_ps_restart); O(suli_ps_restart, hs);
    return 0;
}
"""


def c_layout(tmp_path):
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("gcc not available")
    src = tmp_path / "layout.c"
    src.write_text(C_LAYOUT)
    exe = tmp_path / "layout"
    subprocess.run([gcc, "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of the ABI structs have the sizes and offsets the C
    compiler gives the header's structs."""
    c = c_layout(tmp_path)
    mirrors = {"suli_ps_grid": _lib.Grid, "suli_ps_opts": _lib.Opts, "suli_ps_stats": _lib.Stats,
               "suli_ps_problem": _lib.Problem, "suli_ps_server_opts": _lib.ServerOpts,
               "suli_ps_server_info": _lib.ServerInfo, "suli_ps_restart": _lib.Restart}
    for key, value in c.items():
        struct, _, field = key.partition(".")
        cls = mirrors[struct]
        got = getattr(cls, field).offset if field else ctypes.sizeof(cls)
        assert got == value, (key, got, value)


FORTRAN_LAYOUT = """
program layout
   use iso_c_binding
   use suli_ps_iface
   type(suli_ps_grid) :: g
   type(suli_ps_opts) :: o
   type(suli_ps_stats) :: st
   print '(a,i0)', 'suli_ps_grid ', c_sizeof(g)
   print '(a,i0)', 'suli_ps_opts ', c_sizeof(o)
   print '(a,i0)', 'suli_ps_stats ', c_sizeof(st)
end program layout
"""


def test_fortran_types_match_header(tmp_path):
    """The Fortran bind(C) types of the drop-in have the C sizes."""
    from suli_amd import fortran_build
    try:
        fc = fortran_build.flang()
    except RuntimeError:
        pytest.skip("amdflang not available")
    c = c_layout(tmp_path)
    iface = os.path.join(ROOT, "suli_amd", "fortran", "suli_ps_iface.f90")
    prog = tmp_path / "layout.f90"
    prog.write_text(FORTRAN_LAYOUT)
    exe = tmp_path / "flayout"
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libsuli_ps.so is not built (python -m suli_amd.build)")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run([fc, "-module-dir", str(tmp_path), iface, str(prog), "-o", str(exe), "-L", libdir,
                    "-lsuli_ps", f"-Wl,-rpath,{libdir}"], check=True, cwd=tmp_path)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        name, size = line.split()
        assert int(size) == c[name], (name, size, c[name])


def test_fortran_shim_declares_same_entry_points():
    shim = open(os.path.join(ROOT, "suli_amd", "fortran", "suli_ps_iface.f90")).read()
    for name in ("suli_ps_create", "suli_ps_graddin", "suli_ps_posdin", "suli_ps_last_error",
                 "suli_ps_pressh", "suli_ps_posdin_hydro", "suli_ps_it_graddin", "suli_ps_it_posdin",
                 "suli_ps_it_posdin_layers", "suli_ps_exp_press"):
        assert f'name="{name}"' in shim, name
