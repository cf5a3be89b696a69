"""The C ABI library loads and exports every function include/suli_ps.h
declares (no device calls: these run without a GPU)."""
import ctypes
import os
import re

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


def test_library_exports_every_header_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libsuli_ps.so is not built (python -m suli_amd.build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing
    assert lib.suli_ps_abi_version() == 1


def test_struct_layouts_match_header():
    """ctypes mirrors of the ABI structs have the C sizes (natural alignment)."""
    assert ctypes.sizeof(_lib.Grid) == 9 * 4 + 4 + 5 * 8     # padded before the doubles
    assert _lib.Grid.dx.offset == 40
    assert ctypes.sizeof(_lib.Opts) == 4 + 4 + 8 + 7 * 4 + 4   # trailing pad to 8
    assert _lib.Opts.no_graphs.offset == 36 and _lib.Opts.autotune.offset == 40
    assert ctypes.sizeof(_lib.Stats) == 4 * 4 + 7 * 8
    assert ctypes.sizeof(_lib.Problem) == ctypes.sizeof(_lib.Grid) + 11 * 8 + ctypes.sizeof(_lib.Stats)
    assert _lib.Problem.stats.offset == ctypes.sizeof(_lib.Grid) + 11 * 8
    assert ctypes.sizeof(_lib.ServerOpts) == ctypes.sizeof(_lib.Opts) + 8 + 4 + 4
    assert ctypes.sizeof(_lib.ServerInfo) == 8 + 8 + 4 + 4 + 8 + 8 + 8 + 8


def test_fortran_shim_declares_same_entry_points():
    shim = open(os.path.join(ROOT, "suli_amd", "fortran", "suli_ps_iface.f90")).read()
    for name in ("suli_ps_create", "suli_ps_graddin", "suli_ps_posdin", "suli_ps_last_error",
                 "suli_ps_pressh", "suli_ps_posdin_hydro", "suli_ps_it_graddin", "suli_ps_it_posdin"):
        assert f'name="{name}"' in shim, name
