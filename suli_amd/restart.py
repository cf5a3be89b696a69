"""SuLi restart files: the pressure step's arrays out of the one
list-directed record restart_salva writes (SuLi-IC/11_diversos.f90:581-599,
read back by restart_ini, :500-577), through suli_ps_restart_read (host
code in the library; no GPU needed)."""
from __future__ import annotations

import ctypes as C
from typing import Dict, Iterable

import numpy as np

from . import _lib
from .cases import Fields, Grid

SHAPES = {
    "u": lambda nx, ny, nz: (nx + 3, ny + 2, nz + 2),
    "v": lambda nx, ny, nz: (nx + 2, ny + 3, nz + 2),
    "w": lambda nx, ny, nz: (nx + 2, ny + 2, nz + 3),
    "prd0": lambda nx, ny, nz: (nx + 2, ny + 2, nz + 2),
    "prd1": lambda nx, ny, nz: (nx + 2, ny + 2, nz + 2),
    "prd": lambda nx, ny, nz: (nx + 2, ny + 2, nz + 2),
    "rho": lambda nx, ny, nz: (nx, ny, nz),
    "ls_mu": lambda nx, ny, nz: (nx, ny, nz),
    "ls": lambda nx, ny, nz: (nx, ny, nz),
    "hs": lambda nx, ny, nz: (nx, ny, nz),
}


def read_restart(path: str, nx: int, ny: int, nz: int,
                 names: Iterable[str] = ("u", "v", "w", "prd1", "rho")) -> Dict[str, np.ndarray]:
    """Arrays `names` (Fortran order, reference layouts) from a restart file
    of an nx*ny*nz grid.  Raises ValueError on a malformed or short file."""
    L = _lib.load()
    out = {}
    rs = _lib.Restart()
    for name in names:
        if name not in SHAPES:
            raise KeyError(f"{name} is not one of {sorted(SHAPES)}")
        a = np.zeros(SHAPES[name](nx, ny, nz), dtype=np.float64, order="F")
        out[name] = a
        setattr(rs, name, a.ctypes.data)
    err = C.create_string_buffer(512)
    rc = L.suli_ps_restart_read(str(path).encode(), nx, ny, nz, C.byref(rs), err, len(err))
    if rc != 0:
        raise ValueError(err.value.decode(errors="replace"))
    return out


def fields_from_restart(path: str, grid: Grid, with_hs: bool = False) -> Fields:
    """The pressure step's inputs (u, v, w, rho, prd1; hs on request) from a
    restart file, ready for PressureSolver."""
    names = ["u", "v", "w", "rho", "prd1"] + (["hs"] if with_hs else [])
    a = read_restart(path, grid.nx, grid.ny, grid.nz, names)
    return Fields(u=a["u"], v=a["v"], w=a["w"], rho=a["rho"], prd1=a["prd1"], hs=a.get("hs"))
