"""Compiles the Fortran side of the drop-in that stands alone.

suli_amd/fortran/suli_ps_iface.f90 (the iso_c_binding interface, which
depends on nothing from SuLi) is compiled here, as a check.  The shims
suli_ps_graddin*.F90 use SuLi's own modules (velpre, disc, cond, obst,
ls_param), so they are compiled inside a SuLi build; nothing here compiles
SuLi's sources.  (The test programs that link the shims with SuLi's modules
are built with the checker, `make -C oracle dropin`.)
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build", "fortran")


def flang() -> str:
    for c in (os.environ.get("SULI_FC"), "/opt/rocm/lib/llvm/bin/amdflang", shutil.which("amdflang")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("amdflang not found")


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    src = os.path.join(HERE, "fortran", "suli_ps_iface.f90")
    cmd = [flang(), "-O2", "-fPIC", "-module-dir", BUILD, "-c", src,
           "-o", os.path.join(BUILD, "suli_ps_iface.o")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran interface failed to compile:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built Fortran interface")


if __name__ == "__main__":
    build(verbose=True)
