"""Compiles the Fortran side of the drop-in.

* suli_amd/fortran/suli_ps_iface.f90 (the iso_c_binding interface, which
  depends on nothing from SuLi) is always compiled, as a check.
* suli_ps_graddin.F90 uses SuLi's own modules (velpre, disc, cond, obst,
  ls_param), so it is compiled inside a SuLi build; where the reference
  sources are present, `make -C oracle dropin` builds the check programs
  oracle/_ref/dropin_ic and dropin_icdz that link it with SuLi's modules.
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
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "dropin"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"drop-in check programs failed to build:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built Fortran interface" + (r.stdout.strip() and f" ({r.stdout.strip()})"))


if __name__ == "__main__":
    build(verbose=True)
