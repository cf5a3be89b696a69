"""Writes the golden fixtures in this folder from the reference's own code.

Runs SuLi's graddin() and posdin() as compiled by oracle/Makefile from the
reference sources (SuLi-IC/{0_modules,8_derivadas,9_grad,10_pos}.f90 and the
SuLi-ICdz equivalents) on small seeded inputs, and stores inputs and
outputs as .npz files plus a sha256 manifest.  Needs the reference tree and
oracle/_ref/*.so, so it runs in the development container only:

    make -C oracle && python tests/golden/make_golden.py

Each input is scaled by a factor lam so that the reference's absolute stop
rule, r.r <= 1e-4/(nx*ny*nz) (SuLi-IC/9_grad.f90:393), corresponds to a
relative residual of about 1e-11 (the solution scales with lam).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from suli_amd import cases  # noqa: E402


def case_list():
    out = []
    out.append(("ic_walls_12x10x8", cases.random_case(12, 10, 8, seed=11)))
    out.append(("ic_px_16x16x8", cases.random_case(16, 16, 8, seed=12, periodic_x=True)))
    out.append(("ic_py_16x12x10", cases.random_case(16, 12, 10, seed=13, periodic_y=True)))
    out.append(("ic_ibm_16x16x12", cases.random_case(16, 16, 12, seed=14, ibm=True)))
    out.append(("ic_ipa_12x10x8", cases.random_case(12, 10, 8, seed=15, t_press=2)))
    c = cases.random_case(12, 10, 8, seed=16, stretched=True)
    cases.make_stretched_compatible(c)
    out.append(("icdz_12x10x8", c))
    # warm start: prd1 from a previous solve, ghosts mirrored as contorno(1) sets them
    c = cases.random_case(14, 10, 8, seed=17, ratio=10.0)
    rng = np.random.default_rng(99)
    p = c.fields.prd1
    p[1:-1, 1:-1, 1:-1] = rng.uniform(-5, 5, (14, 10, 8))
    p[0], p[-1] = p[1], p[-2]
    p[:, 0], p[:, -1] = p[:, 1], p[:, -2]
    p[:, :, 0], p[:, :, -1] = p[:, :, 1], p[:, :, -2]
    out.append(("ic_warm_14x10x8", c))
    return out


def scale_for_tight(c: cases.Case, rel: float = 1e-11) -> float:
    _, st, _ = oracle.graddin(c.grid, c.fields, stop_rule=oracle.STOP_RELATIVE, tol=1e-3)
    thresh = 1e-4 / c.grid.ncell
    return float(np.sqrt(thresh / (rel * rel * st.bb)))


def main():
    manifest = {}
    for name, c in case_list():
        lam = scale_for_tight(c)
        f = cases.scale_velocity(c.fields, lam)
        f.prd1 *= lam
        prd1 = oracle.ref_graddin(c.grid, f)
        g = c.grid
        arrays = dict(u=f.u, v=f.v, w=f.w, rho=f.rho, prd1_in=f.prd1, prd1=prd1)
        if g.t_press == 1:
            # posdin's projection branch; SuLi calls no posdin after an IPA
            # solve (SuLi-IC/1_principal.f90:97-111)
            u, v, w, _ = oracle.ref_posdin(c.grid, f, prd1)
            arrays.update(u_out=u, v_out=v, w_out=w)
        if f.obs_lsx is not None:
            arrays.update(obs_lsx=f.obs_lsx, obs_lsy=f.obs_lsy, obs_lsz=f.obs_lsz)
        if g.dzv is not None:
            arrays.update(dzv=g.dzv, dzzv=g.dzzv)
        meta = dict(nx=g.nx, ny=g.ny, nz=g.nz, dx=g.dx, dy=g.dy, dz=g.dz, dt=g.dt, ccx0=g.ccx0,
                    ccy0=g.ccy0, t_press=g.t_press, rho_f1=g.rho_f1, obst_t=g.obst_t,
                    ibm_t=g.ibm_t, lam=lam, seed=c.seed)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, meta=json.dumps(meta), **arrays)
        with open(path, "rb") as fh:
            manifest[f"{name}.npz"] = hashlib.sha256(fh.read()).hexdigest()
        print(name, "lam=%.3e" % lam, "max|p|=%.3e" % np.abs(prd1).max())
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as fh:
        json.dump(manifest, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
