"""Shared helpers for the test suite (fixtures, comparison masks)."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from suli_amd import cases

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    """(grid, fields, expected) from a fixture written by make_golden.py."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    meta = json.loads(str(z["meta"]))
    g = cases.Grid(meta["nx"], meta["ny"], meta["nz"], meta["dx"], meta["dy"], meta["dz"],
                   dt=meta["dt"], ccx0=meta["ccx0"], ccy0=meta["ccy0"], t_press=meta["t_press"],
                   rho_f1=meta["rho_f1"], obst_t=meta["obst_t"], ibm_t=meta["ibm_t"])
    if "dzv" in z:
        g.dzv = np.array(z["dzv"])
        g.dzzv = np.array(z["dzzv"])
    f = cases.Fields(u=np.asfortranarray(z["u"]), v=np.asfortranarray(z["v"]),
                     w=np.asfortranarray(z["w"]), rho=np.asfortranarray(z["rho"]),
                     prd1=np.asfortranarray(z["prd1_in"]))
    if "obs_lsx" in z:
        f.obs_lsx = np.asfortranarray(z["obs_lsx"])
        f.obs_lsy = np.asfortranarray(z["obs_lsy"])
        f.obs_lsz = np.asfortranarray(z["obs_lsz"])
    exp = {k: np.asfortranarray(z[k]) for k in ("prd1", "u_out", "v_out", "w_out") if k in z}
    return g, f, exp


def face_mask(shape):
    """Halo-box cells that feed the 7-point stencil: the interior plus the six
    face ghost layers, without edge and corner ghosts (those race in the
    reference's multi-threaded ghost fills, SuLi-IC/9_grad.f90:188-239)."""
    nx, ny, nz = (s - 2 for s in shape)
    I, J, K = np.meshgrid(np.arange(nx + 2), np.arange(ny + 2), np.arange(nz + 2), indexing="ij")
    ghost = ((I == 0) | (I == nx + 1)).astype(int) + ((J == 0) | (J == ny + 1)) + \
        ((K == 0) | (K == nz + 1))
    return ghost <= 1


def rel_diff(a, b, mask=None):
    if mask is not None:
        a, b = a[mask], b[mask]
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def weighted_mean_free(p, d):
    """p minus its d-weighted interior mean (the all-Neumann null space)."""
    pi = p[1:-1, 1:-1, 1:-1]
    return p - (d * pi).sum() / d.sum()


def it_layer_state(g, f, seed):
    """SuLi-IT state for posdin's layer update: cell thicknesses dz
    (0:nx+1,0:ny+1,0:nz+1) about the layers' own, eta0 (0:nx+1,0:ny+1) and
    a pressure of the size a solve leaves (the halo box)."""
    rng = np.random.default_rng(seed)
    dz0 = float(np.median(f.dzz))
    dz = np.asfortranarray(dz0 * (1.0 + 0.1 * rng.uniform(-1, 1, (g.nx + 2, g.ny + 2, g.nz + 2))))
    eta0 = np.asfortranarray(rng.uniform(-0.01, 0.01, (g.nx + 2, g.ny + 2)))
    prd1 = np.asfortranarray(rng.uniform(-0.05, 0.05, (g.nx + 2, g.ny + 2, g.nz + 2)))
    return dz, eta0, prd1
