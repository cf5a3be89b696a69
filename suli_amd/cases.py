"""Seeded synthetic pressure-step inputs in SuLi's own array layouts.

Every array is a numpy float64 array in Fortran order whose index ``[i, j, k]``
is the reference's ``(i, j, k)`` shifted by the lower bound, exactly as the
module arrays are declared (SuLi-IC/0_modules.f90:170-172,188-189):

    u    (0:nx+2, 0:ny+1, 0:nz+1)   -> shape (nx+3, ny+2, nz+2)
    v    (0:nx+1, 0:ny+2, 0:nz+1)   -> shape (nx+2, ny+3, nz+2)
    w    (0:nx+1, 0:ny+1, 0:nz+2)   -> shape (nx+2, ny+2, nz+3)
    prd1 (0:nx+1, 0:ny+1, 0:nz+1)   -> shape (nx+2, ny+2, nz+2)
    rho  (1:nx, 1:ny, 1:nz)         -> shape (nx, ny, nz)
    obs_lsx (1:nx+1, 1:ny, 1:nz) ... (SuLi-IC/0_modules.f90:135-137)
    dz, dzz (0:nz+1)                 (SuLi-ICdz/0_modules.f90:72-73)

Velocities are drawn i.i.d. on interior faces only; wall-normal boundary
faces are zero, as SuLi's contorno(2) leaves them (SuLi-IC/3_contorno.f90:
395-468), so the right-hand side is compatible with the all-Neumann operator.
With x periodic the two x boundary faces are the same face and share one
value.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np


@dataclasses.dataclass
class Grid:
    """Scalars graddin/posdin read from SuLi's modules (disc, cond, ls_param)."""

    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    dt: float = 1e-3
    ccx0: int = 1          # 0 = periodic in x (SuLi-IC/0_modules.f90:70)
    ccy0: int = 1          # 0 = periodic in y
    t_press: int = 1       # 1 = projection, 2 = IPA with constant rho_f1
    rho_f1: float = 1.2
    obst_t: int = 0
    ibm_t: int = 0
    dzv: Optional[np.ndarray] = None    # dz(0:nz+1) for the stretched (ICdz) grid
    dzzv: Optional[np.ndarray] = None   # dzz(0:nz+1)
    top_bc: int = 0        # 1: p = 0 on the top face (not a reference mode)

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def stretched(self) -> bool:
        return self.dzv is not None

    @property
    def ibm_mask(self) -> bool:
        return self.obst_t != 0 and self.ibm_t == 2


@dataclasses.dataclass
class Fields:
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    rho: np.ndarray
    prd1: np.ndarray
    obs_lsx: Optional[np.ndarray] = None
    obs_lsy: Optional[np.ndarray] = None
    obs_lsz: Optional[np.ndarray] = None
    # solid-cell mask (uint8, (nx, ny, nz), Fortran order; not a reference
    # mode): faces touching a solid cell carry no flux
    solid: Optional[np.ndarray] = None
    # Heaviside field hs (nx, ny, nz), read by posdin's hydrostatic branch
    hs: Optional[np.ndarray] = None

    def copy(self) -> "Fields":
        return Fields(*(None if a is None else a.copy(order="F") for a in dataclasses.astuple(self)))


@dataclasses.dataclass
class Case:
    name: str
    grid: Grid
    fields: Fields
    seed: int


def zeros_fields(g: Grid) -> Fields:
    nx, ny, nz = g.nx, g.ny, g.nz
    z = lambda *s: np.zeros(s, dtype=np.float64, order="F")
    return Fields(u=z(nx + 3, ny + 2, nz + 2), v=z(nx + 2, ny + 3, nz + 2),
                  w=z(nx + 2, ny + 2, nz + 3), rho=z(nx, ny, nz),
                  prd1=z(nx + 2, ny + 2, nz + 2))


def random_velocity(g: Grid, f: Fields, amp: float, rng: np.random.Generator) -> None:
    """u, v, w ~ U(-amp, amp) on interior faces, zero wall-normal walls."""
    nx, ny, nz = g.nx, g.ny, g.nz
    # u faces i = 1..nx+1 (Fortran); interior faces i = 2..nx
    if g.ccx0 == 0:
        f.u[1:nx + 1, 1:ny + 1, 1:nz + 1] = rng.uniform(-amp, amp, (nx, ny, nz))
        f.u[nx + 1, 1:ny + 1, 1:nz + 1] = f.u[1, 1:ny + 1, 1:nz + 1]
    else:
        f.u[2:nx + 1, 1:ny + 1, 1:nz + 1] = rng.uniform(-amp, amp, (nx - 1, ny, nz))
    if g.ccy0 == 0:
        f.v[1:nx + 1, 1:ny + 1, 1:nz + 1] = rng.uniform(-amp, amp, (nx, ny, nz))
        f.v[1:nx + 1, ny + 1, 1:nz + 1] = f.v[1:nx + 1, 1, 1:nz + 1]
    else:
        f.v[1:nx + 1, 2:ny + 1, 1:nz + 1] = rng.uniform(-amp, amp, (nx, ny - 1, nz))
    f.w[1:nx + 1, 1:ny + 1, 2:nz + 1] = rng.uniform(-amp, amp, (nx, ny, nz - 1))


def heaviside_rho(ls: np.ndarray, eps: float, rho1: float, rho2: float) -> np.ndarray:
    """rho = rho1 (1-H) + rho2 H, H the smoothed Heaviside of half-width eps
    (SuLi-IC/14_level_set.f90:555-587; ls > 0 is the rho2 phase)."""
    pi = np.pi
    hs = np.where(ls < -eps, 0.0, np.where(ls > eps, 1.0,
                                           0.5 * (1.0 + ls / eps + 1.0 / pi * np.sin(pi * ls / eps))))
    return rho1 * (1.0 - hs) + rho2 * hs


def cell_centres(n: int, d: float) -> np.ndarray:
    return (np.arange(1, n + 1) - 0.5) * d


def sloshing_level_set(g: Grid, height: float, amp: float) -> np.ndarray:
    """Two-phase tank: phi = z - height - amp sin(2 pi x / Lx), water below.
    Returns the level set ls = -phi (ls > 0 is water), (nx, ny, nz)."""
    lx = g.nx * g.dx
    x = cell_centres(g.nx, g.dx)
    if g.stretched:
        dzc = g.dzv[1:g.nz + 1]
        z = np.cumsum(dzc) - 0.5 * dzc
    else:
        z = cell_centres(g.nz, g.dz)
    phi = z[None, None, :] - height - amp * np.sin(2.0 * np.pi * x[:, None, None] / lx)
    return np.asfortranarray(-np.broadcast_to(phi, (g.nx, g.ny, g.nz)))


def sloshing_rho(g: Grid, height: float, amp: float, alpha1: float,
                 rho_air: float, rho_water: float) -> np.ndarray:
    """Density of the two-phase tank (sloshing_level_set) through the
    smoothed Heaviside of half-width alpha1 max(dx, dy, dz)."""
    eps = alpha1 * max(g.dx, g.dy, g.dz)
    return np.asfortranarray(heaviside_rho(sloshing_level_set(g, height, amp), eps, rho_air, rho_water))


def config1(seed: int = 1, nx: int = 64, ny: int = 64, nz: int = 32,
            top_dirichlet: bool = False) -> Case:
    """Single-phase Poisson: 64x64x32, dx=dy=dz=0.01 m, rho=1000,
    u,v,w ~ U(-0.1, 0.1), dt = 1e-3.  All walls Neumann (the reference BC),
    or with top_dirichlet p = 0 on the top face (an extension), where w is
    then free (drawn like the interior faces)."""
    g = Grid(nx, ny, nz, 0.01, 0.01, 0.01, dt=1e-3, rho_f1=1000.0,
             top_bc=1 if top_dirichlet else 0)
    f = zeros_fields(g)
    rng = np.random.default_rng(seed)
    random_velocity(g, f, 0.1, rng)
    if top_dirichlet:
        f.w[1:nx + 1, 1:ny + 1, nz + 1] = rng.uniform(-0.1, 0.1, (nx, ny))
    f.rho[...] = 1000.0
    return Case(f"cfg1{'-dirichlet' if top_dirichlet else ''}-{nx}x{ny}x{nz}", g, f, seed)


def config2(seed: int = 1, nx: int = 256, ny: int = 128, nz: int = 96,
            dx: float = 5e-3, rho_air: float = 1.2, rho_water: float = 1000.0) -> Case:
    """Two-phase sloshing tank: 256x128x96, dx=5e-3, phi = z - 0.25 -
    0.02 sin(2 pi x/Lx), rho 1000/1.2 via a smoothed Heaviside of half-width
    1.5 dx, u,v,w ~ U(-0.05, 0.05)."""
    g = Grid(nx, ny, nz, dx, dx, dx, dt=1e-3, rho_f1=rho_air)
    f = zeros_fields(g)
    rng = np.random.default_rng(seed)
    random_velocity(g, f, 0.05, rng)
    height = 0.25 * (nz * dx) / 0.48
    f.rho[...] = sloshing_rho(g, height, 0.02 * (nz * dx) / 0.48, 1.5, rho_air, rho_water)
    return Case(f"cfg2-{nx}x{ny}x{nz}", g, f, seed)


def cylinder_cells(g: Grid, radius: float) -> np.ndarray:
    """Solid-cell mask (uint8, Fortran order) of a full-height vertical
    cylinder at the domain centre: cells whose centre lies inside."""
    xc, yc = 0.5 * g.nx * g.dx, 0.5 * g.ny * g.dy
    x = cell_centres(g.nx, g.dx)
    y = cell_centres(g.ny, g.dy)
    inside = (x[:, None] - xc) ** 2 + (y[None, :] - yc) ** 2 < radius * radius
    return np.asfortranarray(np.repeat(inside[:, :, None], g.nz, axis=2).astype(np.uint8))


def config3(seed: int = 1, nx: int = 384, ny: int = 192, nz: int = 128,
            dx: float = 5e-3) -> Case:
    """Stretched vertical grid with an immersed solid: dz finest at the free
    surface z = 0.6 H and growing away from it by at most 1.05 per cell
    (tanh profile), H = nz dx; water (rho 1000) below the surface, air (1.2)
    above, blended over 1.5 max(dx, dy, mean dz); a full-height cylinder of
    radius 0.1 Ly at the centre masked out (faces touching it carry no flux);
    seeded U(-0.05, 0.05) velocities.  Not a reference mode (the reference's
    IBM masks only the right-hand side)."""
    rng = np.random.default_rng(seed)
    H = nz * dx
    g = Grid(nx, ny, nz, dx, dx, dx, dt=1e-3, rho_f1=1.2, ccx0=1, ccy0=1)
    g.dzv = stretched_dz(nz, H, 0.6, 1.05)
    g.dzzv = dzz_from_dz(g.dzv)
    f = zeros_fields(g)
    random_velocity(g, f, 0.05, rng)
    f.rho[...] = sloshing_rho(g, 0.6 * H, 0.0, 1.5, 1.2, 1000.0)
    f.solid = cylinder_cells(g, 0.1 * ny * dx)
    return Case(f"cfg3-{nx}x{ny}x{nz}", g, f, seed)


def stretched_dz(nz: int, height: float, z_focus_frac: float = 0.6,
                 ratio: float = 1.05) -> np.ndarray:
    """dz(0:nz+1) for a vertical grid that is finest at z = z_focus_frac*H
    and grows away from it by at most `ratio` per cell (tanh profile).
    Ghost spacings copy the adjacent cells."""
    # spacing ~ 1 + A tanh^2((z - zf)/L): smooth, minimum at zf
    k = np.arange(nz) + 0.5
    kf = z_focus_frac * nz
    s = np.tanh((k - kf) / (0.25 * nz)) ** 2
    w = 1.0 + 2.0 * s
    # clamp adjacent growth to `ratio`
    for i in range(1, nz):
        w[i] = min(max(w[i], w[i - 1] / ratio), w[i - 1] * ratio)
    dzc = w * (height / w.sum())
    dz = np.empty(nz + 2)
    dz[1:nz + 1] = dzc
    dz[0] = dzc[0]
    dz[nz + 1] = dzc[-1]
    return dz


def dzz_from_dz(dz: np.ndarray) -> np.ndarray:
    """dzz(k) = (dz(k)+dz(k-1))/2, ghosts copied (SuLi-ICdz/2_iniciais.f90:362-366)."""
    nz = dz.size - 2
    dzz = np.empty_like(dz)
    dzz[1:nz + 2] = (dz[1:nz + 2] + dz[0:nz + 1]) * 0.5
    dzz[0] = dzz[1]
    dzz[nz + 1] = dzz[nz]
    return dzz


def random_case(nx: int, ny: int, nz: int, seed: int, *, dx: float = 5e-3,
                ratio: float = 1000.0, periodic_x: bool = False,
                periodic_y: bool = False, surface: Optional[float] = None,
                stretched: bool = False, amp: float = 0.05, ibm: bool = False,
                t_press: int = 1) -> Case:
    """A two-phase case of any size (used for tests and batched workloads).
    surface: free-surface height as a fraction of H (None: sinusoidal 0.52H)."""
    rng = np.random.default_rng(seed)
    g = Grid(nx, ny, nz, dx, dx, dx, dt=1e-3, rho_f1=1000.0 / ratio if ratio else 1.2,
             ccx0=0 if periodic_x else 1, ccy0=0 if periodic_y else 1, t_press=t_press)
    if stretched:
        g.dzv = stretched_dz(nz, nz * dx)
        g.dzzv = dzz_from_dz(g.dzv)
    f = zeros_fields(g)
    random_velocity(g, f, amp, rng)
    H = nz * dx
    if surface is None:
        f.rho[...] = sloshing_rho(g, 0.52 * H, 0.04 * H, 1.5, 1000.0 / ratio, 1000.0)
    else:
        f.rho[...] = sloshing_rho(g, surface * H, 0.0, 1.5, 1000.0 / ratio, 1000.0)
    if ibm:
        g.obst_t, g.ibm_t = 1, 2
        cyl = cylinder_level_sets(g, 0.2 * ny * dx)
        f.obs_lsx, f.obs_lsy, f.obs_lsz = cyl
    return Case(f"rand-{nx}x{ny}x{nz}-s{seed}", g, f, seed)


def make_stretched_compatible(c: Case) -> None:
    """On SuLi-ICdz grids the right-hand side divides w differences by dzz(k)
    (SuLi-ICdz/9_grad.f90:176) while the operator is built with dz(k)^2
    (:141), so the all-Neumann system is solvable only when, per column,
    sum_k w(k) (1/dzz(k) - 1/dzz(k-1)) = 0 over interior faces k = 2..nz.
    Projects w onto that set (in place)."""
    g, f = c.grid, c.fields
    nz = g.nz
    coef = np.zeros(nz + 3)
    kk = np.arange(2, nz + 1)
    coef[kk] = 1.0 / g.dzzv[kk] - 1.0 / g.dzzv[kk - 1]
    wi = f.w[1:g.nx + 1, 1:g.ny + 1, :]
    dot = np.tensordot(wi, coef, axes=([2], [0]))
    wi -= dot[:, :, None] * coef[None, None, :] / (coef @ coef)
    f.w[1:g.nx + 1, 1:g.ny + 1, :] = wi


def cylinder_level_sets(g: Grid, radius: float):
    """Signed distance (>= 0 inside) of a full-height vertical cylinder at the
    domain centre, sampled at x, y and z faces, in the obs_ls* layouts."""
    lx, ly = g.nx * g.dx, g.ny * g.dy
    xc, yc = 0.5 * lx, 0.5 * ly

    def sd(x, y, nzf):
        r = np.sqrt((x[:, None] - xc) ** 2 + (y[None, :] - yc) ** 2)
        return np.asfortranarray(np.repeat((radius - r)[:, :, None], nzf, axis=2))

    xf = np.arange(g.nx + 1) * g.dx
    yf = np.arange(g.ny + 1) * g.dy
    xcn = cell_centres(g.nx, g.dx)
    ycn = cell_centres(g.ny, g.dy)
    return sd(xf, ycn, g.nz), sd(xcn, yf, g.nz), sd(xcn, ycn, g.nz + 1)


def scale_velocity(f: Fields, lam: float) -> Fields:
    out = f.copy()
    out.u *= lam
    out.v *= lam
    out.w *= lam
    return out


def log_uniform_sizes(count: int, seed: int, lo: int = 32, hi: int = 128, mult: int = 8):
    """Per-dimension log-uniform sizes rounded to multiples of `mult`."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        dims = []
        for _ in range(3):
            n = math.exp(rng.uniform(math.log(lo), math.log(hi)))
            dims.append(int(max(lo, min(hi, mult * round(n / mult)))))
        out.append(tuple(dims))
    return out


def batch_cases(seed: int = 1, count: int = 64, dx: float = 5e-3):
    """Uneven batch: nx, ny, nz each log-uniform in [32, 128] rounded to
    multiples of 8; a flat free surface at a height drawn from U(0.3, 0.7) H;
    density ratio drawn from {1, 10, 1000} (water 1000); U(-0.05, 0.05)
    velocities; walls.  Seeds seed*1000 + i."""
    rng = np.random.default_rng(seed)
    sizes = log_uniform_sizes(count, seed, 32, 128, 8)
    out = []
    for i, (nx, ny, nz) in enumerate(sizes):
        ratio = float(rng.choice([1.0, 10.0, 1000.0]))
        h = float(rng.uniform(0.3, 0.7))
        c = random_case(nx, ny, nz, seed=seed * 1000 + i, dx=dx, ratio=ratio, surface=h)
        out.append(c)
    return out


def serving_cases(seed: int = 1, count: int = 32, dx: float = 5e-3):
    """Request stream: nx log-uniform in [32, 256], ny and nz in [32, 128]
    (multiples of 8), two-phase sinusoidal-surface tank fields as the
    256x128x96 workload, seeded per request (seed*1000 + i)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        dims = []
        for lo, hi in ((32, 256), (32, 128), (32, 128)):
            n = math.exp(rng.uniform(math.log(lo), math.log(hi)))
            dims.append(int(max(lo, min(hi, 8 * round(n / 8)))))
        nx, ny, nz = dims
        out.append(config2(seed=seed * 1000 + i, nx=nx, ny=ny, nz=nz))
    return out


# --------------------------------------------------------------------------
# SuLi-IT (non-hydrostatic, Casulli's semi-implicit scheme): the pressure
# system of SuLi-IT/9_grad.f90:10-263.


@dataclasses.dataclass
class ITGrid:
    """Scalars of SuLi-IT's pressure solve (modules disc, cond, parametros)."""

    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dt: float
    tetah: float = 0.6          # implicitness (SuLi-IT/0_modules.f90:150)
    gz: float = 9.80665
    ccx0: int = 2               # 0 = periodic in x
    ccy0: int = 2

    @property
    def ncell(self) -> int:
        return self.nx * self.ny * self.nz


@dataclasses.dataclass
class ITFields:
    """SuLi-IT arrays in its layouts: u, dzx (0:nx+2,0:ny+1,0:nz+1); v, dzy
    (0:nx+1,0:ny+2,0:nz+1); w, dzz (0:nx+1,0:ny+1,0:nz+2); eta1
    (0:nx+1,0:ny+1); deltah (nx,ny); prd1 (0:nx+1,0:ny+1,0:nz+1)."""

    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    dzx: np.ndarray
    dzy: np.ndarray
    dzz: np.ndarray
    eta1: np.ndarray
    deltah: np.ndarray
    prd1: np.ndarray

    def copy(self) -> "ITFields":
        return ITFields(*(a.copy(order="F") for a in dataclasses.astuple(self)))


def it_case(nx: int, ny: int, nz: int, seed: int, *, dx: float = None, dy: float = None,
            dt: float = None, dz: float = None, periodic_x: bool = False,
            periodic_y: bool = False) -> "tuple[ITGrid, ITFields]":
    """A seeded SuLi-IT state: layers of about dz with +-20 % variation per
    face, velocities O(0.1), a free surface eta1 O(1e-2) and its deltah."""
    rng = np.random.default_rng(seed)
    pi = np.pi
    dx = 2.0 * pi / 50.0 if dx is None else dx
    dy = 2.0 * pi / 50.0 if dy is None else dy
    dt = 0.001 * pi if dt is None else dt
    dz = pi / 25.0 if dz is None else dz
    g = ITGrid(nx, ny, nz, dx, dy, dt, ccx0=0 if periodic_x else 2, ccy0=0 if periodic_y else 2)
    F = lambda *s: np.asfortranarray(np.zeros(s))
    u, dzx = F(nx + 3, ny + 2, nz + 2), F(nx + 3, ny + 2, nz + 2)
    v, dzy = F(nx + 2, ny + 3, nz + 2), F(nx + 2, ny + 3, nz + 2)
    w, dzz = F(nx + 2, ny + 2, nz + 3), F(nx + 2, ny + 2, nz + 3)
    for a in (dzx, dzy, dzz):
        a[...] = dz * (1.0 + 0.2 * rng.uniform(-1, 1, a.shape))
    u[1:nx + 2, 1:ny + 1, 1:nz + 1] = rng.uniform(-0.1, 0.1, (nx + 1, ny, nz))
    v[1:nx + 1, 1:ny + 2, 1:nz + 1] = rng.uniform(-0.1, 0.1, (nx, ny + 1, nz))
    w[1:nx + 1, 1:ny + 1, 1:nz + 2] = rng.uniform(-0.02, 0.02, (nx, ny, nz + 1))
    eta1 = np.asfortranarray(rng.uniform(-0.01, 0.01, (nx + 2, ny + 2)))
    deltah = np.asfortranarray(rng.uniform(-0.01, 0.01, (nx, ny)))
    prd1 = F(nx + 2, ny + 2, nz + 2)
    return g, ITFields(u, v, w, dzx, dzy, dzz, eta1, deltah, prd1)
