// Once-per-solve kernels: face densities and coefficients, diagonal and
// right-hand side, symmetric Jacobi scaling, initial residual, un-scaling of
// the solution back to prd1, and the posdin velocity correction.
//
// This file is compiled with -ffp-contract=off: every expression below is
// evaluated in the reference's operation order with no fused multiply-add,
// so the operator and the correction are bit-identical to SuLi's FP64
// arithmetic (SuLi-IC/9_grad.f90:93-301, SuLi-IC/10_pos.f90:17-46,
// SuLi-IC/8_derivadas.f90:305-481).
#include <algorithm>

#include "ps_device.h"

namespace sps {
namespace {

__device__ __forceinline__ long long cidx(const Geo &g, int i, int j, int k)
{
    return ((long long)k * g.ny + j) * g.nx + i;
}

// Face values of a cell field, interpx_cf / interpy_cf / interpz_cf
// (SuLi-IC/8_derivadas.f90:305-357, :381-418, :443-481; dz-weighted z
// faces: SuLi-ICdz/8_derivadas.f90:450-489).  Face index f is 0-based:
// face f of a direction lies between cells f-1 and f.  Boundary faces copy
// the adjacent interior face, except x-periodic where both take the mean of
// the first and last cell.  y and z have no periodic branch.  Used for rho
// and, in posdin's hydrostatic branch, for the Heaviside field hs.
__device__ __forceinline__ double face_x(const Geo &g, const double *field, int f, int j, int k)
{
    const double *r = field + ((long long)k * g.ny + j) * g.nx;
    const int nx = g.nx;
    // the two cells as selects, not branches, so the loads issue together
    // with the caller's others (a sum of two is the same either way round)
    const bool inner = f >= 1 && f <= nx - 1;
    const int a = inner ? f : (g.px ? 0 : (f == 0 ? 1 : nx - 1));
    const int b = inner || !g.px ? a - 1 : nx - 1;
    return (r[a] + r[b]) * 0.5;
}

__device__ __forceinline__ double face_y(const Geo &g, const double *field, int i, int f, int k)
{
    const double *r = field + (long long)k * g.ny * g.nx + i;
    const int ny = g.ny;
    const int a = f == 0 ? 1 : (f == ny ? ny - 1 : f);
    return (r[(long long)a * g.nx] + r[(long long)(a - 1) * g.nx]) * 0.5;
}

// z face f takes cells a - 1 and a (0-based), a = f clamped to 1 .. nz - 1,
// r0 and r1 their values
__device__ __forceinline__ int face_z_upper(const Geo &g, int f) { return f == 0 ? 1 : (f == g.nz ? g.nz - 1 : f); }
// with weights d1 = dz(a+1), d0 = dz(a) (Fortran face a+1 between cells a
// and a+1, 1-based)
__device__ __forceinline__ double face_z_weighted(double r1, double r0, double d1, double d0)
{
    return (r1 * d1 + r0 * d0) / (d1 + d0);
}
__device__ __forceinline__ double face_z_of(double r1, double r0, const double *dz, int a)
{
    if (dz) return face_z_weighted(r1, r0, dz[a + 1], dz[a]);
    return (r1 + r0) * 0.5;
}
__device__ __forceinline__ double face_z(const Geo &g, const double *field, const double *dz, int i, int j, int f)
{
    const long long plane = (long long)g.nx * g.ny;
    const double *r = field + (long long)j * g.nx + i;
    const int a = face_z_upper(g, f);
    const double r1 = r[a * plane], r0 = r[(a - 1) * plane];
    return face_z_of(r1, r0, dz, a);
}

__device__ __forceinline__ double rho_face_x(const Geo &g, const SetupIn &in, int f, int j, int k)
{
    return face_x(g, in.rho, f, j, k);
}
__device__ __forceinline__ double rho_face_y(const Geo &g, const SetupIn &in, int i, int f, int k)
{
    return face_y(g, in.rho, i, f, k);
}
__device__ __forceinline__ double rho_face_z(const Geo &g, const SetupIn &in, int i, int j, int f)
{
    return face_z(g, in.rho, in.dz, i, j, f);
}

// Solid-cell mask (an extension with no reference counterpart): a face
// with a solid cell on either side (ghost cells by the mirror/periodic rule)
// carries no flux: its coefficient is 0, so it drops out of the diagonal as
// well, and its velocity counts as 0 in the divergence; posdin leaves it
// alone.  A solid cell's own row becomes the identity with zero right-hand
// side (d = 1, b = 0, e = 0).
__device__ __forceinline__ bool solid_cell(const Geo &g, const SetupIn &in, int i, int j, int k)
{
    return in.solid && in.solid[cidx(g, i, j, k)] != 0;
}
// (both cells' bytes are loaded and combined with |, not ||, so the two
// loads are issued together instead of the second waiting on the first)
__device__ __forceinline__ bool closed_x(const Geo &g, const SetupIn &in, int f, int j, int k)
{
    if (!in.solid) return false;
    return (in.solid[cidx(g, box_to_cell(f + 1, g.nx, g.px), j, k)] | in.solid[cidx(g, box_to_cell(f, g.nx, g.px), j, k)]) !=
           0;
}
__device__ __forceinline__ bool closed_y(const Geo &g, const SetupIn &in, int i, int f, int k)
{
    if (!in.solid) return false;
    return (in.solid[cidx(g, i, box_to_cell(f + 1, g.ny, g.py), k)] | in.solid[cidx(g, i, box_to_cell(f, g.ny, g.py), k)]) !=
           0;
}
__device__ __forceinline__ bool closed_z(const Geo &g, const SetupIn &in, int i, int j, int f)
{
    if (!in.solid) return false;
    return (in.solid[cidx(g, i, j, box_to_cell(f + 1, g.nz, 0))] | in.solid[cidx(g, i, j, box_to_cell(f, g.nz, 0))]) != 0;
}

// Cell tiles of the setup and correction kernels: 64 x 4 cells of a plane
// per workgroup pass, one cell per thread, numbered x fastest, then y, then
// the planes.  Workgroup b starts at tile xcd_tile(b, G) and strides by the
// grid G, so each XCD walks contiguous runs of tiles: adjacent tiles share
// cache lines at row ends (rows of the face arrays, u, v, w and the prd1
// box do not start on a line) and halo rows, and read them into one L2.  A
// wave is one row of its tile, so j and k are the same on all its lanes.
__host__ __device__ inline int cell_tiles(const Geo &g)
{
    return ((g.nx + 63) / 64) * ((g.ny + 3) / 4) * g.nz;   // at most a quarter of the cells
}

template <class F>
__device__ __forceinline__ void for_cells(const Geo &g, F &&f)
{
    const int tx = (g.nx + 63) / 64, ty = (g.ny + 3) / 4, ntile = cell_tiles(g);
    const int G = gridDim.x;
    const int row = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int t = xcd_tile(blockIdx.x, G); t < ntile; t += G) {
        const int bx = t % tx, r = t / tx;
        const int i = bx * 64 + (threadIdx.x & 63), j = (r % ty) * 4 + row, k = r / ty;
        if (i < g.nx && j < g.ny) f(i, j, k);
    }
}

// s = dt/(delta^2 rho_face) (9_grad.f90:118-148; SuLi-ICdz/9_grad.f90:141)
// of x face f of row (j, k), y face f of column (i, k), z face f of (i, j).
__device__ __forceinline__ double coef_x(const Geo &g, const SetupIn &in, int f, int j, int k)
{
    // SuLi-IT: tetah dt/dx^2 dzx (SuLi-IT/9_grad.f90:44-50)
    if (in.it) return in.aux1 * in.dzx[(f + 1) + (long long)(g.nx + 3) * ((j + 1) + (long long)(g.ny + 2) * (k + 1))];
    const double rf = in.ipa ? in.rho_f1 : rho_face_x(g, in, f, j, k);
    return closed_x(g, in, f, j, k) ? 0.0 : in.aux1 / rf;
}
__device__ __forceinline__ double coef_y(const Geo &g, const SetupIn &in, int i, int f, int k)
{
    if (in.it) return in.aux2 * in.dzy[(i + 1) + (long long)(g.nx + 2) * ((f + 1) + (long long)(g.ny + 3) * (k + 1))];
    const double rf = in.ipa ? in.rho_f1 : rho_face_y(g, in, i, f, k);
    return closed_y(g, in, i, f, k) ? 0.0 : in.aux2 / rf;
}
__device__ __forceinline__ double coef_z(const Geo &g, const SetupIn &in, int i, int j, int f)
{
    if (in.it) {
        // tetah dt/dzz; no face above the top plane (:51-60)
        return f == g.nz ? 0.
                         : in.tetah * in.dt /
                               in.dz3[(i + 1) + (long long)(g.nx + 2) * ((j + 1) + (long long)(g.ny + 2) * (f + 1))];
    }
    const double rf = in.ipa ? in.rho_f1 : rho_face_z(g, in, i, j, f);
    if (closed_z(g, in, i, j, f)) return 0.0;
    if (in.dz) {
        const double dzk = in.dz[f + 1];
        return in.dt / (dzk * dzk * rf);
    }
    return in.aux3 / rf;
}

// Cell (i, j, k) writes its faces on the low side in x, y and z, and the
// cells of the last column, row and plane the face on their high side too:
// every face once.  In the usual case (not SuLi-IT, no IPA constant) the
// four face densities a cell needs are loaded together before any branch
// or store; the values are those of coef_x/y/z.
__global__ __launch_bounds__(kThreads) void k_face_coefs(Geo g, SetupIn in, double *sx, double *sy, double *sz)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    for_cells(g, [&](int i, int j, int k) {
        const long long fx = ((long long)k * ny + j) * (nx + 1) + i;
        const long long fy = ((long long)k * (ny + 1) + j) * nx + i;
        const long long fz = ((long long)k * ny + j) * nx + i;
        double x0, x1, y0, z0;
        if (in.it || in.ipa) {
            x0 = coef_x(g, in, i, j, k);
            x1 = coef_x(g, in, i + 1, j, k);
            y0 = coef_y(g, in, i, j, k);
            z0 = coef_z(g, in, i, j, k);
        } else {
            // face i + 1 is in range for every lane; only the last column
            // stores it
            const double rx0 = rho_face_x(g, in, i, j, k), rx1 = rho_face_x(g, in, i + 1, j, k);
            const double ry0 = rho_face_y(g, in, i, j, k), rz0 = rho_face_z(g, in, i, j, k);
            const double dzk = in.dz ? in.dz[k + 1] : 0.0;
            x0 = in.aux1 / rx0;
            x1 = in.aux1 / rx1;
            y0 = in.aux2 / ry0;
            z0 = in.dz ? in.dt / (dzk * dzk * rz0) : in.aux3 / rz0;
            if (in.solid) {
                if (closed_x(g, in, i, j, k)) x0 = 0.0;
                if (closed_x(g, in, i + 1, j, k)) x1 = 0.0;
                if (closed_y(g, in, i, j, k)) y0 = 0.0;
                if (closed_z(g, in, i, j, k)) z0 = 0.0;
            }
        }
        sx[fx] = x0;
        if (i == nx - 1) sx[fx + 1] = x1;
        sy[fy] = y0;
        sz[fz] = z0;
        if (j == ny - 1) sy[fy + nx] = coef_y(g, in, i, ny, k);
        if (k == nz - 1) sz[fz + (long long)nx * ny] = coef_z(g, in, i, j, nz);
    });
}

// d = sum of the six face coefficients (9_grad.f90:152-160) and the scaled
// right-hand side b = q/sqrt(d) with q = -div u (9_grad.f90:163-183,297;
// SuLi-ICdz/9_grad.f90:166,176; IBM face mask 9_grad.f90:45-88).
__global__ __launch_bounds__(kThreads) void k_diag_rhs(Geo g, SetupIn in, const double *sx, const double *sy,
                                                       const double *sz, double *d, double *b, double *e)
{
    const int nx = g.nx, ny = g.ny;
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long c = (long long)blockIdx.x * kThreads + threadIdx.x; c < g.n; c += stride) {
        const int i = (int)(c % nx);
        const long long jk = c / nx;
        const int j = (int)(jk % ny), k = (int)(jk / ny);
        const long long fx = ((long long)k * ny + j) * (nx + 1) + i;
        const long long fy = ((long long)k * (ny + 1) + j) * nx + i;
        const long long fz = ((long long)k * ny + j) * nx + i;
        const long long pl = (long long)nx * ny;
        const double dd = sx[fx + 1] + sx[fx] + sy[fy + nx] + sy[fy] + sz[fz + pl] + sz[fz];
        d[c] = dd;

        const int I = i + 1, J = j + 1, K = k + 1;
        const long long iu = I + (long long)(nx + 3) * (J + (long long)(ny + 2) * K);
        const long long iv = I + (long long)(nx + 2) * (J + (long long)(ny + 3) * K);
        const long long iw = I + (long long)(nx + 2) * (J + (long long)(ny + 2) * K);
        double u0 = in.u[iu], u1 = in.u[iu + 1];
        double v0 = in.v[iv], v1 = in.v[iv + (nx + 2)];
        double w0 = in.w[iw], w1 = in.w[iw + (long long)(nx + 2) * (ny + 2)];
        if (in.obsx) {
            if (in.obsx[fx] >= 0.0) u0 = 0.0;
            if (in.obsx[fx + 1] >= 0.0) u1 = 0.0;
            if (in.obsy[fy] >= 0.0) v0 = 0.0;
            if (in.obsy[fy + nx] >= 0.0) v1 = 0.0;
            if (in.obsz[fz] >= 0.0) w0 = 0.0;
            if (in.obsz[fz + pl] >= 0.0) w1 = 0.0;
        }
        if (in.solid) {
            // closed faces carry no flux in the right-hand side either
            if (closed_x(g, in, i, j, k)) u0 = 0.0;
            if (closed_x(g, in, i + 1, j, k)) u1 = 0.0;
            if (closed_y(g, in, i, j, k)) v0 = 0.0;
            if (closed_y(g, in, i, j + 1, k)) v1 = 0.0;
            if (closed_z(g, in, i, j, k)) w0 = 0.0;
            if (closed_z(g, in, i, j, k + 1)) w1 = 0.0;
        }
        if (in.it) {
            // SuLi-IT continuity in fluxes (SuLi-IT/9_grad.f90:65-81): the
            // top plane adds the free surface to the diagonal and the rhs
            const double ax = (u0 * in.dzx[iu] - u1 * in.dzx[iu + 1]) / in.dx;
            const double ay = (v0 * in.dzy[iv] - v1 * in.dzy[iv + (nx + 2)]) / in.dy;
            double q, dt = dd;
            if (k < g.nz - 1) {
                q = ax + ay + w0 - w1;
            } else {
                q = ax + ay + w0 - (in.eta1[I + (long long)(nx + 2) * J] - in.deltah[i + (long long)nx * j]) /
                                       (in.tetah * in.dt);
                dt = dd + 1. / (in.gz * in.tetah * in.dt);
            }
            d[c] = dt;
            b[c] = q / sqrt(dt);
            e[c] = in.prd1[I + (long long)(nx + 2) * (J + (long long)(ny + 2) * K)] * sqrt(dt);
            continue;
        }
        const double dzk = in.dzz ? in.dzz[K] : in.dzu;
        const double q = (u0 - u1) / in.dx + (v0 - v1) / in.dy + (w0 - w1) / dzk;
        if (solid_cell(g, in, i, j, k)) {
            d[c] = 1.0;
            b[c] = 0.0;
            e[c] = 0.0;
        } else {
            b[c] = q / sqrt(dd);
            // the scaled unknown e = prd1 sqrt(d) (9_grad.f90:242-252), as e_box0
            e[c] = in.prd1[I + (long long)(nx + 2) * (J + (long long)(ny + 2) * K)] * sqrt(dd);
        }
    }
}

// s_hat = s / sqrt(d_a d_b) with the ghost diagonal mirrored or periodic
// (9_grad.f90:188-239 ghosts, :256-288 scaling).  Also writes FP32 copies
// for the mixed-precision inner solve when requested.
__global__ __launch_bounds__(kThreads) void k_scale_faces(Geo g, const double *d, double *sx, double *sy,
                                                          double *sz, float *sx32, float *sy32, float *sz32)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long nfx = (long long)(nx + 1) * ny * nz;
    const long long nfy = (long long)nx * (ny + 1) * nz;
    const long long nfz = (long long)nx * ny * (nz + 1);
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long t = (long long)blockIdx.x * kThreads + threadIdx.x; t < nfx + nfy + nfz; t += stride) {
        if (t < nfx) {
            const int f = (int)(t % (nx + 1));
            const long long jk = t / (nx + 1);
            const int j = (int)(jk % ny), k = (int)(jk / ny);
            const double da = d[cidx(g, box_to_cell(f + 1, nx, g.px), j, k)];
            const double db = d[cidx(g, box_to_cell(f, nx, g.px), j, k)];
            const double s = sx[t] / sqrt(da * db);
            sx[t] = s;
            if (sx32) sx32[t] = (float)s;
        } else if (t < nfx + nfy) {
            const long long u = t - nfx;
            const int i = (int)(u % nx);
            const long long jk = u / nx;
            const int f = (int)(jk % (ny + 1)), k = (int)(jk / (ny + 1));
            const double da = d[cidx(g, i, box_to_cell(f + 1, ny, g.py), k)];
            const double db = d[cidx(g, i, box_to_cell(f, ny, g.py), k)];
            const double s = sy[u] / sqrt(da * db);
            sy[u] = s;
            if (sy32) sy32[u] = (float)s;
        } else {
            const long long u = t - nfx - nfy;
            const int i = (int)(u % nx);
            const long long jk = u / nx;
            const int j = (int)(jk % ny), f = (int)(jk / ny);
            const double da = d[cidx(g, i, j, box_to_cell(f + 1, nz, 0))];
            const double db = d[cidx(g, i, j, box_to_cell(f, nz, 0))];
            const double s = sz[u] / sqrt(da * db);
            sz[u] = s;
            if (sz32) sz32[u] = (float)s;
        }
    }
}

// Scaled unknown e = prd1 sqrt(d) at Fortran box index (I, J, K), ghosts of
// d taken from the interior by the mirror/periodic rule (9_grad.f90:242-252).
// Solid cells (and their ghost images) start at e = 0; with a Dirichlet
// lid the top ghost plane is minus the cell below.
__device__ __forceinline__ double e_box0(const Geo &g, const SetupIn &in, const double *d, int I, int J, int K)
{
    const int i = box_to_cell(I, g.nx, g.px), j = box_to_cell(J, g.ny, g.py), k = box_to_cell(K, g.nz, 0);
    if (solid_cell(g, in, i, j, k)) return 0.0;
    const double p = in.prd1[I + (long long)(g.nx + 2) * (J + (long long)(g.ny + 2) * K)];
    return p * sqrt(d[cidx(g, i, j, k)]);
}

__device__ __forceinline__ double e_box(const Geo &g, const SetupIn &in, const double *d, int I, int J, int K)
{
    if (g.dtop && K == g.nz + 1) return -e_box0(g, in, d, I, J, g.nz);
    return e_box0(g, in, d, I, J, K);
}

// Initial residual r = A e - b with the caller's prd1 ghost values
// (9_grad.f90:242-301).  b is replaced by b_eff = b + (A_mirror e -
// A_ghosts e): the right-hand side under which later FP64 residual
// recomputations (which use mirrored ghosts, as every CG iteration does)
// reproduce this residual.  With mirror-consistent prd1 ghosts (SuLi's
// contorno(1) leaves them so), b_eff == b bit for bit.
// Partials: r.r, b.b, r.sqrt(d), sum of d, e.sqrt(d).
__global__ __launch_bounds__(kThreads) void k_init_residual(Geo g, SetupIn in, Op<double> op, const double *d,
                                                            double *b, double *e, double *r0, Partials P)
{
    const int nx = g.nx, ny = g.ny;
    const long long pl = (long long)nx * ny;
    double acc_rr = 0.0, acc_bb = 0.0, acc_rn = 0.0, acc_dd = 0.0, acc_en = 0.0;
    for_cells(g, [&](int i, int j, int k) {
        const long long c = cidx(g, i, j, k);
        const int I = i + 1, J = j + 1, K = k + 1;
        const long long fx = ((long long)k * ny + j) * (nx + 1) + i;
        const long long fy = ((long long)k * (ny + 1) + j) * nx + i;
        const long long fz = c;
        const bool xe = i + 1 < nx, xw = i > 0, yn = j + 1 < ny, ys = j > 0, zt = k + 1 < g.nz, zb = k > 0;
        // e of the cells (written with d by k_diag_rhs); ghost positions from
        // the caller's prd1 ghosts, as the reference's e over the whole box
        const double ec = e[c];
        const double eE = xe ? e[c + 1] : e_box(g, in, d, I + 1, J, K);
        const double eW = xw ? e[c - 1] : e_box(g, in, d, I - 1, J, K);
        const double eN = yn ? e[c + nx] : e_box(g, in, d, I, J + 1, K);
        const double eS = ys ? e[c - nx] : e_box(g, in, d, I, J - 1, K);
        const double eT = zt ? e[c + pl] : e_box(g, in, d, I, J, K + 1);
        const double eB = zb ? e[c - pl] : e_box(g, in, d, I, J, K - 1);
        // the same neighbours with mirrored/periodic ghosts: the cells above
        // wherever the neighbour is a cell
        const double mE = xe ? eE : e[cidx(g, box_to_cell(I + 1, nx, g.px), j, k)];
        const double mW = xw ? eW : e[cidx(g, box_to_cell(I - 1, nx, g.px), j, k)];
        const double mN = yn ? eN : e[cidx(g, i, box_to_cell(J + 1, ny, g.py), k)];
        const double mS = ys ? eS : e[cidx(g, i, box_to_cell(J - 1, ny, g.py), k)];
        const double mT = (g.dtop && K == g.nz) ? -ec : (zt ? eT : e[cidx(g, i, j, box_to_cell(K + 1, g.nz, 0))]);
        const double mB = zb ? eB : e[cidx(g, i, j, box_to_cell(K - 1, g.nz, 0))];
        const double sE = op.sx[fx + 1], sW = op.sx[fx], sN = op.sy[fy + nx], sS = op.sy[fy];
        const double sT = op.sz[fz + pl], sB = op.sz[fz];
        const double bc = b[c];
        // null vector: sqrt(d) on fluid cells, 0 on solid ones (d and the
        // mask byte loaded together, then selected)
        const double dl = d[c];
        const double dc = solid_cell(g, in, i, j, k) ? 0.0 : dl;
        const double a_in = ec - sE * eE - sW * eW - sN * eN - sS * eS - sT * eT - sB * eB;
        const double a_mir = ec - sE * mE - sW * mW - sN * mN - sS * mS - sT * mT - sB * mB;
        const double r = a_in - bc;
        b[c] = bc + (a_mir - a_in);
        r0[c] = r;
        const double sdc = sqrt(dc);
        acc_rr += r * r;
        acc_bb += bc * bc;
        acc_rn += r * sdc;
        acc_dd += dc;
        acc_en += ec * sdc;
    });
    const double s0 = block_sum(acc_rr), s1 = block_sum(acc_bb), s2 = block_sum(acc_rn), s3 = block_sum(acc_dd);
    const double s4 = block_sum(acc_en);
    if (threadIdx.x == 0) {
        const int nb = gridDim.x;
        P.init[blockIdx.x] = s0;
        P.init[nb + blockIdx.x] = s1;
        P.init[2 * nb + blockIdx.x] = s2;
        P.init[3 * nb + blockIdx.x] = s3;
        P.init[4 * nb + blockIdx.x] = s4;
    }
}

// Every reference boundary condition leaves the operator singular: its rows
// sum to zero (zero-flux walls, periodic sides), so the scaled operator
// annihilates n = sqrt(d).  A right-hand side computed in floating point is
// consistent only to rounding, and at tight tolerances CG then drifts along
// n.  When S->project is set the n-component gamma n of the initial
// residual is removed (gamma = r.n / n.n) and b_eff adjusted to match; this
// changes the solution only by a constant (the pressure gauge).  The first
// search direction p = r (9_grad.f90:372-380) is formed by the first cg_a.
// One workgroup: the totals of k_init_residual's nb0 partials of each sum,
// taken once; the null-space component gamma and the stop threshold.
__global__ __launch_bounds__(kThreads) void k_init_totals(Partials P, int nb0, CgScalars *S)
{
    // each thread its partials t, t + kThreads, ... of all five sums in one
    // loop (unrolled, so many loads are in flight), then the block sums:
    // the additions of block_total_of, in its order
    const gptr<const double> p = glob(static_cast<const double *>(P.init));
    double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
    for (int t = threadIdx.x; t < nb0; t += kThreads)
#pragma unroll
        for (int q = 0; q < 5; ++q) a[q] += p[(long long)q * nb0 + t];
    const double rr = block_sum(a[0]), bb = block_sum(a[1]), rn = block_sum(a[2]), dd = block_sum(a[3]);
    const double en = block_sum(a[4]);
    if (threadIdx.x == 0) {
        S->gamma = S->project ? rn / dd : 0.0;
        S->bb = bb;
        S->rr0 = rr;
        S->en0 = en;
        S->dd = dd;
        S->thresh = S->rel_rule ? S->tol * S->tol * bb : S->thresh_ref;
    }
}

// r -= gamma sqrt(d) (and b += gamma sqrt(d)), the FP64 or FP32 residual
// the CG starts from, and its r.r partials (one per workgroup).
__global__ __launch_bounds__(kThreads) void k_init_project(Geo g, const uint8_t *solid, const double *d, double *b,
                                                           const double *r0, double *r64, float *r32, float *c32,
                                                           Partials P, const CgScalars *S)
{
    const double gamma = S->gamma;
    const long long stride = (long long)gridDim.x * kThreads;
    double acc = 0.0, acc32 = 0.0;
    for (long long c = (long long)blockIdx.x * kThreads + threadIdx.x; c < g.n; c += stride) {
        const double dl = d[c];
        const double sq = (solid && solid[c]) ? 0.0 : sqrt(dl);
        const double r = gamma != 0.0 ? r0[c] - gamma * sq : r0[c];
        if (gamma != 0.0) b[c] = b[c] + gamma * sq;
        acc += r * r;
        if (r64) {
            r64[c] = r;
        } else {
            const float rf = (float)r;
            r32[c] = rf;
            c32[c] = 0.0f;
            acc32 += (double)rf * (double)rf;
        }
    }
    const double s0 = block_sum(acc), s1 = block_sum(acc32);
    if (threadIdx.x == 0) {
        P.proj[blockIdx.x] = s0;
        P.proj[gridDim.x + blockIdx.x] = s1;
    }
}

// One workgroup: totals after the projection; S ready for iteration 1.
// The reference tests the stop rule before the first iteration
// (do while (|r.r| > thresh .and. cont < 10000), 9_grad.f90:398).
__global__ __launch_bounds__(kThreads) void k_init_finalize(Partials P, int np, CgScalars *S)
{
    const double rr = block_total_of(P.proj, np);
    const double rs = block_total_of(P.proj + np, np);
    if (threadIdx.x == 0) {
        const double rrc = S->mixed ? rs : rr;
        S->rr[0] = rrc;
        S->rr[1] = rrc;
        S->rr0 = rr;
        S->rr_true = rr;
        S->iters = 0;
        S->refreshes = 0;
        S->gshift = 0.0;
        S->skip_e_at = -1;
        S->next_refresh = kFirstRefresh;
        S->prev_iters = 0;
        S->prev_rr = 0.0;
        S->slope = 0.0;
        const bool conv = !(fabs(rr) > S->thresh);
        const bool done = conv || S->max_iter <= 0;
        S->done_at = done ? 0 : kNotDone;
        S->final_ = done ? 1 : 0;
        S->status = conv ? 0 : 1;
    }
}

// prd1 = e / sqrt(d) on the whole halo box; ghosts copy the interior by the
// mirror/periodic rule in x, then y, then z order (9_grad.f90:594-659).
// gshift (0 unless the null space is projected) restores the initial gauge.
__global__ __launch_bounds__(kThreads) void k_unscale(Geo g, const double *e, const double *d, double *prd1,
                                                      const CgScalars *S, const uint8_t *solid)
{
    const double gshift = S->gshift;
    const int bx = g.nx + 2, by = g.ny + 2;
    const long long nbox = (long long)bx * by * (g.nz + 2);
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long t = (long long)blockIdx.x * kThreads + threadIdx.x; t < nbox; t += stride) {
        const int I = (int)(t % bx);
        const long long jk = t / bx;
        const int J = (int)(jk % by), K = (int)(jk / by);
        const long long c = cidx(g, box_to_cell(I, g.nx, g.px), box_to_cell(J, g.ny, g.py), box_to_cell(K, g.nz, 0));
        const double p = e[c] / sqrt(d[c]);
        // solid cells (and their ghost images) are 0 exactly; a Dirichlet
        // lid's ghost plane is minus the cell below
        const double v = (solid && solid[c]) ? 0.0 : (gshift != 0.0 ? p - gshift : p);
        prd1[t] = (g.dtop && K == g.nz + 1) ? -v : v;
    }
}

// posdin, projection branch (SuLi-IC/10_pos.f90:17-46; SuLi-ICdz/10_pos.f90:41):
//   u(i,j,k) -= dt/dx (p(i)-p(i-1)) / rhox,   i = 1..nx+1, j = 1..ny, k = 1..nz
//   v(i,j,k) -= dt/dy (p(j)-p(j-1)) / rhoy,   j = 1..ny+1
//   w(i,j,k) -= dt/dz (p(k)-p(k-1)) / rhoz,   k = 1..nz+1 (dz -> dzz(k) on stretched grids)
// With_z false: the u and v faces only, the part both branches share.
//
// One cell (i, j, k) corrects its faces at its low side in x, y and z (face
// index f = i, j, k), and the cells of the last column, row and plane the
// face at their high side as well (f = nx, ny, nz): every face once, with no
// index divisions.  Each face's arithmetic is the reference's, in its order.
__device__ __forceinline__ void posdin_cell(const Geo &g, const SetupIn &in, double *u, double *v, double *w,
                                            bool with_z, int i, int j, int k, double aux1, double aux2)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long pb = nx + 2, pbb = (long long)(nx + 2) * (ny + 2);
    const long long ip = (i + 1) + pb * (j + 1) + pbb * (k + 1);   // the cell in the prd1 box
    const long long iu = (i + 1) + (long long)(nx + 3) * ((j + 1) + (long long)(ny + 2) * (k + 1));
    const long long iv = (i + 1) + (long long)(nx + 2) * ((j + 1) + (long long)(ny + 3) * (k + 1));
    const long long iw = (i + 1) + (long long)(nx + 2) * ((j + 1) + (long long)(ny + 2) * (k + 1));
    const long long wplane = (long long)(nx + 2) * (ny + 2);
    const bool hx = i == nx - 1, hy = j == ny - 1, hz = with_z && k == nz - 1;
    const double *P = in.prd1;

    // Every load before the first store: the stores could alias the inputs
    // as far as the compiler knows, so loads after a store would wait for it
    // and each face would cost the cell another round trip.
    const double dz0 = with_z ? (in.dzz ? in.dzz[k + 1] : in.dzu) : 1.0;
    const double pc = P[ip], pxm = P[ip - 1], pym = P[ip - pb];
    const double u0 = u[iu], v0 = v[iv];
    const double rx0 = rho_face_x(g, in, i, j, k), ry0 = rho_face_y(g, in, i, j, k);
    const bool cx0 = closed_x(g, in, i, j, k), cy0 = closed_y(g, in, i, j, k);
    // face i + 1 in x, stored by the last column only: in bounds for every
    // lane (the box and u have a column beyond it) and on the lines just
    // read, so every lane loads it rather than branch
    const double pxp = P[ip + 1], u1 = u[iu + 1];
    const double rx1 = rho_face_x(g, in, i + 1, j, k);
    const bool cx1 = !hx || closed_x(g, in, nx, j, k);
    double pzm = 0, w0 = 0, rz0 = 1;
    bool cz0 = true;
    if (with_z) {
        pzm = P[ip - pbb];
        w0 = w[iw];
        rz0 = rho_face_z(g, in, i, j, k);
        cz0 = closed_z(g, in, i, j, k);
    }

    // face f between box cells f and f + 1 of its direction (0-based f),
    // in the reference's operation order
    if (!cx0) u[iu] = u0 - aux1 * (pc - pxm) / rx0;
    if (!cx1) u[iu + 1] = u1 - aux1 * (pxp - pc) / rx1;
    if (!cy0) v[iv] = v0 - aux2 * (pc - pym) / ry0;
    if (!cz0) w[iw] = w0 - in.dt / dz0 * (pc - pzm) / rz0;

    // The high faces in y and z: j and k are the same on every lane of a
    // wave, so only the waves of the last row and plane come here, and
    // their loads after the stores above cost them one more round trip.
    if (hy) {
        const double pyp = P[ip + pb], v1 = v[iv + (nx + 2)];
        const double ry1 = rho_face_y(g, in, i, ny, k);
        if (!closed_y(g, in, i, ny, k)) v[iv + (nx + 2)] = v1 - aux2 * (pyp - pc) / ry1;
    }
    if (hz) {
        const double pzp = P[ip + pbb], w1 = w[iw + wplane];
        const double rz1 = rho_face_z(g, in, i, j, nz);
        const double dz1 = in.dzz ? in.dzz[nz + 1] : in.dzu;
        if (!closed_z(g, in, i, j, nz)) w[iw + wplane] = w1 - in.dt / dz1 * (pzp - pc) / rz1;
    }
}

// One thread per cell of the cell tiles (for_cells).
__global__ __launch_bounds__(kThreads) void k_posdin(Geo g, SetupIn in, double *u, double *v, double *w, bool with_z)
{
    const double aux1 = in.dt / in.dx, aux2 = in.dt / in.dy;
    for_cells(g, [&](int i, int j, int k) { posdin_cell(g, in, u, v, w, with_z, i, j, k, aux1, aux2); });
}

// Planes whose loads a column sweep (k_pressh, k_posdin_hydro_w) issues
// together before it runs its recurrence over them.
constexpr int kColumnAhead = 8;

// pressh, the hydrostatic pressure of t_press = 0 (SuLi-IC/9_grad.f90:
// 667-702; dz(k) weights: SuLi-ICdz/9_grad.f90:665-700).  One lane per
// (i, j) column walks k from the top: the top cell takes dz g rho, the one
// below the trapezoid of the top two, every other cell Simpson's rule over
// the two cells above it plus prd1(k+2).  The two values above stay in
// registers; lanes of a wavefront are consecutive i, so every k step reads
// and writes contiguous rows.  Only prd1's interior is written.
__global__ __launch_bounds__(kThreads) void k_pressh(Geo g, const double *__restrict__ rho, const double *dz, double dzu,
                                                     double gz, double *__restrict__ prd1)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long col = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (col >= (long long)nx * ny) return;
    const int i = (int)(col % nx), j = (int)(col / nx);
    const long long plane = (long long)nx * ny;
    const long long pbb = (long long)(nx + 2) * (ny + 2);
    const double *rc = rho + col;
    double *pc = prd1 + (i + 1) + (long long)(nx + 2) * (j + 1);
    double p1 = 0.0, p2 = 0.0;  // prd1 at k+1, k+2
    double r1 = 0.0, r2 = 0.0;  // rho at k+1, k+2
    // rho of kColumnAhead planes loaded together, then the recurrence over
    // them: one round trip per block of planes instead of one per plane
    for (int K0 = nz; K0 >= 1; K0 -= kColumnAhead) {
        double rb[kColumnAhead];
#pragma unroll
        for (int m = 0; m < kColumnAhead; ++m) rb[m] = K0 - m >= 1 ? rc[(long long)(K0 - m - 1) * plane] : 0.0;
#pragma unroll
        for (int m = 0; m < kColumnAhead; ++m) {
            const int K = K0 - m;
            if (K < 1) break;
            const double r0 = rb[m];
            double p;
            if (dz) {
                if (K == nz) {
                    p = dz[K] * gz * r0;
                } else if (K == nz - 1) {
                    p = gz * (r0 * dz[K] + r1 * dz[K + 1]) * 0.5 + p1;
                } else {
                    const double intz = (r0 * dz[K] + 4. * r1 * dz[K + 1] + r2 * dz[K + 2]) / 3.;
                    p = gz * intz + p2;
                }
            } else {
                if (K == nz) {
                    p = dzu * gz * r0;
                } else if (K == nz - 1) {
                    p = dzu * gz * (r0 + r1) * 0.5 + p1;
                } else {
                    const double intz = dzu * (r0 + 4. * r1 + r2) / 3.;
                    p = gz * intz + p2;
                }
            }
            pc[(long long)K * pbb] = p;
            p2 = p1;
            p1 = p;
            r2 = r1;
            r1 = r0;
        }
    }
}

// posdin's hydrostatic branch, after the shared u, v correction
// (SuLi-IC/10_pos.f90:48-57; dz(k): SuLi-ICdz/10_pos.f90:52): w(:,:,1) = 0
// over the whole bottom plane, then w(k+1) = w(k) - du dz/dx - dv dz/dy up
// each interior column from the corrected u, v.  The same lane then zeroes
// its column's w faces where the face value of hs is below 1e-4 (:84-91);
// the chain itself runs on the values before zeroing, as in the reference,
// where the zeroing is a later sweep.  Lanes over the (nx+2)(ny+2) plane;
// the ring only clears its bottom face.
template <bool DZ>
__global__ __launch_bounds__(kThreads) void k_posdin_hydro_w(Geo g, const double *u, const double *v,
                                                             const double *hs, const double *dz, double dzu,
                                                             double dx, double dy, double *w)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long t = (long long)blockIdx.x * kThreads + threadIdx.x;
    if (t >= (long long)(nx + 2) * (ny + 2)) return;
    const int I = (int)(t % (nx + 2)), J = (int)(t / (nx + 2));
    const long long wbb = (long long)(nx + 2) * (ny + 2);
    double *wc = w + t;
    wc[wbb] = 0.;  // w(:,:,1) = 0
    if (I < 1 || I > nx || J < 1 || J > ny) return;
    const long long ubb = (long long)(nx + 3) * (ny + 2), vbb = (long long)(nx + 2) * (ny + 3);
    const double *uc = u + I + (long long)(nx + 3) * J;
    const double *vc = v + I + (long long)(nx + 2) * J;
    double wk = 0.;
    if (face_z(g, hs, dz, I - 1, J - 1, 0) < 0.0001) wc[wbb] = 0.;
    // u, v and the two hs cells of the z face of kColumnAhead planes loaded
    // together (no branch between the loads), then the chain over them
    // (DZ: the stretched grid's dz(K) and the face weights too)
    const long long plane = (long long)nx * ny;
    const double *hc = hs + (long long)(J - 1) * nx + (I - 1);
    for (int K0 = 1; K0 <= nz; K0 += kColumnAhead) {
        double u0[kColumnAhead], u1[kColumnAhead], v0[kColumnAhead], v1[kColumnAhead];
        double h1[kColumnAhead], h0[kColumnAhead], dk[kColumnAhead], d1[kColumnAhead], d0[kColumnAhead];
#pragma unroll
        for (int m = 0; m < kColumnAhead; ++m) {
            const int K = K0 + m <= nz ? K0 + m : nz;
            const int a = face_z_upper(g, K);
            u1[m] = uc[(long long)K * ubb + 1];
            u0[m] = uc[(long long)K * ubb];
            v1[m] = vc[(long long)K * vbb + (nx + 2)];
            v0[m] = vc[(long long)K * vbb];
            h1[m] = hc[a * plane];
            h0[m] = hc[(a - 1) * plane];
            dk[m] = DZ ? dz[K] : dzu;
            d1[m] = DZ ? dz[a + 1] : 0.0;
            d0[m] = DZ ? dz[a] : 0.0;
        }
#pragma unroll
        for (int m = 0; m < kColumnAhead; ++m) {
            const int K = K0 + m;
            if (K > nz) break;
            const double dzk = dk[m];
            const double du = u1[m] - u0[m], dv = v1[m] - v0[m];
            wk = wk - du * dzk / dx - dv * dzk / dy;
            const double hf = DZ ? face_z_weighted(h1[m], h0[m], d1[m], d0[m]) : (h1[m] + h0[m]) * 0.5;
            wc[(long long)(K + 1) * wbb] = hf < 0.0001 ? 0. : wk;
        }
    }
}

// posdin's hydrostatic branch, last: u and v zeroed on the faces where hs
// is below 1e-4 (SuLi-IC/10_pos.f90:64-82).  Runs after the w sweep, which
// reads the values before zeroing.
__global__ __launch_bounds__(kThreads) void k_posdin_hydro_uv(Geo g, const double *hs, double *u, double *v)
{
    const int nx = g.nx, ny = g.ny;
    // one thread per cell of the cell tiles: its low x and y faces, and in
    // the last column and row the high one too
    for_cells(g, [&](int i, int j, int k) {
        const long long iu = (i + 1) + (long long)(nx + 3) * ((j + 1) + (long long)(g.ny + 2) * (k + 1));
        const long long iv = (i + 1) + (long long)(nx + 2) * ((j + 1) + (long long)(g.ny + 3) * (k + 1));
        const bool cx0 = face_x(g, hs, i, j, k) < 0.0001, cy0 = face_y(g, hs, i, j, k) < 0.0001;
        // face i + 1 is in range for every lane; the last column acts on it
        const bool cx1 = face_x(g, hs, i + 1, j, k) < 0.0001;
        if (cx0) u[iu] = 0.;
        if (i == nx - 1 && cx1) u[iu + 1] = 0.;
        if (cy0) v[iv] = 0.;
        if (j == ny - 1 && face_y(g, hs, i, ny, k) < 0.0001) v[iv + (nx + 2)] = 0.;
    });
}

// heaviside with t_tens = 0 (SuLi-IC/14_level_set.f90:555-590): the
// smoothed Heaviside of the level set over a band of half-width eps =
// alpha1*dx1 (dx1 = max(dx,dy,dz), :41), then rho and the viscosity as its
// linear blends between the two fluids.  hs and mu may be null.
__global__ __launch_bounds__(kThreads) void k_heaviside(long long n, const double *__restrict__ ls, double eps,
                                                        double rho1, double rho2, double mu1, double mu2,
                                                        double *__restrict__ hs_out, double *__restrict__ rho,
                                                        double *__restrict__ mu)
{
    const double pi = 3.141592653589793;  // acos(-1.) in FP64
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long c = (long long)blockIdx.x * kThreads + threadIdx.x; c < n; c += stride) {
        const double l = ls[c];
        double hs;
        if (l < -eps)
            hs = 0.;
        else if (l > eps)
            hs = 1.;
        else
            hs = 0.5 * (1. + l / eps + 1. / pi * sin(pi * l / eps));
        if (hs_out) hs_out[c] = hs;
        rho[c] = rho1 * (1. - hs) + rho2 * hs;
        if (mu) mu[c] = mu1 * (1. - hs) + mu2 * hs;
    }
}

// ---- heaviside with t_tens = 1 (SuLi-IC/14_level_set.f90:519-550,563-567)

// phi1(m), m = -2..n+3 (1-based), of the grid line through a cell: the
// line's values (base[(m-1)*stride]), its periodic wrap, or its ends
// extrapolated with zero derivative, as the WENO routines fill them
// (SuLi-IC/8_derivadas.f90:702-714 and :742-756 for x; y and z alike).
struct Line {
    const double *base;
    long long stride;
    int n;
    bool per;
    __device__ __forceinline__ double at(int m) const
    {
        if (m >= 1 && m <= n) return base[(long long)(m - 1) * stride];
        if (per) return base[(long long)((m <= 0 ? m + n - 1 : m - n + 1) - 1) * stride];
        const double c11 = 1. / 11.;
        if (m <= 0) {
            const double p1 = base[0], p2 = base[stride], p3 = base[2 * stride];
            const double g0 = c11 * (18. * p1 - 9. * p2 + 2. * p3);
            if (m == 0) return g0;
            const double g1 = c11 * (18. * g0 - 9. * p1 + 2. * p2);
            if (m == -1) return g1;
            return c11 * (18. * g1 - 9. * g0 + 2. * p1);
        }
        const double p1 = base[(long long)(n - 1) * stride], p2 = base[(long long)(n - 2) * stride],
                     p3 = base[(long long)(n - 3) * stride];
        const double g0 = c11 * (18. * p1 - 9. * p2 + 2. * p3);
        if (m == n + 1) return g0;
        const double g1 = c11 * (18. * g0 - 9. * p1 + 2. * p2);
        if (m == n + 2) return g1;
        return c11 * (18. * g1 - 9. * g0 + 2. * p1);
    }
};

// The WENO5 one-sided derivatives at cell i of a line (wenox's loop body,
// SuLi-IC/8_derivadas.f90:759-809, in its operation order), returned as the
// one of larger magnitude, as heaviside picks (14_level_set.f90:523-527).
__device__ double weno_pick(const Line &L, int i, double h)
{
    double ph[7];   // phi1(i-3 .. i+3)
#pragma unroll
    for (int m = 0; m < 7; ++m) ph[m] = L.at(i - 3 + m);
    // up(m) = (phi1(m+3) - phi1(m+2))/h, m = i-4..i -> up[m - i + 4]
    // un(m) = (phi1(m-2) - phi1(m-3))/h, m = i..i+4 -> un[m - i]
    double up[5], un[5];
#pragma unroll
    for (int t = 0; t < 5; ++t) {
        up[t] = (ph[t + 2] - ph[t + 1]) / h;
        un[t] = (ph[t + 1] - ph[t]) / h;
    }
    const double a1 = 13. / 12., a2 = 1. / 4., a3 = 1. / 6., eps = 0.00000001;
    const double wgt[3] = {0.1, 0.6, 0.3};
    // up(i - d) = up[4 - d]; un(i + d) = un[d]
#define UP(d) up[4 - (d)]
#define UN(d) un[(d)]
    double isp[3], isn[3], t;
    t = UP(0) - 2 * UP(1) + UP(2);
    isp[0] = a1 * t * t;
    t = UP(0) - 4 * UP(1) + 3 * UP(2);
    isp[0] = isp[0] + a2 * t * t;
    t = UN(0) - 2 * UN(1) + UN(2);
    isn[0] = a1 * t * t;
    t = UN(0) - 4 * UN(1) + 3 * UN(2);
    isn[0] = isn[0] + a2 * t * t;
    t = UP(1) - 2 * UP(2) + UP(3);
    isp[1] = a1 * t * t;
    t = UP(1) - UP(3);
    isp[1] = isp[1] + a2 * t * t;
    t = UN(1) - 2 * UN(2) + UN(3);
    isn[1] = a1 * t * t;
    t = UN(1) - UN(3);
    isn[1] = isn[1] + a2 * t * t;
    t = UP(2) - 2 * UP(3) + UP(4);
    isp[2] = a1 * t * t;
    t = 3 * UP(2) - 4 * UP(3) + UP(4);
    isp[2] = isp[2] + a2 * t * t;
    t = UN(2) - 2 * UN(3) + UN(4);
    isn[2] = a1 * t * t;
    t = 3 * UN(2) - 4 * UN(3) + UN(4);
    isn[2] = isn[2] + a2 * t * t;
    double alp[3], aln[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        alp[q] = wgt[q] / ((eps + isp[q]) * (eps + isp[q]));
        aln[q] = wgt[q] / ((eps + isn[q]) * (eps + isn[q]));
    }
    const double sp = alp[0] + alp[1] + alp[2], sn = aln[0] + aln[1] + aln[2];
    const double dp = a3 * (alp[0] / sp * (2 * UP(0) - 7 * UP(1) + 11 * UP(2)) +
                            alp[1] / sp * (-UP(1) + 5 * UP(2) + 2 * UP(3)) +
                            alp[2] / sp * (2 * UP(2) + 5 * UP(3) - UP(4)));
    const double dn = a3 * (aln[0] / sn * (2 * UN(0) - 7 * UN(1) + 11 * UN(2)) +
                            aln[1] / sn * (-UN(1) + 5 * UN(2) + 2 * UN(3)) +
                            aln[2] / sn * (2 * UN(2) + 5 * UN(3) - UN(4)));
#undef UP
#undef UN
    return fabs(dp) > fabs(dn) ? dp : dn;
}

// heaviside with the surface-tension gradients: as k_heaviside, and hsx,
// hsy, hsz = 0.5 g (1 + cos(pi ls/eps))/eps inside the band, 0 outside,
// with g the level set's WENO5 derivative along each axis (x, y periodic
// when g.px / g.py, z never).  One thread per cell.
__global__ __launch_bounds__(kThreads) void k_heaviside_tens(Geo g, const double *__restrict__ ls, double eps,
                                                             double rho1, double rho2, double mu1, double mu2,
                                                             double dx, double dy, double dz,
                                                             double *__restrict__ hs_out, double *__restrict__ rho,
                                                             double *__restrict__ mu, double *__restrict__ hsx,
                                                             double *__restrict__ hsy, double *__restrict__ hsz)
{
    const double pi = 3.141592653589793;  // acos(-1.) in FP64
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long plane = (long long)nx * ny;
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long c = (long long)blockIdx.x * kThreads + threadIdx.x; c < g.n; c += stride) {
        const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / plane);
        const double l = ls[c];
        double hs;
        double gx = 0., gy = 0., gz = 0.;
        const bool band = !(l < -eps) && !(l > eps);
        if (l < -eps)
            hs = 0.;
        else if (l > eps)
            hs = 1.;
        else
            hs = 0.5 * (1. + l / eps + 1. / pi * sin(pi * l / eps));
        if (band) {
            const double cs = 1. + cos(pi * l / eps);
            const Line lx{ls + (c - i), 1, nx, g.px != 0};
            const Line ly{ls + (c - (long long)j * nx), nx, ny, g.py != 0};
            const Line lz{ls + (c - (long long)k * plane), plane, nz, false};
            gx = 0.5 * weno_pick(lx, i + 1, dx) * cs / eps;
            gy = 0.5 * weno_pick(ly, j + 1, dy) * cs / eps;
            gz = 0.5 * weno_pick(lz, k + 1, dz) * cs / eps;
        }
        if (hs_out) hs_out[c] = hs;
        rho[c] = rho1 * (1. - hs) + rho2 * hs;
        if (mu) mu[c] = mu1 * (1. - hs) + mu2 * hs;
        hsx[c] = gx;
        hsy[c] = gy;
        hsz[c] = gz;
    }
}

// ---- exp_press (SuLi-IC/11_diversos.f90:455-495) ----------------------------

// px, py, pz = -(prdaux - prdaux_behind)/h/rho_face on every face, with
// prdaux = prd1 + (prd1 + prd0) formed from prd0 before its increment (the
// reference forms prd0 + prd1 first, :466-467: the same operations).  z
// differences divide by dzz(k) when dzz is given (SuLi-ICdz), else dzu.
// One thread per cell of the cell tiles (for_cells): its faces on the low
// side in x, y and z, and in the last column, row and plane the face on the
// high side too; prdaux of the cell is formed once for its faces.
__global__ __launch_bounds__(kThreads) void k_exp_press(Geo g, const double *__restrict__ rho,
                                                        const double *__restrict__ dz,
                                                        const double *__restrict__ dzz, double dx, double dy,
                                                        double dzu, const double *__restrict__ prd1,
                                                        const double *__restrict__ prd0, double *__restrict__ px,
                                                        double *__restrict__ py, double *__restrict__ pz)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long pb = nx + 2, pbb = (long long)(nx + 2) * (ny + 2);
    auto aux = [&](long long x) { return prd1[x] + (prd1[x] + prd0[x]); };
    for_cells(g, [&](int i, int j, int k) {
        const long long ip = (i + 1) + pb * (j + 1) + pbb * (k + 1);
        const long long fx = ((long long)k * ny + j) * (nx + 1) + i;
        const long long fy = ((long long)k * (ny + 1) + j) * nx + i;
        const long long fz = ((long long)k * ny + j) * nx + i;
        const double a0 = aux(ip);
        const double hk = dzz ? dzz[k + 1] : dzu;
        px[fx] = -(a0 - aux(ip - 1)) / dx / face_x(g, rho, i, j, k);
        // face i + 1 is in the box for every lane; the last column stores it
        const double x1 = -(aux(ip + 1) - a0) / dx / face_x(g, rho, i + 1, j, k);
        if (i == nx - 1) px[fx + 1] = x1;
        py[fy] = -(a0 - aux(ip - pb)) / dy / face_y(g, rho, i, j, k);
        pz[fz] = -(a0 - aux(ip - pbb)) / hk / face_z(g, rho, dz, i, j, k);
        if (j == ny - 1) py[fy + nx] = -(aux(ip + pb) - a0) / dy / face_y(g, rho, i, ny, k);
        if (k == nz - 1) {
            const double ht = dzz ? dzz[nz + 1] : dzu;
            pz[fz + (long long)nx * ny] = -(aux(ip + pbb) - a0) / ht / face_z(g, rho, dz, i, j, nz);
        }
    });
}

// prd0 = prd1 + prd0 over the whole halo box (:466).
__global__ __launch_bounds__(kThreads) void k_add_box(long long n, const double *__restrict__ prd1,
                                                      double *__restrict__ prd0)
{
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long c = (long long)blockIdx.x * kThreads + threadIdx.x; c < n; c += stride) prd0[c] = prd1[c] + prd0[c];
}

// ---- SuLi-IT posdin's surface layer (SuLi-IT/10_pos.f90:84-93) --------------

// eta1 += prd1(:,:,nz)/gz and dz(:,:,nz) += eta1 - eta0 over the
// (0:nx+1, 0:ny+1) box; dz in prd1's layout.
__global__ __launch_bounds__(kThreads) void k_it_layer(Geo g, double gz, const double *__restrict__ prd1,
                                                       const double *__restrict__ eta0, double *__restrict__ eta1,
                                                       double *__restrict__ dz)
{
    const long long ne = (long long)(g.nx + 2) * (g.ny + 2);
    const long long top = ne * g.nz;
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long e = (long long)blockIdx.x * kThreads + threadIdx.x; e < ne; e += stride) {
        const double et = eta1[e] + prd1[top + e] / gz;
        eta1[e] = et;
        dz[top + e] = dz[top + e] + (et - eta0[e]);
    }
}

// dz_it (SuLi-IT/11_diversos.f90:80-141): dzx (u layout) and dzy (v layout)
// on planes nz and nz+1, dzz (w layout) on planes nz, nz+1 and nz+2, from
// the cell thicknesses dz.  End faces: periodic pairs (x when g.px, y when
// g.py) or the linear extrapolation of the two faces inside.
__global__ __launch_bounds__(kThreads) void k_dz_it(Geo g, const double *__restrict__ dz, double *__restrict__ dzx,
                                                    double *__restrict__ dzy, double *__restrict__ dzz)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long pb = nx + 2, pbb = (long long)(nx + 2) * (ny + 2);
    const long long nux = (long long)(nx + 3) * (ny + 2), nvy = (long long)(nx + 2) * (ny + 3);
    const long long nx2 = 2 * nux, ny2 = 2 * nvy, nz3 = 3 * pbb;
    auto D = [&](int i, int j, int k) { return dz[i + pb * j + pbb * k]; };
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long t = (long long)blockIdx.x * kThreads + threadIdx.x; t < nx2 + ny2 + nz3; t += stride) {
        if (t < nx2) {
            const int k = nz + (int)(t / nux);
            const long long r = t % nux;
            const int i = (int)(r % (nx + 3)), j = (int)(r / (nx + 3));
            auto mx = [&](int f) { return (D(f, j, k) + D(f - 1, j, k)) * 0.5; };
            double v;
            if (i >= 1 && i <= nx + 1)
                v = mx(i);
            else if (g.px)
                v = i == 0 ? (D(0, j, k) + D(nx, j, k)) * 0.5 : (D(1, j, k) + D(nx + 1, j, k)) * 0.5;
            else if (i == 0)
                v = mx(1) + (mx(1) - mx(2));
            else
                v = mx(nx + 1) + (mx(nx + 1) - mx(nx));
            dzx[i + (nx + 3) * (j + (long long)(ny + 2) * k)] = v;
        } else if (t < nx2 + ny2) {
            const long long q = t - nx2;
            const int k = nz + (int)(q / nvy);
            const long long r = q % nvy;
            const int i = (int)(r % (nx + 2)), j = (int)(r / (nx + 2));
            auto my = [&](int f) { return (D(i, f, k) + D(i, f - 1, k)) * 0.5; };
            double v;
            if (j >= 1 && j <= ny + 1)
                v = my(j);
            else if (g.py)
                v = j == 0 ? (D(i, 0, k) + D(i, ny, k)) * 0.5 : (D(i, 1, k) + D(i, ny + 1, k)) * 0.5;
            else if (j == 0)
                v = my(1) + (my(1) - my(2));
            else
                v = my(ny + 1) + (my(ny + 1) - my(ny));
            dzy[i + (nx + 2) * (j + (long long)(ny + 3) * k)] = v;
        } else {
            const long long q = t - nx2 - ny2;
            const int k = nz + (int)(q / pbb);
            const long long e = q % pbb;
            const int i = (int)(e % pb), j = (int)(e / pb);
            auto mz = [&](int f) { return (D(i, j, f) + D(i, j, f - 1)) * 0.5; };
            const double v = k <= nz + 1 ? mz(k) : mz(nz + 1) + (mz(nz + 1) - mz(nz));
            dzz[e + pbb * k] = v;
        }
    }
}

// Tables of a constant-density operator: face 0 of row 0 of each plane
// (every face of the plane holds the same value).
__global__ __launch_bounds__(kThreads) void k_face_tables(Geo g, const double *sx, const double *sy,
                                                          const double *sz, const float *sx32, const float *sy32,
                                                          const float *sz32, double *tx, double *ty, double *tz,
                                                          float *tx32, float *ty32, float *tz32)
{
    const long long plane = (long long)g.nx * g.ny;
    for (int k = threadIdx.x; k <= g.nz; k += kThreads) {
        const long long fx = (long long)k * g.ny * (g.nx + 1), fy = (long long)k * (g.ny + 1) * g.nx;
        if (k < g.nz) {
            tx[k] = sx[fx];
            ty[k] = sy[fy];
            if (sx32) {
                tx32[k] = sx32[fx];
                ty32[k] = sy32[fy];
            }
        }
        tz[k] = sz[k * plane];
        if (sz32) tz32[k] = sz32[k * plane];
    }
}

// Row tables of the FP32 operator (mixed solver): one value per plane for
// the x and y faces of plane k (rx[8k .. 8k+7], ry[4k .. 4k+3], the value
// repeated) and per z face kf (rz[4kf .. 4kf+3]),
// taken where a centred obstacle or a surface band is least likely (the
// cell (nx/4, ny/4)), and one bit per row (j, k): set when every x face of
// the row, its y faces j and j + 1 and its z faces k and k + 1 hold exactly
// (bit for bit) those values.  Where the density is uniform across a plane
// (one phase, walls included) the scaled coefficients are the same sums in
// the same order everywhere, so most rows of a two-phase problem have the
// bit set.  The step kernels take such a row's coefficients from the tables
// and read the face arrays only for the others: the same values, the same
// results.  Bits: word k * row_words(ny) + j / 32, bit j % 32.  One
// workgroup per word (each wave checks every fourth of its 32 rows); the
// tables are written by the workgroup of word 0 of each plane.
__device__ __forceinline__ int row_words(int ny) { return (ny + 31) / 32; }

__global__ __launch_bounds__(kThreads) void k_row_tables(Geo g, const float *sx32, const float *sy32,
                                                         const float *sz32, float *rx, float *ry, float *rz,
                                                         unsigned *rowreg)
{
    __shared__ unsigned bits[kThreads / 64];
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int wpp = row_words(ny);
    const int k = blockIdx.x / wpp, word = blockIdx.x % wpp;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i0 = nx / 4, j0 = ny / 4;
    const long long plane = (long long)nx * ny;
    const long long fxk = (long long)k * ny * (nx + 1), fyk = (long long)k * (ny + 1) * nx;
    const float vx = sx32[fxk + (long long)j0 * (nx + 1) + i0 + 1];
    const float vy = sy32[fyk + (long long)(j0 + 1) * nx + i0];
    const float vz0 = sz32[(long long)k * plane + (long long)j0 * nx + i0];
    const float vz1 = sz32[(long long)(k + 1) * plane + (long long)j0 * nx + i0];
    unsigned mine = 0;
    for (int r = wave; r < 32; r += kThreads / 64) {
        const int j = word * 32 + r;
        if (j >= ny) break;   // whole waves
        // & rather than &&: every load is issued without waiting for the
        // comparisons before it
        bool ok = true;
        for (int i = lane; i <= nx; i += 64) ok &= sx32[fxk + (long long)j * (nx + 1) + i] == vx;
        for (int i = lane; i < nx; i += 64) {
            ok &= (sy32[fyk + (long long)j * nx + i] == vy) & (sy32[fyk + (long long)(j + 1) * nx + i] == vy);
            ok &= (sz32[(long long)k * plane + (long long)j * nx + i] == vz0) &
                  (sz32[(long long)(k + 1) * plane + (long long)j * nx + i] == vz1);
        }
        if (__all(ok)) mine |= 1u << r;
    }
    if (lane == 0) bits[wave] = mine;
    __syncthreads();
    if (threadIdx.x == 0) rowreg[blockIdx.x] = bits[0] | bits[1] | bits[2] | bits[3];
    if (word == 0 && threadIdx.x < 8) {
        rx[8 * k + threadIdx.x] = vx;
        if (threadIdx.x < 4) {
            ry[4 * k + threadIdx.x] = vy;
            rz[4 * k + threadIdx.x] = vz0;
            if (k == nz - 1) rz[4 * nz + threadIdx.x] = vz1;
        }
    }
}

// SuLi-IT posdin (SuLi-IT/10_pos.f90:73-132) without dz_it, which updates
// the caller's layer thicknesses: eta1 += prd1(:,:,nz)/gz over its whole
// (0:nx+1,0:ny+1) box (skipped when eta1 is null); u, v -= tetah dt/dx
// (dy) times the pressure difference; w -= tetah dt/dzz times it,
// k = 1..nz+1.
__global__ __launch_bounds__(kThreads) void k_it_posdin(Geo g, double dx, double dy, double dt, double tetah,
                                                        double gz, const double *prd1, const double *dz3,
                                                        double *eta1, double *u, double *v, double *w)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long nfx = (long long)(nx + 1) * ny * nz;
    const long long nfy = (long long)nx * (ny + 1) * nz;
    const long long nfz = (long long)nx * ny * (nz + 1);
    const long long neta = eta1 ? (long long)(nx + 2) * (ny + 2) : 0;
    const long long stride = (long long)gridDim.x * kThreads;
    const double aux1 = tetah * dt / dx, aux2 = tetah * dt / dy;
    const long long pb = nx + 2, pbb = (long long)(nx + 2) * (ny + 2);
    for (long long t = (long long)blockIdx.x * kThreads + threadIdx.x; t < nfx + nfy + nfz + neta; t += stride) {
        if (t < nfx) {
            const int f = (int)(t % (nx + 1));
            const long long jk = t / (nx + 1);
            const int j = (int)(jk % ny), k = (int)(jk / ny);
            const long long ip = (f + 1) + pb * (j + 1) + pbb * (k + 1);
            const long long iu = (f + 1) + (long long)(nx + 3) * ((j + 1) + (long long)(ny + 2) * (k + 1));
            u[iu] = u[iu] - aux1 * (prd1[ip] - prd1[ip - 1]);
        } else if (t < nfx + nfy) {
            const long long s = t - nfx;
            const int i = (int)(s % nx);
            const long long jk = s / nx;
            const int f = (int)(jk % (ny + 1)), k = (int)(jk / (ny + 1));
            const long long ip = (i + 1) + pb * (f + 1) + pbb * (k + 1);
            const long long iv = (i + 1) + (long long)(nx + 2) * ((f + 1) + (long long)(ny + 3) * (k + 1));
            v[iv] = v[iv] - aux2 * (prd1[ip] - prd1[ip - pb]);
        } else if (t < nfx + nfy + nfz) {
            const long long s = t - nfx - nfy;
            const int i = (int)(s % nx);
            const long long jk = s / nx;
            const int j = (int)(jk % ny), f = (int)(jk / ny);
            const long long ip = (i + 1) + pb * (j + 1) + pbb * (f + 1);
            const long long iw = (i + 1) + (long long)(nx + 2) * ((j + 1) + (long long)(ny + 2) * (f + 1));
            w[iw] = w[iw] - tetah * dt / dz3[iw] * (prd1[ip] - prd1[ip - pbb]);
        } else {
            const long long e = t - nfx - nfy - nfz;   // eta1 box index I + (nx+2) J
            eta1[e] = eta1[e] + prd1[e + pbb * nz] / gz;
        }
    }
}

int grid_for(long long n)
{
    long long b = (n + kThreads - 1) / kThreads;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}

// Grid of a for_cells kernel without partial sums: a workgroup per cell
// tile, up to kCellMaxGrid (further tiles take further passes).
constexpr int kCellMaxGrid = 1 << 20;
dim3 cell_grid(const Geo &g) { return dim3((unsigned)std::min(cell_tiles(g), kCellMaxGrid)); }

}  // namespace

int init_blocks(const Geo &g) { return std::max(1, std::min(cell_tiles(g), kInitBlocks)); }

void launch_face_coefs(const Geo &g, const SetupIn &in, double *sx, double *sy, double *sz, hipStream_t s)
{
    hipLaunchKernelGGL(k_face_coefs, cell_grid(g), dim3(kThreads), 0, s, g, in, sx, sy, sz);
}

void launch_diag_rhs(const Geo &g, const SetupIn &in, const double *sx, const double *sy, const double *sz,
                     double *d, double *b, double *e, hipStream_t s)
{
    hipLaunchKernelGGL(k_diag_rhs, dim3(grid_for(g.n)), dim3(kThreads), 0, s, g, in, sx, sy, sz, d, b, e);
}

void launch_scale_faces(const Geo &g, const double *d, double *sx, double *sy, double *sz, float *sx32,
                        float *sy32, float *sz32, hipStream_t s)
{
    const long long nf = (long long)(g.nx + 1) * g.ny * g.nz + (long long)g.nx * (g.ny + 1) * g.nz +
                         (long long)g.nx * g.ny * (g.nz + 1);
    hipLaunchKernelGGL(k_scale_faces, dim3(grid_for(nf)), dim3(kThreads), 0, s, g, d, sx, sy, sz, sx32, sy32,
                       sz32);
}

void launch_face_tables(const Geo &g, const double *sx, const double *sy, const double *sz, const float *sx32,
                        const float *sy32, const float *sz32, double *tx, double *ty, double *tz, float *tx32,
                        float *ty32, float *tz32, hipStream_t s)
{
    hipLaunchKernelGGL(k_face_tables, dim3(1), dim3(kThreads), 0, s, g, sx, sy, sz, sx32, sy32, sz32, tx, ty, tz,
                       tx32, ty32, tz32);
}

void launch_row_tables(const Geo &g, const float *sx32, const float *sy32, const float *sz32, float *rx, float *ry,
                       float *rz, unsigned *rowreg, hipStream_t s)
{
    const int words = (g.ny + 31) / 32 * g.nz;
    hipLaunchKernelGGL(k_row_tables, dim3(words), dim3(kThreads), 0, s, g, sx32, sy32, sz32, rx, ry, rz, rowreg);
}

void launch_init(const Geo &g, const SetupIn &in, Op<double> op, const double *d, double *b, double *e,
                 double *r0, double *r64, float *r32, float *c32, const Partials &P, int nb0, CgScalars *S,
                 hipStream_t s)
{
    const int np = nb0;   // the projection's workgroups (and partials)
    hipLaunchKernelGGL(k_init_residual, dim3(nb0), dim3(kThreads), 0, s, g, in, op, d, b, e, r0, P);
    hipLaunchKernelGGL(k_init_totals, dim3(1), dim3(kThreads), 0, s, P, nb0, S);
    hipLaunchKernelGGL(k_init_project, dim3(np), dim3(kThreads), 0, s, g, in.solid, d, b, r0, r64, r32, c32, P, S);
    hipLaunchKernelGGL(k_init_finalize, dim3(1), dim3(kThreads), 0, s, P, np, S);
}

void launch_unscale(const Geo &g, const double *e, const double *d, double *prd1, const CgScalars *S,
                    const uint8_t *solid, hipStream_t s)
{
    const long long nbox = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    hipLaunchKernelGGL(k_unscale, dim3(grid_for(nbox)), dim3(kThreads), 0, s, g, e, d, prd1, S, solid);
}

void launch_posdin(const Geo &g, const SetupIn &in, double *u, double *v, double *w, hipStream_t s)
{
    hipLaunchKernelGGL(k_posdin, cell_grid(g), dim3(kThreads), 0, s, g, in, u, v, w, true);
}

void launch_heaviside(long long n, const double *ls, double eps, double rho1, double rho2, double mu1, double mu2,
                      double *hs, double *rho, double *mu, hipStream_t s)
{
    hipLaunchKernelGGL(k_heaviside, dim3(grid_for(n)), dim3(kThreads), 0, s, n, ls, eps, rho1, rho2, mu1, mu2, hs,
                       rho, mu);
}

void launch_heaviside_tens(const Geo &g, const double *ls, double eps, double rho1, double rho2, double mu1,
                           double mu2, double dx, double dy, double dz, double *hs, double *rho, double *mu,
                           double *hsx, double *hsy, double *hsz, hipStream_t s)
{
    hipLaunchKernelGGL(k_heaviside_tens, dim3(grid_for(g.n)), dim3(kThreads), 0, s, g, ls, eps, rho1, rho2, mu1, mu2,
                       dx, dy, dz, hs, rho, mu, hsx, hsy, hsz);
}

void launch_exp_press(const Geo &g, const double *rho, const double *dz, const double *dzz, double dx, double dy,
                      double dzu, const double *prd1, double *prd0, double *px, double *py, double *pz, hipStream_t s)
{
    const long long nbox = (long long)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
    hipLaunchKernelGGL(k_exp_press, cell_grid(g), dim3(kThreads), 0, s, g, rho, dz, dzz, dx, dy, dzu, prd1, prd0, px,
                       py, pz);
    hipLaunchKernelGGL(k_add_box, dim3(grid_for(nbox)), dim3(kThreads), 0, s, nbox, prd1, prd0);
}

void launch_it_layers(const Geo &g, double gz, const double *prd1, const double *eta0, double *eta1, double *dz,
                      double *dzx, double *dzy, double *dzz, hipStream_t s)
{
    const long long ne = (long long)(g.nx + 2) * (g.ny + 2);
    const long long nt = 2 * (long long)(g.nx + 3) * (g.ny + 2) + 2 * (long long)(g.nx + 2) * (g.ny + 3) + 3 * ne;
    hipLaunchKernelGGL(k_it_layer, dim3(grid_for(ne)), dim3(kThreads), 0, s, g, gz, prd1, eta0, eta1, dz);
    hipLaunchKernelGGL(k_dz_it, dim3(grid_for(nt)), dim3(kThreads), 0, s, g, dz, dzx, dzy, dzz);
}

void launch_it_posdin(const Geo &g, double dx, double dy, double dt, double tetah, double gz, const double *prd1,
                      const double *dz3, double *eta1, double *u, double *v, double *w, hipStream_t s)
{
    const long long n = (long long)(g.nx + 1) * g.ny * g.nz + (long long)g.nx * (g.ny + 1) * g.nz +
                        (long long)g.nx * g.ny * (g.nz + 1) + (eta1 ? (long long)(g.nx + 2) * (g.ny + 2) : 0);
    hipLaunchKernelGGL(k_it_posdin, dim3(grid_for(n)), dim3(kThreads), 0, s, g, dx, dy, dt, tetah, gz, prd1, dz3,
                       eta1, u, v, w);
}

void launch_pressh(const Geo &g, const double *rho, const double *dz, double dzu, double gz, double *prd1,
                   hipStream_t s)
{
    const long long ncol = (long long)g.nx * g.ny;
    hipLaunchKernelGGL(k_pressh, dim3((unsigned)((ncol + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, g, rho,
                       dz, dzu, gz, prd1);
}

void launch_posdin_hydro(const Geo &g, const SetupIn &in, const double *hs, double *u, double *v, double *w,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_posdin, cell_grid(g), dim3(kThreads), 0, s, g, in, u, v, w, false);
    const long long nplane = (long long)(g.nx + 2) * (g.ny + 2);
    const dim3 grid((unsigned)((nplane + kThreads - 1) / kThreads));
    if (in.dz)
        hipLaunchKernelGGL(k_posdin_hydro_w<true>, grid, dim3(kThreads), 0, s, g, u, v, hs, in.dz, in.dzu, in.dx, in.dy,
                           w);
    else
        hipLaunchKernelGGL(k_posdin_hydro_w<false>, grid, dim3(kThreads), 0, s, g, u, v, hs, in.dz, in.dzu, in.dx,
                           in.dy, w);
    hipLaunchKernelGGL(k_posdin_hydro_uv, cell_grid(g), dim3(kThreads), 0, s, g, hs, u, v);
}

}  // namespace sps
