/*
 * CPU restatement of SuLi's pressure step: graddin() and posdin().
 *
 * TEST INFRASTRUCTURE ONLY (see suli_oracle.h).  It restates the reference
 * algorithm sweep by sweep, in the reference's operation order, so that on
 * x86-64 without FMA contraction it reproduces the compiled Fortran bit for
 * bit when omp_t = 0 (serial dot products).  Each block cites the reference
 * lines it follows.  Edge and corner cells of the halo box follow the
 * reference's sequential (one-thread) order: x ghosts, then y, then z.
 */
#include "suli_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Fortran layouts -> 0-based linear offsets (see suli_oracle.h). */
#define IP(i, j, k) ((size_t)(i) + (size_t)(nx + 2) * ((size_t)(j) + (size_t)(ny + 2) * (size_t)(k)))
#define IU(i, j, k) ((size_t)(i) + (size_t)(nx + 3) * ((size_t)(j) + (size_t)(ny + 2) * (size_t)(k)))
#define IV(i, j, k) ((size_t)(i) + (size_t)(nx + 2) * ((size_t)(j) + (size_t)(ny + 3) * (size_t)(k)))
#define IW(i, j, k) ((size_t)(i) + (size_t)(nx + 2) * ((size_t)(j) + (size_t)(ny + 2) * (size_t)(k)))
#define IC(i, j, k) ((size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1)))
#define IFX(i, j, k) ((size_t)((i) - 1) + (size_t)(nx + 1) * ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1)))
#define IFY(i, j, k) ((size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)(ny + 1) * (size_t)((k) - 1)))
#define IFZ(i, j, k) ((size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1)))

static int nthreads(const oracle_grid *g)
{
#ifdef _OPENMP
    return g->threads > 0 ? g->threads : omp_get_max_threads();
#else
    (void)g;
    return 1;
#endif
}

/* SuLi-IC/8_derivadas.f90:305-357 (x), :381-418 (y), :443-481 (z);
 * SuLi-ICdz/8_derivadas.f90:450-489 (dz-weighted z). */
void oracle_interp_faces(const oracle_grid *g, const double *rho, double *rhox,
                         double *rhoy, double *rhoz)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int nt = nthreads(g);
    int i, j, k;

#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++) {
            for (i = 2; i <= nx; i++)
                rhox[IFX(i, j, k)] = (rho[IC(i, j, k)] + rho[IC(i - 1, j, k)]) * 0.5;
            if (g->ccx0 != 0) {
                rhox[IFX(1, j, k)] = rhox[IFX(2, j, k)];
                rhox[IFX(nx + 1, j, k)] = rhox[IFX(nx, j, k)];
            } else {
                rhox[IFX(1, j, k)] = (rho[IC(1, j, k)] + rho[IC(nx, j, k)]) * 0.5;
                rhox[IFX(nx + 1, j, k)] = (rho[IC(1, j, k)] + rho[IC(nx, j, k)]) * 0.5;
            }
        }

#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz; k++) {
        for (j = 2; j <= ny; j++)
            for (i = 1; i <= nx; i++)
                rhoy[IFY(i, j, k)] = (rho[IC(i, j, k)] + rho[IC(i, j - 1, k)]) * 0.5;
        /* no periodic branch in y (8_derivadas.f90:403-414) */
        for (i = 1; i <= nx; i++) {
            rhoy[IFY(i, 1, k)] = rhoy[IFY(i, 2, k)];
            rhoy[IFY(i, ny + 1, k)] = rhoy[IFY(i, ny, k)];
        }
    }

#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 2; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                if (g->dzv)
                    rhoz[IFZ(i, j, k)] = (rho[IC(i, j, k)] * g->dzv[k] + rho[IC(i, j, k - 1)] * g->dzv[k - 1]) /
                                         (g->dzv[k] + g->dzv[k - 1]);
                else
                    rhoz[IFZ(i, j, k)] = (rho[IC(i, j, k)] + rho[IC(i, j, k - 1)]) * 0.5;
            }
    for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            rhoz[IFZ(i, j, 1)] = rhoz[IFZ(i, j, 2)];
            rhoz[IFZ(i, j, nz + 1)] = rhoz[IFZ(i, j, nz)];
        }
}

/* Ghost fill of a halo-box array in the reference's order: x faces (mirror
 * or periodic), then y, then z (always mirror).  9_grad.f90:188-239 (d),
 * :322-370 (r), :531-579 (p), :594-643 (e).  With top_sign = -1 the top
 * ghost plane is minus the interior (the Dirichlet lid extension; used for
 * r, p and e, never for d). */
static void fill_ghosts_s(const oracle_grid *g, double *a, double top_sign)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    int i, j, k;
    for (k = 0; k <= nz + 1; k++)
        for (j = 0; j <= ny + 1; j++) {
            if (g->ccx0 == 0) {
                a[IP(0, j, k)] = a[IP(nx, j, k)];
                a[IP(nx + 1, j, k)] = a[IP(1, j, k)];
            } else {
                a[IP(0, j, k)] = a[IP(1, j, k)];
                a[IP(nx + 1, j, k)] = a[IP(nx, j, k)];
            }
        }
    for (k = 0; k <= nz + 1; k++)
        for (i = 0; i <= nx + 1; i++) {
            if (g->ccy0 == 0) {
                a[IP(i, 0, k)] = a[IP(i, ny, k)];
                a[IP(i, ny + 1, k)] = a[IP(i, 1, k)];
            } else {
                a[IP(i, 0, k)] = a[IP(i, 1, k)];
                a[IP(i, ny + 1, k)] = a[IP(i, ny, k)];
            }
        }
    for (j = 0; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) {
            a[IP(i, j, 0)] = a[IP(i, j, 1)];
            a[IP(i, j, nz + 1)] = top_sign * a[IP(i, j, nz)];
        }
}

static void fill_ghosts(const oracle_grid *g, double *a) { fill_ghosts_s(g, a, 1.0); }

/* ghosts of a pressure-like array: the lid is Dirichlet when top_dirichlet */
static void fill_ghosts_p(const oracle_grid *g, double *a) { fill_ghosts_s(g, a, g->top_dirichlet ? -1.0 : 1.0); }

/* Sum of a[c]*b[c] over interior cells.  Serial k-j-i order (omp_t = 0,
 * 9_grad.f90:422-433, :482-493) or an OpenMP reduction (omp_t = 1,
 * :436-447, :496-507). */
static double dot_interior(const oracle_grid *g, const double *a, const double *b, int serial)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int nt = nthreads(g);
    double s = 0.0;
    int i, j, k;
    if (serial || nt <= 1) {
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++)
                    s = s + a[IP(i, j, k)] * b[IP(i, j, k)];
    } else {
#pragma omp parallel for num_threads(nt) private(i, j) reduction(+ : s)
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++)
                    s = s + a[IP(i, j, k)] * b[IP(i, j, k)];
    }
    return s;
}

/* Solid-cell mask (extension, not in the reference).  Fortran halo index
 * -> 1-based cell by the mirror/periodic rule; a face is closed when the
 * cell on either side is solid. */
static int cell_map(int I, int n, int per)
{
    if (I == 0) return per ? n : 1;
    if (I == n + 1) return per ? 1 : n;
    return I;
}

static int solid_at(const oracle_grid *g, int i, int j, int k)
{
    const int nx = g->nx, ny = g->ny;
    return g->solid && g->solid[IC(cell_map(i, nx, g->ccx0 == 0), cell_map(j, ny, g->ccy0 == 0),
                                   cell_map(k, g->nz, 0))];
}

/* x face i lies between cells i-1 and i (Fortran indices), likewise y, z. */
static int closed_x(const oracle_grid *g, int i, int j, int k) { return solid_at(g, i - 1, j, k) || solid_at(g, i, j, k); }
static int closed_y(const oracle_grid *g, int i, int j, int k) { return solid_at(g, i, j - 1, k) || solid_at(g, i, j, k); }
static int closed_z(const oracle_grid *g, int i, int j, int k) { return solid_at(g, i, j, k - 1) || solid_at(g, i, j, k); }

/* Everything graddin computes before the CG loop that tests want to see:
 * face coefficients, diagonal with ghosts, q = -div u (9_grad.f90:45-239). */
static void build_unscaled(const oracle_grid *g, const double *u, const double *v, const double *w,
                           const double *rho, double *sx, double *sy, double *sz, double *d, double *q)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int nt = nthreads(g);
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz,
                 nfz = (size_t)nx * ny * (nz + 1);
    int i, j, k;

    /* IBM: zero face velocities inside obstacles (9_grad.f90:45-88) */
    double *u1 = NULL, *v1 = NULL, *w1 = NULL;
    if (g->ibm_mask) {
        u1 = malloc(nfx * sizeof(double));
        v1 = malloc(nfy * sizeof(double));
        w1 = malloc(nfz * sizeof(double));
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx + 1; i++)
                    u1[IFX(i, j, k)] = g->obsx[IFX(i, j, k)] >= 0.0 ? 0.0 : u[IU(i, j, k)];
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny + 1; j++)
                for (i = 1; i <= nx; i++)
                    v1[IFY(i, j, k)] = g->obsy[IFY(i, j, k)] >= 0.0 ? 0.0 : v[IV(i, j, k)];
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
        for (k = 1; k <= nz + 1; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++)
                    w1[IFZ(i, j, k)] = g->obsz[IFZ(i, j, k)] >= 0.0 ? 0.0 : w[IW(i, j, k)];
    }

    /* aux = dt / delta^2 (9_grad.f90:93-100) */
    const double aux1 = g->dt / (g->dx * g->dx);
    const double aux2 = g->dt / (g->dy * g->dy);
    const double aux3 = g->dt / (g->dz * g->dz);

    /* face densities (9_grad.f90:103-116) */
    double *rhox = malloc(nfx * sizeof(double));
    double *rhoy = malloc(nfy * sizeof(double));
    double *rhoz = malloc(nfz * sizeof(double));
    if (g->t_press == 1) {
        oracle_interp_faces(g, rho, rhox, rhoy, rhoz);
    } else {
        for (size_t c = 0; c < nfx; c++) rhox[c] = g->rho_f1;
        for (size_t c = 0; c < nfy; c++) rhoy[c] = g->rho_f1;
        for (size_t c = 0; c < nfz; c++) rhoz[c] = g->rho_f1;
    }

    /* face coefficients (9_grad.f90:118-148; SuLi-ICdz/9_grad.f90:141) */
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx + 1; i++)
                sx[IFX(i, j, k)] = aux1 / rhox[IFX(i, j, k)];
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny + 1; j++)
            for (i = 1; i <= nx; i++)
                sy[IFY(i, j, k)] = aux2 / rhoy[IFY(i, j, k)];
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz + 1; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                if (g->dzv)
                    sz[IFZ(i, j, k)] = g->dt / (g->dzv[k] * g->dzv[k] * rhoz[IFZ(i, j, k)]);
                else
                    sz[IFZ(i, j, k)] = aux3 / rhoz[IFZ(i, j, k)];
            }
    free(rhox);
    free(rhoy);
    free(rhoz);

    /* solid mask: closed faces carry no flux */
    if (g->solid) {
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx + 1; i++)
                    if (closed_x(g, i, j, k)) sx[IFX(i, j, k)] = 0.0;
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny + 1; j++)
                for (i = 1; i <= nx; i++)
                    if (closed_y(g, i, j, k)) sy[IFY(i, j, k)] = 0.0;
        for (k = 1; k <= nz + 1; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++)
                    if (closed_z(g, i, j, k)) sz[IFZ(i, j, k)] = 0.0;
    }

    /* diagonal (9_grad.f90:152-160) */
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++)
                d[IP(i, j, k)] = sx[IFX(i + 1, j, k)] + sx[IFX(i, j, k)] + sy[IFY(i, j + 1, k)] +
                                 sy[IFY(i, j, k)] + sz[IFZ(i, j, k + 1)] + sz[IFZ(i, j, k)];

    /* q = -div u (9_grad.f90:163-183; SuLi-ICdz/9_grad.f90:166,176) */
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                double du, dv, dw;
                if (g->ibm_mask) {
                    du = u1[IFX(i, j, k)] - u1[IFX(i + 1, j, k)];
                    dv = v1[IFY(i, j, k)] - v1[IFY(i, j + 1, k)];
                    dw = w1[IFZ(i, j, k)] - w1[IFZ(i, j, k + 1)];
                } else {
                    du = u[IU(i, j, k)] - u[IU(i + 1, j, k)];
                    dv = v[IV(i, j, k)] - v[IV(i, j + 1, k)];
                    dw = w[IW(i, j, k)] - w[IW(i, j, k + 1)];
                }
                if (g->solid) {
                    /* closed faces count with zero velocity; solid cells: d = 1, q = 0 */
                    const double u0 = closed_x(g, i, j, k) ? 0.0 : u[IU(i, j, k)];
                    const double u1 = closed_x(g, i + 1, j, k) ? 0.0 : u[IU(i + 1, j, k)];
                    const double v0 = closed_y(g, i, j, k) ? 0.0 : v[IV(i, j, k)];
                    const double v1 = closed_y(g, i, j + 1, k) ? 0.0 : v[IV(i, j + 1, k)];
                    const double w0 = closed_z(g, i, j, k) ? 0.0 : w[IW(i, j, k)];
                    const double w1 = closed_z(g, i, j, k + 1) ? 0.0 : w[IW(i, j, k + 1)];
                    du = u0 - u1;
                    dv = v0 - v1;
                    dw = w0 - w1;
                }
                const double dzk = g->dzzv ? g->dzzv[k] : g->dz;
                q[IC(i, j, k)] = du / g->dx + dv / g->dy + dw / dzk;
                if (solid_at(g, i, j, k)) {
                    d[IP(i, j, k)] = 1.0;
                    q[IC(i, j, k)] = 0.0;
                }
            }
    free(u1);
    free(v1);
    free(w1);

    /* ghosts of d (9_grad.f90:188-239) */
    fill_ghosts(g, d);
}

/* Symmetric Jacobi scaling of the face coefficients (9_grad.f90:256-288). */
static void scale_faces(const oracle_grid *g, const double *d, double *sx, double *sy, double *sz)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int nt = nthreads(g);
    int i, j, k;
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx + 1; i++)
                sx[IFX(i, j, k)] = sx[IFX(i, j, k)] / sqrt(d[IP(i, j, k)] * d[IP(i - 1, j, k)]);
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny + 1; j++)
            for (i = 1; i <= nx; i++)
                sy[IFY(i, j, k)] = sy[IFY(i, j, k)] / sqrt(d[IP(i, j, k)] * d[IP(i, j - 1, k)]);
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz + 1; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++)
                sz[IFZ(i, j, k)] = sz[IFZ(i, j, k)] / sqrt(d[IP(i, j, k)] * d[IP(i, j, k - 1)]);
}

void oracle_scaled_operator(const oracle_grid *g, const double *u, const double *v, const double *w,
                            const double *rho, double *shx, double *shy, double *shz, double *d,
                            double *q, double *b)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    memset(d, 0, (size_t)(nx + 2) * (ny + 2) * (nz + 2) * sizeof(double));
    build_unscaled(g, u, v, w, rho, shx, shy, shz, d, q);
    scale_faces(g, d, shx, shy, shz);
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
                b[IC(i, j, k)] = q[IC(i, j, k)] / sqrt(d[IP(i, j, k)]);
}

int oracle_graddin(const oracle_grid *g, const double *u, const double *v, const double *w,
                   const double *rho, double *prd1, oracle_stats *st)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    if (nx < 2 || ny < 2 || nz < 2) return -1;
    const int nt = nthreads(g);
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz,
                 nfz = (size_t)nx * ny * (nz + 1), ncell = (size_t)nx * ny * nz;
    int i, j, k;

    double *sx = malloc(nfx * sizeof(double));
    double *sy = malloc(nfy * sizeof(double));
    double *sz = malloc(nfz * sizeof(double));
    double *q = malloc(ncell * sizeof(double));
    double *d = calloc(nbox, sizeof(double));
    double *e = calloc(nbox, sizeof(double));
    double *r = calloc(nbox, sizeof(double));
    double *p = calloc(nbox, sizeof(double));
    double *mp = calloc(nbox, sizeof(double));
    if (!sx || !sy || !sz || !q || !d || !e || !r || !p || !mp) {
        free(sx); free(sy); free(sz); free(q); free(d); free(e); free(r); free(p); free(mp);
        return -2;
    }

    build_unscaled(g, u, v, w, rho, sx, sy, sz, d, q);

    /* e = prd1 * sqrt(d) on the whole box (9_grad.f90:242-252) */
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (size_t c = 0; c < nbox; c++) e[c] = prd1[c] * sqrt(d[c]);
    /* Dirichlet lid (extension): the top ghosts are minus the interior */
    if (g->top_dirichlet)
        for (j = 0; j <= ny + 1; j++)
            for (i = 0; i <= nx + 1; i++) e[IP(i, j, nz + 1)] = -e[IP(i, j, nz)];
    /* solid mask: solid cells (and their ghost images) start at 0 */
    if (g->solid)
        for (k = 0; k <= nz + 1; k++)
            for (j = 0; j <= ny + 1; j++)
                for (i = 0; i <= nx + 1; i++)
                    if (solid_at(g, i, j, k)) e[IP(i, j, k)] = 0.0;

    scale_faces(g, d, sx, sy, sz);

    /* initial residual r = A e - q/sqrt(d) (9_grad.f90:290-301) */
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++)
                r[IP(i, j, k)] = e[IP(i, j, k)] - sx[IFX(i + 1, j, k)] * e[IP(i + 1, j, k)] -
                                 sx[IFX(i, j, k)] * e[IP(i - 1, j, k)] - sy[IFY(i, j + 1, k)] * e[IP(i, j + 1, k)] -
                                 sy[IFY(i, j, k)] * e[IP(i, j - 1, k)] - sz[IFZ(i, j, k + 1)] * e[IP(i, j, k + 1)] -
                                 sz[IFZ(i, j, k)] * e[IP(i, j, k - 1)] - q[IC(i, j, k)] / sqrt(d[IP(i, j, k)]);

    /* Extension (project_null, not in the reference): every reference BC
     * leaves A singular with null vector 1, so the scaled operator has null
     * vector sqrt(d); remove the residual's component along it, making the
     * floating-point right-hand side exactly consistent.  Serial sums. */
    if (g->project_null && !g->top_dirichlet) {   /* a Dirichlet lid leaves no null space */
        /* with a solid mask the null vector is sqrt(d) on fluid cells only */
        double rn = 0.0, dd = 0.0;
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++) {
                    const double dc = solid_at(g, i, j, k) ? 0.0 : d[IP(i, j, k)];
                    rn = rn + r[IP(i, j, k)] * sqrt(dc);
                    dd = dd + dc;
                }
        const double gamma = rn / dd;
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++) {
                    const double dc = solid_at(g, i, j, k) ? 0.0 : d[IP(i, j, k)];
                    r[IP(i, j, k)] = r[IP(i, j, k)] - gamma * sqrt(dc);
                }
    }

    /* r.r, always serial (9_grad.f90:303-319) */
    double rr = dot_interior(g, r, r, 1);

    /* b.b for the relative stop rule (not in the reference) */
    double bb = 0.0;
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                const double bi = q[IC(i, j, k)] / sqrt(d[IP(i, j, k)]);
                bb = bb + bi * bi;
            }

    /* ghosts of r, p = r (9_grad.f90:322-380) */
    fill_ghosts_p(g, r);
    memcpy(p, r, nbox * sizeof(double));

    /* stop rule (9_grad.f90:384-398) */
    int cont = 0;
    double aux1 = fabs(rr);
    double thresh;
    if (g->stop_rule == ORACLE_STOP_RELATIVE)
        thresh = g->tol * g->tol * bb;
    else
        thresh = 0.0001 / (double)(nx * ny * nz);
    const int max_iter = g->max_iter > 0 ? g->max_iter : 10000;
    const int serial = g->omp_t == 0;
    if (st) st->rr0 = rr;

    /* CG loop (9_grad.f90:398-587) */
    while (aux1 > thresh && cont < max_iter) {
        cont = cont + 1;

        /* mp = A p (:408-420) */
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++)
                    mp[IP(i, j, k)] = p[IP(i, j, k)] - p[IP(i + 1, j, k)] * sx[IFX(i + 1, j, k)] -
                                      p[IP(i - 1, j, k)] * sx[IFX(i, j, k)] - p[IP(i, j + 1, k)] * sy[IFY(i, j + 1, k)] -
                                      p[IP(i, j - 1, k)] * sy[IFY(i, j, k)] - p[IP(i, j, k + 1)] * sz[IFZ(i, j, k + 1)] -
                                      p[IP(i, j, k - 1)] * sz[IFZ(i, j, k)];

        /* alpha (:422-455) */
        const double pap = dot_interior(g, p, mp, serial);
        const double alpha = rr / pap;

        /* e -= alpha p, r -= alpha A p (:458-480) */
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++) {
                    e[IP(i, j, k)] = e[IP(i, j, k)] - alpha * p[IP(i, j, k)];
                    r[IP(i, j, k)] = r[IP(i, j, k)] - alpha * mp[IP(i, j, k)];
                }

        /* beta (:482-515) */
        const double rrn = dot_interior(g, r, r, serial);
        const double beta = rrn / rr;
        rr = rrn;

        /* p = r + beta p (:518-527), ghosts (:531-579) */
#pragma omp parallel for num_threads(nt) private(i, j) if (nt > 1)
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++)
                    p[IP(i, j, k)] = r[IP(i, j, k)] + beta * p[IP(i, j, k)];
        fill_ghosts_p(g, p);

        aux1 = fabs(rr);
    }

    /* ghosts of e and un-scaling on the whole box (9_grad.f90:594-659) */
    fill_ghosts_p(g, e);
#pragma omp parallel for num_threads(nt) if (nt > 1)
    for (size_t c = 0; c < nbox; c++) prd1[c] = e[c] / sqrt(d[c]);

    if (st) {
        st->iters = cont;
        st->converged = aux1 <= thresh;
        st->bb = bb;
        st->rr = rr;
        st->thresh = thresh;
    }
    free(sx); free(sy); free(sz); free(q); free(d); free(e); free(r); free(p); free(mp);
    return aux1 <= thresh ? 0 : 1;
}

/* SuLi-IC/10_pos.f90:17-46,95; SuLi-ICdz/10_pos.f90:21,41. */
/* posdin's horizontal correction, taken in both branches
 * (SuLi-IC/10_pos.f90:24-37): u and v lose dt/dx (dt/dy) times the
 * pressure difference over the face density.  With a solid mask, closed
 * faces are left alone (extension). */
static void posdin_uv(const oracle_grid *g, double *u, double *v, const double *rhox,
                      const double *rhoy, const double *prd1)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double aux1 = g->dt / g->dx;
    const double aux2 = g->dt / g->dy;
    int i, j, k;
    for (k = 1; k <= nz; k++) {
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx + 1; i++)
                if (!closed_x(g, i, j, k))
                    u[IU(i, j, k)] = u[IU(i, j, k)] - aux1 * (prd1[IP(i, j, k)] - prd1[IP(i - 1, j, k)]) / rhox[IFX(i, j, k)];
        for (j = 1; j <= ny + 1; j++)
            for (i = 1; i <= nx; i++)
                if (!closed_y(g, i, j, k))
                    v[IV(i, j, k)] = v[IV(i, j, k)] - aux2 * (prd1[IP(i, j, k)] - prd1[IP(i, j - 1, k)]) / rhoy[IFY(i, j, k)];
    }
}

void oracle_posdin(const oracle_grid *g, double *u, double *v, double *w, const double *rho,
                   const double *prd1, double *prd)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz,
                 nfz = (size_t)nx * ny * (nz + 1);
    int i, j, k;
    double *rhox = malloc(nfx * sizeof(double));
    double *rhoy = malloc(nfy * sizeof(double));
    double *rhoz = malloc(nfz * sizeof(double));
    oracle_interp_faces(g, rho, rhox, rhoy, rhoz);
    posdin_uv(g, u, v, rhox, rhoy, prd1);
    for (k = 1; k <= nz + 1; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                const double dzk = g->dzzv ? g->dzzv[k] : g->dz;
                if (!closed_z(g, i, j, k))
                    w[IW(i, j, k)] = w[IW(i, j, k)] - g->dt / dzk * (prd1[IP(i, j, k)] - prd1[IP(i, j, k - 1)]) / rhoz[IFZ(i, j, k)];
            }
    memcpy(prd, prd1, (size_t)(nx + 2) * (ny + 2) * (nz + 2) * sizeof(double));
    free(rhox);
    free(rhoy);
    free(rhoz);
}

/* Hydrostatic pressure, t_press = 0 (SuLi-IC/9_grad.f90:667-702; dz(k)
 * weights: SuLi-ICdz/9_grad.f90:665-700).  Column integral from the top
 * (Simpson over two cells below nz-1); prd1's halo is not touched. */
void oracle_pressh(const oracle_grid *g, const double *rho, double gz, double *prd1)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    int i, j, k;
    for (k = nz; k >= 1; k--)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                if (g->dzv) {
                    const double *dz = g->dzv;
                    if (k == nz) {
                        prd1[IP(i, j, k)] = dz[k] * gz * rho[IC(i, j, k)];
                    } else if (k == nz - 1) {
                        prd1[IP(i, j, k)] = gz * (rho[IC(i, j, k)] * dz[k] + rho[IC(i, j, k + 1)] * dz[k + 1]) * 0.5 +
                                            prd1[IP(i, j, k + 1)];
                    } else {
                        const double intz = (rho[IC(i, j, k)] * dz[k] + 4. * rho[IC(i, j, k + 1)] * dz[k + 1] +
                                             rho[IC(i, j, k + 2)] * dz[k + 2]) / 3.;
                        prd1[IP(i, j, k)] = gz * intz + prd1[IP(i, j, k + 2)];
                    }
                } else {
                    const double dz = g->dz;
                    if (k == nz) {
                        prd1[IP(i, j, k)] = dz * gz * rho[IC(i, j, k)];
                    } else if (k == nz - 1) {
                        prd1[IP(i, j, k)] = dz * gz * (rho[IC(i, j, k)] + rho[IC(i, j, k + 1)]) * 0.5 +
                                            prd1[IP(i, j, k + 1)];
                    } else {
                        const double intz = dz * (rho[IC(i, j, k)] + 4. * rho[IC(i, j, k + 1)] + rho[IC(i, j, k + 2)]) / 3.;
                        prd1[IP(i, j, k)] = gz * intz + prd1[IP(i, j, k + 2)];
                    }
                }
            }
}

/* posdin's hydrostatic branch, t_press /= 1 (SuLi-IC/10_pos.f90:24-37 and
 * :48-93; dz(k): SuLi-ICdz/10_pos.f90:48-56): the same u, v correction as
 * the other branch, w from continuity column by column from w = 0 at the
 * bottom, then u, v, w zeroed on the faces where the
 * face-interpolated hs (8_derivadas.f90 interp*_cf) is below 1e-4; prd =
 * prd1. */
void oracle_posdin_hydro(const oracle_grid *g, double *u, double *v, double *w, const double *rho,
                         const double *hs, const double *prd1, double *prd)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz,
                 nfz = (size_t)nx * ny * (nz + 1);
    int i, j, k;
    double *hx = malloc(nfx * sizeof(double));
    double *hy = malloc(nfy * sizeof(double));
    double *hz = malloc(nfz * sizeof(double));
    oracle_interp_faces(g, rho, hx, hy, hz);
    posdin_uv(g, u, v, hx, hy, prd1);
    for (j = 0; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) w[IW(i, j, 1)] = 0.;
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                const double dzk = g->dzv ? g->dzv[k] : g->dz;
                w[IW(i, j, k + 1)] = w[IW(i, j, k)] - (u[IU(i + 1, j, k)] - u[IU(i, j, k)]) * dzk / g->dx -
                                     (v[IV(i, j + 1, k)] - v[IV(i, j, k)]) * dzk / g->dy;
            }
    oracle_interp_faces(g, hs, hx, hy, hz);
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx + 1; i++)
                if (hx[IFX(i, j, k)] < 0.0001) u[IU(i, j, k)] = 0.;
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny + 1; j++)
            for (i = 1; i <= nx; i++)
                if (hy[IFY(i, j, k)] < 0.0001) v[IV(i, j, k)] = 0.;
    for (k = 1; k <= nz + 1; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++)
                if (hz[IFZ(i, j, k)] < 0.0001) w[IW(i, j, k)] = 0.;
    free(hx);
    free(hy);
    free(hz);
    memcpy(prd, prd1, (size_t)(nx + 2) * (ny + 2) * (nz + 2) * sizeof(double));
}

/* heaviside with t_tens = 0 (SuLi-IC/14_level_set.f90:555-590, with
 * dx1 = max(dx,dy,dz) from :41 and eps = alpha1 * dx1 supplied by the
 * caller): hs from the level set, then rho and ls_mu as its linear blends.
 * mu and hs may be NULL. */
void oracle_heaviside(int n, const double *ls, double eps, double rho1, double rho2, double mu1, double mu2,
                      double *hs_out, double *rho, double *mu)
{
    const double pi = acos(-1.0);
    for (int c = 0; c < n; c++) {
        double hs;
        if (ls[c] < -eps)
            hs = 0.0;
        else if (ls[c] > eps)
            hs = 1.0;
        else
            hs = 0.5 * (1.0 + ls[c] / eps + 1.0 / pi * sin(pi * ls[c] / eps));
        if (hs_out) hs_out[c] = hs;
        rho[c] = rho1 * (1.0 - hs) + rho2 * hs;
        if (mu) mu[c] = mu1 * (1.0 - hs) + mu2 * hs;
    }
}

void oracle_heaviside_rho(int n, const double *ls, double eps, double rho1, double rho2, double *rho)
{
    oracle_heaviside(n, ls, eps, rho1, rho2, 0.0, 0.0, NULL, rho, NULL);
}

/* ------------------------------------------------------------------------
 * SuLi-IT non-hydrostatic pressure (SuLi-IT/9_grad.f90:10-263, Casulli's
 * semi-implicit scheme).  Same Jacobi-scaled CG as SuLi-IC, with:
 *   - face coefficients from per-face layer thicknesses (:44-62):
 *     sx = tetah dt/dx^2 dzx, sy = tetah dt/dy^2 dzy, sz = tetah dt/dzz,
 *     and no top face (sz(nz+1) = 0);
 *   - the continuity right-hand side in fluxes (:65-81): u dzx and v dzy
 *     differences, w unscaled; the top plane adds the free surface,
 *     d += 1/(gz tetah dt), q -= (eta1 - deltah)/(tetah dt);
 *   - the stop rule |r.r| > 1e-11 on the r.r that entered the previous
 *     iteration (:165-177: the loop tests alfamupr, summed at the top of
 *     the iteration before), and no cap.
 * u, dzx: (0:nx+2,0:ny+1,0:nz+1); v, dzy: (0:nx+1,0:ny+2,0:nz+1); w, dzz:
 * (0:nx+1,0:ny+1,0:nz+2); eta1 (0:nx+1,0:ny+1); deltah (nx,ny); prd1 the
 * halo box.  Ghost fills: fill_ghosts_s (x, then y, then z). */
#define IE(i, j) ((size_t)(i) + (size_t)(nx + 2) * (size_t)(j))
#define IDH(i, j) ((size_t)((i) - 1) + (size_t)nx * (size_t)((j) - 1))

int oracle_it_graddin(const oracle_grid *g, double tetah, double gz, const double *u, const double *v,
                      const double *w, const double *dzx, const double *dzy, const double *dzz,
                      const double *eta1, const double *deltah, double *prd1, oracle_stats *st)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    if (nx < 2 || ny < 2 || nz < 2) return -1;
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz,
                 nfz = (size_t)nx * ny * (nz + 1), ncell = (size_t)nx * ny * nz;
    int i, j, k;
    double *sx = malloc(nfx * sizeof(double));
    double *sy = malloc(nfy * sizeof(double));
    double *sz = malloc(nfz * sizeof(double));
    double *q = malloc(ncell * sizeof(double));
    double *d = calloc(nbox, sizeof(double));
    double *e = calloc(nbox, sizeof(double));
    double *r = calloc(nbox, sizeof(double));
    double *p = calloc(nbox, sizeof(double));
    double *mp = calloc(nbox, sizeof(double));
    if (!sx || !sy || !sz || !q || !d || !e || !r || !p || !mp) {
        free(sx); free(sy); free(sz); free(q); free(d); free(e); free(r); free(p); free(mp);
        return -2;
    }
    const double dt = g->dt, dx = g->dx, dy = g->dy;

    /* face coefficients (:44-62) */
    const double aux1 = tetah * dt / (dx * dx);
    const double aux2 = tetah * dt / (dy * dy);
    for (k = 1; k <= nz; k++) {
        for (j = 1; j <= ny; j++) {
            for (i = 1; i <= nx + 1; i++) sx[IFX(i, j, k)] = aux1 * dzx[IU(i, j, k)];
            for (i = 1; i <= nx; i++) {
                sy[IFY(i, j, k)] = aux2 * dzy[IV(i, j, k)];
                sz[IFZ(i, j, k)] = tetah * dt / dzz[IW(i, j, k)];
            }
        }
        for (i = 1; i <= nx; i++) sy[IFY(i, ny + 1, k)] = aux2 * dzy[IV(i, ny + 1, k)];
    }
    for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) sz[IFZ(i, j, nz + 1)] = 0.;

    /* diagonal and right-hand side (:65-81) */
    for (k = 1; k <= nz - 1; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                d[IP(i, j, k)] = sx[IFX(i + 1, j, k)] + sx[IFX(i, j, k)] + sy[IFY(i, j + 1, k)] + sy[IFY(i, j, k)] +
                                 sz[IFZ(i, j, k + 1)] + sz[IFZ(i, j, k)];
                q[IC(i, j, k)] = (u[IU(i, j, k)] * dzx[IU(i, j, k)] - u[IU(i + 1, j, k)] * dzx[IU(i + 1, j, k)]) / dx +
                                 (v[IV(i, j, k)] * dzy[IV(i, j, k)] - v[IV(i, j + 1, k)] * dzy[IV(i, j + 1, k)]) / dy +
                                 w[IW(i, j, k)] - w[IW(i, j, k + 1)];
            }
    k = nz;
    for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            d[IP(i, j, k)] = sx[IFX(i + 1, j, k)] + sx[IFX(i, j, k)] + sy[IFY(i, j + 1, k)] + sy[IFY(i, j, k)] +
                             sz[IFZ(i, j, k)] + 1. / (gz * tetah * dt);
            q[IC(i, j, k)] = (u[IU(i, j, k)] * dzx[IU(i, j, k)] - u[IU(i + 1, j, k)] * dzx[IU(i + 1, j, k)]) / dx +
                             (v[IV(i, j, k)] * dzy[IV(i, j, k)] - v[IV(i, j + 1, k)] * dzy[IV(i, j + 1, k)]) / dy +
                             w[IW(i, j, k)] - (eta1[IE(i, j)] - deltah[IDH(i, j)]) / (tetah * dt);
        }
    /* ghosts of d (:84-106), e = prd1 sqrt(d) (:108-114) */
    fill_ghosts_s(g, d, 1.0);
    for (size_t c = 0; c < nbox; c++) e[c] = prd1[c] * sqrt(d[c]);

    /* scaled faces and the initial residual (:116-140); a face between
     * cells a and b takes s / sqrt(d_a d_b) from either side */
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                const double dc = d[IP(i, j, k)];
                const double api = sx[IFX(i + 1, j, k)] / sqrt(dc * d[IP(i + 1, j, k)]);
                const double ani = sx[IFX(i, j, k)] / sqrt(dc * d[IP(i - 1, j, k)]);
                const double apj = sy[IFY(i, j + 1, k)] / sqrt(dc * d[IP(i, j + 1, k)]);
                const double anj = sy[IFY(i, j, k)] / sqrt(dc * d[IP(i, j - 1, k)]);
                const double apk = sz[IFZ(i, j, k + 1)] / sqrt(dc * d[IP(i, j, k + 1)]);
                const double ank = sz[IFZ(i, j, k)] / sqrt(dc * d[IP(i, j, k - 1)]);
                r[IP(i, j, k)] = e[IP(i, j, k)] - api * e[IP(i + 1, j, k)] - ani * e[IP(i - 1, j, k)] -
                                 apj * e[IP(i, j + 1, k)] - anj * e[IP(i, j - 1, k)] - apk * e[IP(i, j, k + 1)] -
                                 ank * e[IP(i, j, k - 1)] - q[IC(i, j, k)] / sqrt(dc);
            }
    /* the scaled coefficients replace the unscaled ones for the iteration
     * (the product d_a d_b is the same in either order) */
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx + 1; i++) {
                const double da = d[IP(i - 1, j, k)], db = d[IP(i, j, k)];
                sx[IFX(i, j, k)] = sx[IFX(i, j, k)] / sqrt(db * da);
            }
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny + 1; j++)
            for (i = 1; i <= nx; i++) {
                const double da = d[IP(i, j - 1, k)], db = d[IP(i, j, k)];
                sy[IFY(i, j, k)] = sy[IFY(i, j, k)] / sqrt(db * da);
            }
    for (k = 1; k <= nz + 1; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                const double da = d[IP(i, j, k - 1)], db = d[IP(i, j, k)];
                sz[IFZ(i, j, k)] = sz[IFZ(i, j, k)] / sqrt(db * da);
            }

    double rr = dot_interior(g, r, r, 1);
    double bb = 0.0;
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                const double bi = q[IC(i, j, k)] / sqrt(d[IP(i, j, k)]);
                bb = bb + bi * bi;
            }
    fill_ghosts_s(g, r, 1.0);
    memcpy(p, r, nbox * sizeof(double));

    const double thresh = g->stop_rule == ORACLE_STOP_RELATIVE ? g->tol * g->tol * bb : 0.00000000001;
    const int max_iter = g->max_iter > 0 ? g->max_iter : 10000;
    double alfamu = rr;   /* the loop tests this (:165) */
    int cont = 0;
    if (st) st->rr0 = rr;
    while (fabs(alfamu) > thresh && cont < max_iter) {
        cont = cont + 1;
        /* mp = A p, alfamu = r.r, alfadi = p.mp in one serial sweep (:176-190) */
        double adi = 0.0;
        alfamu = 0.0;
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++) {
                    mp[IP(i, j, k)] = p[IP(i, j, k)] - p[IP(i + 1, j, k)] * sx[IFX(i + 1, j, k)] -
                                      p[IP(i - 1, j, k)] * sx[IFX(i, j, k)] - p[IP(i, j + 1, k)] * sy[IFY(i, j + 1, k)] -
                                      p[IP(i, j - 1, k)] * sy[IFY(i, j, k)] - p[IP(i, j, k + 1)] * sz[IFZ(i, j, k + 1)] -
                                      p[IP(i, j, k - 1)] * sz[IFZ(i, j, k)];
                    alfamu = alfamu + r[IP(i, j, k)] * r[IP(i, j, k)];
                    adi = adi + p[IP(i, j, k)] * mp[IP(i, j, k)];
                }
        const double alpha = alfamu / adi;
        /* e -= alpha p, r -= alpha mp, beta (:194-208) */
        double bmu = 0.0;
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++) {
                    e[IP(i, j, k)] = e[IP(i, j, k)] - alpha * p[IP(i, j, k)];
                    r[IP(i, j, k)] = r[IP(i, j, k)] - alpha * mp[IP(i, j, k)];
                    bmu = bmu + r[IP(i, j, k)] * r[IP(i, j, k)];
                }
        const double beta = bmu / alfamu;
        rr = bmu;
        for (k = 1; k <= nz; k++)
            for (j = 1; j <= ny; j++)
                for (i = 1; i <= nx; i++) p[IP(i, j, k)] = r[IP(i, j, k)] + beta * p[IP(i, j, k)];
        /* ghosts of e and p (:220-253) */
        fill_ghosts_s(g, e, 1.0);
        fill_ghosts_s(g, p, 1.0);
    }
    /* un-scaling on the whole box (:257-263) */
    for (size_t c = 0; c < nbox; c++) prd1[c] = e[c] / sqrt(d[c]);
    if (st) {
        st->iters = cont;
        st->converged = !(fabs(alfamu) > thresh);
        st->bb = bb;
        st->rr = rr;
        st->thresh = thresh;
    }
    free(sx); free(sy); free(sz); free(q); free(d); free(e); free(r); free(p); free(mp);
    return st && !st->converged ? 1 : (fabs(alfamu) > thresh ? 1 : 0);
}

/* SuLi-IT posdin (SuLi-IT/10_pos.f90:73-132, without the layer-thickness
 * update dz_it): eta1 += prd1(:,:,nz)/gz on its whole (0:nx+1,0:ny+1)
 * box (skipped when eta1 is NULL); u, v, w -= tetah dt grad(prd1) with w
 * over dzz; prd = prd1. */
void oracle_it_posdin(const oracle_grid *g, double tetah, double gz, double *u, double *v, double *w,
                      const double *dzz, double *eta1, const double *prd1, double *prd)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    int i, j, k;
    if (eta1)
        for (j = 0; j <= ny + 1; j++)
            for (i = 0; i <= nx + 1; i++) eta1[IE(i, j)] = eta1[IE(i, j)] + prd1[IP(i, j, nz)] / gz;
    const double aux1 = tetah * g->dt / g->dx;
    const double aux2 = tetah * g->dt / g->dy;
    for (k = 1; k <= nz; k++) {
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx + 1; i++)
                u[IU(i, j, k)] = u[IU(i, j, k)] - aux1 * (prd1[IP(i, j, k)] - prd1[IP(i - 1, j, k)]);
        for (j = 1; j <= ny + 1; j++)
            for (i = 1; i <= nx; i++)
                v[IV(i, j, k)] = v[IV(i, j, k)] - aux2 * (prd1[IP(i, j, k)] - prd1[IP(i, j - 1, k)]);
    }
    for (k = 1; k <= nz + 1; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++)
                w[IW(i, j, k)] = w[IW(i, j, k)] - tetah * g->dt / dzz[IW(i, j, k)] * (prd1[IP(i, j, k)] - prd1[IP(i, j, k - 1)]);
    memcpy(prd, prd1, (size_t)(nx + 2) * (ny + 2) * (nz + 2) * sizeof(double));
}

/* dz_it (SuLi-IT/11_diversos.f90:80-141): the face layer thicknesses of the
 * two top planes from the cell thicknesses dz (IP layout).  dzx (IU layout)
 * and dzy (IV) on planes nz and nz+1: face means, the end faces periodic
 * (ccx0 / ccy0 = 0) or linearly extrapolated; dzz (IW) on planes nz and
 * nz+1 the means, nz+2 extrapolated. */
static void it_dz_it(const oracle_grid *g, const double *dz, double *dzx, double *dzy, double *dzz)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    int i, j, k;
    for (k = nz; k <= nz + 1; k++)             /* :89-106 */
        for (j = 0; j <= ny + 1; j++) {
            for (i = 1; i <= nx + 1; i++) dzx[IU(i, j, k)] = (dz[IP(i, j, k)] + dz[IP(i - 1, j, k)]) * 0.5;
            if (g->ccx0 == 0) {
                dzx[IU(0, j, k)] = (dz[IP(0, j, k)] + dz[IP(nx, j, k)]) * 0.5;
                dzx[IU(nx + 2, j, k)] = (dz[IP(1, j, k)] + dz[IP(nx + 1, j, k)]) * 0.5;
            } else {
                dzx[IU(0, j, k)] = dzx[IU(1, j, k)] + (dzx[IU(1, j, k)] - dzx[IU(2, j, k)]);
                dzx[IU(nx + 2, j, k)] = dzx[IU(nx + 1, j, k)] + (dzx[IU(nx + 1, j, k)] - dzx[IU(nx, j, k)]);
            }
        }
    for (k = nz; k <= nz + 1; k++)             /* :108-125 */
        for (i = 0; i <= nx + 1; i++) {
            for (j = 1; j <= ny + 1; j++) dzy[IV(i, j, k)] = (dz[IP(i, j, k)] + dz[IP(i, j - 1, k)]) * 0.5;
            if (g->ccy0 == 0) {
                dzy[IV(i, 0, k)] = (dz[IP(i, 0, k)] + dz[IP(i, ny, k)]) * 0.5;
                dzy[IV(i, ny + 2, k)] = (dz[IP(i, 1, k)] + dz[IP(i, ny + 1, k)]) * 0.5;
            } else {
                dzy[IV(i, 0, k)] = dzy[IV(i, 1, k)] + (dzy[IV(i, 1, k)] - dzy[IV(i, 2, k)]);
                dzy[IV(i, ny + 2, k)] = dzy[IV(i, ny + 1, k)] + (dzy[IV(i, ny + 1, k)] - dzy[IV(i, ny, k)]);
            }
        }
    for (j = 0; j <= ny + 1; j++)              /* :127-138 */
        for (i = 0; i <= nx + 1; i++) {
            dzz[IW(i, j, nz)] = (dz[IP(i, j, nz)] + dz[IP(i, j, nz - 1)]) * 0.5;
            dzz[IW(i, j, nz + 1)] = (dz[IP(i, j, nz + 1)] + dz[IP(i, j, nz)]) * 0.5;
            dzz[IW(i, j, nz + 2)] = dzz[IW(i, j, nz + 1)] + (dzz[IW(i, j, nz + 1)] - dzz[IW(i, j, nz)]);
        }
}

/* SuLi-IT posdin as the reference runs it (SuLi-IT/10_pos.f90:73-132):
 * eta1 += prd1(:,:,nz)/gz and dz(:,:,nz) += eta1 - eta0 over the
 * (0:nx+1,0:ny+1) box (:84-90), dz_it (:93), then the velocity correction
 * with the new dzz (:99-121); prd = prd1. */
void oracle_it_posdin_dz(const oracle_grid *g, double tetah, double gz, double *u, double *v, double *w,
                         double *dz, double *dzx, double *dzy, double *dzz, const double *eta0, double *eta1,
                         const double *prd1, double *prd)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    int i, j;
    for (j = 0; j <= ny + 1; j++)
        for (i = 0; i <= nx + 1; i++) {
            eta1[IE(i, j)] = eta1[IE(i, j)] + prd1[IP(i, j, nz)] / gz;
            dz[IP(i, j, nz)] = dz[IP(i, j, nz)] + (eta1[IE(i, j)] - eta0[IE(i, j)]);
        }
    it_dz_it(g, dz, dzx, dzy, dzz);
    oracle_it_posdin(g, tetah, gz, u, v, w, dzz, NULL, prd1, prd);
}

/* ------------------------------------------------------------------------
 * exp_press, the IPA scheme's pressure increment (SuLi-IC/11_diversos.f90:
 * 455-495; SuLi-ICdz/11_diversos.f90:455-495 divides the z difference by
 * dzz(k)): prd0 = prd1 + prd0 over the halo box (:466), prdaux = prd1 + prd0
 * (:467), the face densities of graddin (:469-471), then
 *   px = -(prdaux(i) - prdaux(i-1))/dx/rhox   on x faces 1..nx+1 (:473-478)
 *   py = -(prdaux(j) - prdaux(j-1))/dy/rhoy   on y faces 1..ny+1 (:480-485)
 *   pz = -(prdaux(k) - prdaux(k-1))/dz/rhoz   on z faces 1..nz+1 (:487-493)
 * px, py, pz in the 1-based face layouts (nx+1,ny,nz), (nx,ny+1,nz),
 * (nx,ny,nz+1). */
void oracle_exp_press(const oracle_grid *g, const double *rho, const double *prd1, double *prd0, double *px,
                      double *py, double *pz)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    double *aux = malloc(nbox * sizeof(double));
    double *rhox = malloc((size_t)(nx + 1) * ny * nz * sizeof(double));
    double *rhoy = malloc((size_t)nx * (ny + 1) * nz * sizeof(double));
    double *rhoz = malloc((size_t)nx * ny * (nz + 1) * sizeof(double));
    int i, j, k;
    for (size_t c = 0; c < nbox; c++) {
        prd0[c] = prd1[c] + prd0[c];
        aux[c] = prd1[c] + prd0[c];
    }
    oracle_interp_faces(g, rho, rhox, rhoy, rhoz);
    for (k = 1; k <= nz; k++) {
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx + 1; i++)
                px[IFX(i, j, k)] = -(aux[IP(i, j, k)] - aux[IP(i - 1, j, k)]) / g->dx / rhox[IFX(i, j, k)];
        for (j = 1; j <= ny + 1; j++)
            for (i = 1; i <= nx; i++)
                py[IFY(i, j, k)] = -(aux[IP(i, j, k)] - aux[IP(i, j - 1, k)]) / g->dy / rhoy[IFY(i, j, k)];
    }
    for (k = 1; k <= nz + 1; k++)
        for (j = 1; j <= ny; j++)
            for (i = 1; i <= nx; i++) {
                const double h = g->dzzv ? g->dzzv[k] : g->dz;
                pz[IFZ(i, j, k)] = -(aux[IP(i, j, k)] - aux[IP(i, j, k - 1)]) / h / rhoz[IFZ(i, j, k)];
            }
    free(aux); free(rhox); free(rhoy); free(rhoz);
}

/* ------------------------------------------------------------------------
 * WENO5 derivatives of the level set along one grid line (SuLi-IC/
 * 8_derivadas.f90:675-811 wenox, :815-951 wenoy, :955-1095 wenoz, which
 * differ only in the direction): ph[m + 2] holds phi1(m), m = -2..n+3, with
 * phi1(1..n) the line's values; the ends are filled here, periodic (per)
 * as :702-714 or, otherwise, extrapolated with zero derivative as
 * :742-756.  dp, dn (1..n at [0..n-1]) get the two one-sided derivatives
 * (dphidxp, dphidxn). */
static void weno_line(double *ph, int n, double h, int per, double *dp, double *dn)
{
#define PH(m) ph[(m) + 2]
    const double c11 = 1. / 11.;
    if (per) {
        PH(0) = PH(n - 1);
        PH(-1) = PH(n - 2);
        PH(-2) = PH(n - 3);
        PH(n + 1) = PH(2);
        PH(n + 2) = PH(3);
        PH(n + 3) = PH(4);
    } else {
        PH(0) = c11 * (18. * PH(1) - 9. * PH(2) + 2. * PH(3));
        PH(-1) = c11 * (18. * PH(0) - 9. * PH(1) + 2. * PH(2));
        PH(-2) = c11 * (18. * PH(-1) - 9. * PH(0) + 2. * PH(1));
        PH(n + 1) = c11 * (18. * PH(n) - 9. * PH(n - 1) + 2. * PH(n - 2));
        PH(n + 2) = c11 * (18. * PH(n + 1) - 9. * PH(n) + 2. * PH(n - 1));
        PH(n + 3) = c11 * (18. * PH(n + 2) - 9. * PH(n + 1) + 2. * PH(n));
    }
    /* up(m) = (phi1(m+3)-phi1(m+2))/h, m = -3..n (:759-765);
     * un(m) = (phi1(m-2)-phi1(m-3))/h, m = 1..n+4 (:767-773) */
    double *up = malloc((size_t)(n + 4) * sizeof(double)); /* up[m + 3] */
    double *un = malloc((size_t)(n + 4) * sizeof(double)); /* un[m - 1] */
    for (int m = -3; m <= n; m++) up[m + 3] = (PH(m + 3) - PH(m + 2)) / h;
    for (int m = 1; m <= n + 4; m++) un[m - 1] = (PH(m - 2) - PH(m - 3)) / h;
#define UP(m) up[(m) + 3]
#define UN(m) un[(m) - 1]
    const double a1 = 13. / 12., a2 = 1. / 4., a3 = 1. / 6., eps = 0.00000001;
    const double wgt[3] = {0.1, 0.6, 0.3};
    for (int i = 1; i <= n; i++) {   /* :775-809 */
        double isp[3], isn[3], alp[3], aln[3];
        double t;
        t = UP(i) - 2 * UP(i - 1) + UP(i - 2);
        isp[0] = a1 * t * t;
        t = UP(i) - 4 * UP(i - 1) + 3 * UP(i - 2);
        isp[0] = isp[0] + a2 * t * t;
        t = UN(i) - 2 * UN(i + 1) + UN(i + 2);
        isn[0] = a1 * t * t;
        t = UN(i) - 4 * UN(i + 1) + 3 * UN(i + 2);
        isn[0] = isn[0] + a2 * t * t;
        t = UP(i - 1) - 2 * UP(i - 2) + UP(i - 3);
        isp[1] = a1 * t * t;
        t = UP(i - 1) - UP(i - 3);
        isp[1] = isp[1] + a2 * t * t;
        t = UN(i + 1) - 2 * UN(i + 2) + UN(i + 3);
        isn[1] = a1 * t * t;
        t = UN(i + 1) - UN(i + 3);
        isn[1] = isn[1] + a2 * t * t;
        t = UP(i - 2) - 2 * UP(i - 3) + UP(i - 4);
        isp[2] = a1 * t * t;
        t = 3 * UP(i - 2) - 4 * UP(i - 3) + UP(i - 4);
        isp[2] = isp[2] + a2 * t * t;
        t = UN(i + 2) - 2 * UN(i + 3) + UN(i + 4);
        isn[2] = a1 * t * t;
        t = 3 * UN(i + 2) - 4 * UN(i + 3) + UN(i + 4);
        isn[2] = isn[2] + a2 * t * t;
        for (int q = 0; q < 3; q++) {
            alp[q] = wgt[q] / ((eps + isp[q]) * (eps + isp[q]));
            aln[q] = wgt[q] / ((eps + isn[q]) * (eps + isn[q]));
        }
        const double sp = alp[0] + alp[1] + alp[2], sn = aln[0] + aln[1] + aln[2];
        const double op0 = alp[0] / sp, op1 = alp[1] / sp, op2 = alp[2] / sp;
        const double on0 = aln[0] / sn, on1 = aln[1] / sn, on2 = aln[2] / sn;
        dp[i - 1] = a3 * (op0 * (2 * UP(i) - 7 * UP(i - 1) + 11 * UP(i - 2)) +
                          op1 * (-UP(i - 1) + 5 * UP(i - 2) + 2 * UP(i - 3)) +
                          op2 * (2 * UP(i - 2) + 5 * UP(i - 3) - UP(i - 4)));
        dn[i - 1] = a3 * (on0 * (2 * UN(i) - 7 * UN(i + 1) + 11 * UN(i + 2)) +
                          on1 * (-UN(i + 1) + 5 * UN(i + 2) + 2 * UN(i + 3)) +
                          on2 * (2 * UN(i + 2) + 5 * UN(i + 3) - UN(i + 4)));
    }
#undef UP
#undef UN
#undef PH
    free(up);
    free(un);
}

/* The level set's WENO5 derivative of larger magnitude per direction, as
 * heaviside's t_tens = 1 branch forms it (SuLi-IC/14_level_set.f90:519-544
 * over der_weno, :338-370): x and y lines periodic when ccx0 / ccy0 = 0,
 * else extrapolated; z lines always extrapolated.  gx, gy, gz (nx,ny,nz). */
void oracle_ls_gradient(const oracle_grid *g, const double *ls, double *gx, double *gy, double *gzo)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
    double *ph = malloc((size_t)(nmax + 6) * sizeof(double));
    double *dp = malloc((size_t)nmax * sizeof(double));
    double *dn = malloc((size_t)nmax * sizeof(double));
    int i, j, k;
    /* |ta1| > |td1| keeps ta1, else td1 (:523-527) */
#define PICK(a, b) (fabs(a) > fabs(b) ? (a) : (b))
    for (k = 1; k <= nz; k++)
        for (j = 1; j <= ny; j++) {
            for (i = 1; i <= nx; i++) ph[i + 2] = ls[IC(i, j, k)];
            weno_line(ph, nx, g->dx, g->ccx0 == 0, dp, dn);
            for (i = 1; i <= nx; i++) gx[IC(i, j, k)] = PICK(dp[i - 1], dn[i - 1]);
        }
    for (k = 1; k <= nz; k++)
        for (i = 1; i <= nx; i++) {
            for (j = 1; j <= ny; j++) ph[j + 2] = ls[IC(i, j, k)];
            weno_line(ph, ny, g->dy, g->ccy0 == 0, dp, dn);
            for (j = 1; j <= ny; j++) gy[IC(i, j, k)] = PICK(dp[j - 1], dn[j - 1]);
        }
    for (j = 1; j <= ny; j++)
        for (i = 1; i <= nx; i++) {
            for (k = 1; k <= nz; k++) ph[k + 2] = ls[IC(i, j, k)];
            weno_line(ph, nz, g->dz, 0, dp, dn);
            for (k = 1; k <= nz; k++) gzo[IC(i, j, k)] = PICK(dp[k - 1], dn[k - 1]);
        }
#undef PICK
    free(ph);
    free(dp);
    free(dn);
}

/* heaviside with t_tens = 1 (SuLi-IC/14_level_set.f90:508-594): hs, rho
 * and ls_mu as oracle_heaviside, and the surface-tension gradients hsx,
 * hsy, hsz: 0.5*g*(1 + cos(pi*ls/eps))/eps inside the band (:563-567) with
 * g from oracle_ls_gradient, 0 outside (:546-548); eps = alpha1*dx1. */
void oracle_heaviside_tens(const oracle_grid *g, const double *ls, double eps, double rho1, double rho2,
                           double mu1, double mu2, double *hs, double *rho, double *mu, double *hsx,
                           double *hsy, double *hsz)
{
    const int n = g->nx * g->ny * g->nz;
    const double pi = acos(-1.0);
    double *tx = malloc((size_t)n * sizeof(double));
    double *ty = malloc((size_t)n * sizeof(double));
    double *tz = malloc((size_t)n * sizeof(double));
    oracle_ls_gradient(g, ls, tx, ty, tz);
    oracle_heaviside(n, ls, eps, rho1, rho2, mu1, mu2, hs, rho, mu);
    for (int c = 0; c < n; c++) {
        hsx[c] = hsy[c] = hsz[c] = 0.;
        if (!(ls[c] < -eps) && !(ls[c] > eps)) {
            const double cs = 1. + cos(pi * ls[c] / eps);
            hsx[c] = 0.5 * tx[c] * cs / eps;
            hsy[c] = 0.5 * ty[c] * cs / eps;
            hsz[c] = 0.5 * tz[c] * cs / eps;
        }
    }
    free(tx);
    free(ty);
    free(tz);
}
