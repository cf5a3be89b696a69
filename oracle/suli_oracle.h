/*
 * CPU restatement of SuLi's pressure step (graddin + posdin).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product in suli_amd/ and the host-core CPU baseline in bench.py.  Nothing
 * in the product links, loads or calls it; only tests/, the smoke check
 * and bench.py's cpu_baseline leg do.
 *
 * All arrays use the reference's Fortran layouts (column-major, x fastest):
 *   u    (0:nx+2, 0:ny+1, 0:nz+1)      SuLi-IC/0_modules.f90:170
 *   v    (0:nx+1, 0:ny+2, 0:nz+1)      :171
 *   w    (0:nx+1, 0:ny+1, 0:nz+2)      :172
 *   prd1 (0:nx+1, 0:ny+1, 0:nz+1)      :188
 *   rho  (1:nx, 1:ny, 1:nz)            :189
 *   obs_lsx (nx+1,ny,nz), obs_lsy (nx,ny+1,nz), obs_lsz (nx,ny,nz+1)  :135-137
 *   dz, dzz (0:nz+1)                   SuLi-ICdz/0_modules.f90:72-73
 */
#ifndef SULI_ORACLE_H
#define SULI_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_STOP_REFERENCE = 0, ORACLE_STOP_RELATIVE = 1 };

typedef struct oracle_grid {
    int nx, ny, nz;
    double dx, dy, dz;      /* dz is used when dzv == NULL (SuLi-IC) */
    const double *dzv;      /* dz(0:nz+1), SuLi-ICdz; NULL for uniform */
    const double *dzzv;     /* dzz(0:nz+1), SuLi-ICdz; NULL for uniform */
    double dt;
    int ccx0, ccy0;         /* 0 = periodic in that direction */
    int t_press;            /* 1 = projection, 2 = IPA (faces use rho_f1) */
    double rho_f1;
    int ibm_mask;           /* obst_t != 0 && ibm_t == 2 */
    const double *obsx, *obsy, *obsz;
    int omp_t;              /* 0 = serial k-j-i dot products, 1 = OpenMP reductions */
    int stop_rule;          /* ORACLE_STOP_* */
    double tol;             /* relative tolerance for ORACLE_STOP_RELATIVE */
    int max_iter;           /* reference: 10000 */
    int threads;            /* OpenMP threads for the sweeps; <= 0 = runtime default */
    int project_null;       /* not in the reference: remove the initial residual's
                               component along the null vector sqrt(d) */
    const unsigned char *solid; /* not in the reference: cells (nx,ny,nz, x fastest)
                                   masked out; faces touching one carry no flux;
                                   NULL = none */
    int top_dirichlet;      /* not in the reference: p = 0 on the top face (ghost =
                               -interior for the pressure-like arrays) instead of
                               the zero-flux lid */
} oracle_grid;

typedef struct oracle_stats {
    int iters;          /* CG iterations performed (reference `cont`) */
    int converged;      /* 1 if the stop rule was met before max_iter */
    double rr0;         /* r.r of the initial scaled residual */
    double bb;          /* b.b of the scaled right-hand side */
    double rr;          /* final iterated r.r (reference `alfamupr`) */
    double thresh;      /* r.r threshold that was used */
} oracle_stats;

/* SuLi-IC/9_grad.f90:10-663 (SuLi-ICdz/9_grad.f90 when dzv != NULL).
 * prd1 is read (initial guess, full halo box) and written (full box).
 * Returns 0 when converged, 1 when max_iter was hit, <0 on bad input. */
int oracle_graddin(const oracle_grid *g, const double *u, const double *v,
                   const double *w, const double *rho, double *prd1,
                   oracle_stats *st);

/* SuLi-IC/10_pos.f90:17-46,95 (t_press == 1 branch): u, v, w updated in
 * place on every face the reference loops over; prd = prd1 (full box). */
void oracle_posdin(const oracle_grid *g, double *u, double *v, double *w,
                   const double *rho, const double *prd1, double *prd);

/* SuLi-IC/8_derivadas.f90:305-357, :381-418, :443-481 (SuLi-ICdz z variant
 * :450-489 when dzv != NULL).  Outputs are 1-based face arrays laid out as
 * (nx+1,ny,nz), (nx,ny+1,nz), (nx,ny,nz+1). */
void oracle_interp_faces(const oracle_grid *g, const double *rho,
                         double *rhox, double *rhoy, double *rhoz);

/* Scaled operator pieces, exposed so that tests can assemble the matrix:
 * face coefficients after symmetric Jacobi scaling (9_grad.f90:256-288),
 * diagonal d with mirror/periodic ghosts (:152-160, :188-239) and the
 * scaled right-hand side b = q / sqrt(d) (:297).  Face arrays as in
 * oracle_interp_faces; d is (0:nx+1,0:ny+1,0:nz+1); b and q are (nx,ny,nz). */
void oracle_scaled_operator(const oracle_grid *g, const double *u,
                            const double *v, const double *w,
                            const double *rho, double *shx, double *shy,
                            double *shz, double *d, double *q, double *b);

/* Smoothed Heaviside density, SuLi-IC/14_level_set.f90:41,555-587:
 * rho = rho1 (1-H) + rho2 H with H(ls) over half-width eps. */
/* SuLi-IC/9_grad.f90:667-702 (SuLi-ICdz with dzv): hydrostatic prd1
 * interior from rho, gz (module param) and dz; the halo is not touched. */
void oracle_pressh(const oracle_grid *g, const double *rho, double gz, double *prd1);

/* SuLi-IC/10_pos.f90:24-37 then :48-93 (t_press /= 1 branch): u, v
 * corrected from prd1 and rho, w from continuity, then u, v, w zeroed where
 * hs (nx,ny,nz) is below 1e-4 on the face; prd = prd1. */
void oracle_posdin_hydro(const oracle_grid *g, double *u, double *v, double *w, const double *rho,
                         const double *hs, const double *prd1, double *prd);

/* SuLi-IC/14_level_set.f90:555-590 (t_tens = 0): hs, rho and ls_mu from
 * the level set; hs_out and mu may be NULL. */
void oracle_heaviside(int n, const double *ls, double eps, double rho1, double rho2, double mu1, double mu2,
                      double *hs_out, double *rho, double *mu);
void oracle_heaviside_rho(int n, const double *ls, double eps, double rho1,
                          double rho2, double *rho);

/* SuLi-IT/9_grad.f90:10-263: the non-hydrostatic pressure of SuLi-IT
 * (coefficients from per-face layer thicknesses, free-surface top row,
 * stop rule |r.r| > 1e-11 on the previous iteration's r.r).  g supplies
 * nx..dt, ccx0, ccy0, stop_rule, tol, max_iter; u and dzx in u's layout,
 * v and dzy in v's, w and dzz in w's, eta1 (0:nx+1,0:ny+1), deltah
 * (nx,ny).  Returns 0 when converged, 1 at max_iter. */
int oracle_it_graddin(const oracle_grid *g, double tetah, double gz, const double *u, const double *v,
                      const double *w, const double *dzx, const double *dzy, const double *dzz,
                      const double *eta1, const double *deltah, double *prd1, oracle_stats *st);

/* SuLi-IT/10_pos.f90:73-132 without dz_it: eta1 += prd1(:,:,nz)/gz (eta1
 * may be NULL: skipped); u, v, w corrected in place; prd = prd1. */
void oracle_it_posdin(const oracle_grid *g, double tetah, double gz, double *u, double *v, double *w,
                      const double *dzz, double *eta1, const double *prd1, double *prd);

/* SuLi-IT/10_pos.f90:73-132 with dz_it (SuLi-IT/11_diversos.f90:80-141):
 * eta1 and dz(:,:,nz) from prd1 and eta0, the top planes' face thicknesses
 * (dzx, dzy, dzz), then the correction above with the new dzz.  dz in
 * prd1's layout, dzx in u's, dzy in v's, dzz in w's, eta0/eta1
 * (0:nx+1,0:ny+1). */
void oracle_it_posdin_dz(const oracle_grid *g, double tetah, double gz, double *u, double *v, double *w,
                         double *dz, double *dzx, double *dzy, double *dzz, const double *eta0, double *eta1,
                         const double *prd1, double *prd);

/* SuLi-IC/11_diversos.f90:455-495 (SuLi-ICdz with dzv/dzzv): prd0 += prd1
 * over the halo box; px, py, pz = -grad(prd1 + prd0)/rho_face on the face
 * layouts (nx+1,ny,nz), (nx,ny+1,nz), (nx,ny,nz+1). */
void oracle_exp_press(const oracle_grid *g, const double *rho, const double *prd1, double *prd0, double *px,
                      double *py, double *pz);

/* The level set's WENO5 derivatives as heaviside's t_tens = 1 branch picks
 * them (SuLi-IC/14_level_set.f90:338-370,519-544; 8_derivadas.f90:675-1095):
 * per direction the one-sided derivative of larger magnitude.  Uses nx..dz,
 * ccx0, ccy0.  All arrays (nx,ny,nz). */
void oracle_ls_gradient(const oracle_grid *g, const double *ls, double *gx, double *gy, double *gz);

/* SuLi-IC/14_level_set.f90:508-594 with t_tens = 1: oracle_heaviside plus
 * hsx, hsy, hsz (nx,ny,nz); eps = alpha1*dx1; hs and mu may be NULL. */
void oracle_heaviside_tens(const oracle_grid *g, const double *ls, double eps, double rho1, double rho2,
                           double mu1, double mu2, double *hs, double *rho, double *mu, double *hsx,
                           double *hsy, double *hsz);

#ifdef __cplusplus
}
#endif
#endif
