/*
 * suli_ps: MI355X (gfx950) pressure-projection solver for SuLi.
 *
 * C ABI that replaces the reference's pressure step.  SuLi's time loop calls
 *     CALL graddin()   (SuLi-IC/1_principal.f90:102, SuLi-IC/9_grad.f90:10-663)
 *     CALL posdin()    (SuLi-IC/1_principal.f90:103, SuLi-IC/10_pos.f90:1-99)
 * with every operand in module variables.  suli_amd/fortran/suli_ps_graddin.f90
 * keeps those argument-less signatures and forwards the module arrays here.
 *
 * Array layouts are the reference's (Fortran column-major, x fastest, lower
 * bounds as declared in SuLi-IC/0_modules.f90:135-137,170-172,188-189):
 *   u    (0:nx+2, 0:ny+1, 0:nz+1)       v (0:nx+1, 0:ny+2, 0:nz+1)
 *   w    (0:nx+1, 0:ny+1, 0:nz+2)       prd1, prd (0:nx+1, 0:ny+1, 0:nz+1)
 *   rho  (1:nx, 1:ny, 1:nz)
 *   obs_lsx (nx+1,ny,nz)  obs_lsy (nx,ny+1,nz)  obs_lsz (nx,ny,nz+1)
 *   dz, dzz (0:nz+1)  (SuLi-ICdz/0_modules.f90:72-73; NULL on uniform grids)
 *
 * Return codes: 0 success, 1 not converged within max_iter (the reference
 * prints "pulou pressão" and continues, 9_grad.f90:590), < 0 error (message
 * from suli_ps_last_error).  A context is used by one host thread at a time.
 */
#ifndef SULI_PS_H
#define SULI_PS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SULI_PS_ABI_VERSION 2

enum suli_ps_precision {
    SULI_PS_FP64 = 0,   /* FP64 CG, as the reference */
    SULI_PS_MIXED = 1   /* FP32 SpMV/axpy inside FP64 residual replacement */
};

enum suli_ps_stop_rule {
    SULI_PS_STOP_REFERENCE = 0, /* r.r <= 1e-4/(nx*ny*nz), 9_grad.f90:393,398 */
    SULI_PS_STOP_RELATIVE = 1   /* ||r|| <= tol * ||b|| on the scaled system   */
};

enum suli_ps_status {
    SULI_PS_OK = 0,
    SULI_PS_NOT_CONVERGED = 1,
    SULI_PS_EINVAL = -1,
    SULI_PS_EHIP = -2,
    SULI_PS_ENOMEM = -3,
    SULI_PS_EAGAIN = -4     /* server: the request is not finished yet */
};

/* Scalars graddin/posdin read from SuLi's modules: disc (nx..dt, t_press,
 * obst_t, ibm_t; SuLi-IC/0_modules.f90:2-58), cond (ccx0, ccy0; :68-80),
 * ls_param (rho_f1; :282-295). */
typedef struct suli_ps_grid {
    int32_t nx, ny, nz;
    int32_t ccx0, ccy0;     /* 0 = periodic in x / y */
    int32_t t_press;        /* 1 = projection; 2 = IPA, faces use rho_f1 (9_grad.f90:109-116) */
    int32_t obst_t, ibm_t;  /* obst_t != 0 && ibm_t == 2: masked RHS (9_grad.f90:45-88) */
    int32_t top_bc;         /* 0: zero-flux lid, as the reference; 1: p = 0 on the top
                               face (an extension: ghost = -interior, w free there) */
    double dx, dy, dz, dt;  /* dz unused when a dz(0:nz+1) array is passed */
    double rho_f1;
} suli_ps_grid;

typedef struct suli_ps_opts {
    int32_t precision;     /* enum suli_ps_precision */
    int32_t stop_rule;     /* enum suli_ps_stop_rule */
    double tol;            /* relative tolerance for SULI_PS_STOP_RELATIVE */
    int32_t max_iter;      /* <= 0: 10000, the reference cap (9_grad.f90:398) */
    int32_t chunk;         /* CG iterations between host polls; <= 0: automatic */
    int32_t verbose;       /* 1: print the reference's non-convergence line */
    int32_t timing;        /* 1: time CG steps with HIP events (stats) */
    int32_t keep_null;     /* 1: keep the initial residual's component along the
                              operator's null space (constant pressure), as the
                              reference does; 0 (default): remove it, which
                              changes the solution by a constant only */
    int32_t no_graphs;     /* 1: launch every kernel from the host instead of
                              replaying captured HIP graphs (same results) */
    int32_t autotune;      /* 1: pick each grid size's tile depth by timing the
                              CG step kernel on its first solve (see
                              suli_ps_autotune); 0: a fixed rule */
    int32_t reuse_staged;  /* 1: suli_ps_posdin on host pointers, called right
                              after suli_ps_graddin on the same grid with the
                              same u, v, w, rho and prd1 pointers, takes those
                              arrays from the device copies graddin left (the
                              inputs as uploaded, prd1 as solved) instead of
                              uploading them again.  The caller promises not
                              to change them in between, as SuLi's loop does
                              (SuLi-IC/1_principal.f90:102-103).  0: upload. */
    int32_t stored_residual; /* mixed precision: 1 = each FP32 step stores its
                              residual r_{n-1} (two FP32 vectors alternate)
                              and the next step reads it; 0 (default) = the
                              step recovers r_{n-2} from the directions,
                              r_{n-2} = p_{n-1} - beta_{n-1} p_{n-2}, and
                              stores no residual.  Both are CG; they differ
                              in FP32 rounding (one more word written per
                              cell and step for stored) */
} suli_ps_opts;

typedef struct suli_ps_stats {
    int32_t status;        /* enum suli_ps_status */
    int32_t iters;         /* CG iterations (reference `cont`) */
    int32_t refreshes;     /* FP64 residual replacements (mixed) */
    int32_t spmv_timed;    /* timing windows measured (opts.timing) */
    double rr0;            /* r.r of the initial scaled residual */
    double bb;             /* b.b of the scaled right-hand side */
    double rr;             /* final r.r the stop rule saw */
    double thresh;         /* the stop rule's threshold on r.r */
    double rel_true;       /* FP64 ||b - A e|| / ||b|| recomputed at exit */
    double solve_ms;       /* device time of the solve, setup included */
    double spmv_ms;        /* mean time of one CG step (its launches, SpMV
                              included) over windows of 8 consecutive steps,
                              HIP events on the context stream (opts.timing) */
    int32_t graph_replays; /* captured HIP graphs launched during the call (each
                              runs 8 CG steps); 0 when every kernel was
                              launched directly.  0 for served requests. */
    int32_t reserved;
} suli_ps_stats;

typedef struct suli_ps_ctx suli_ps_ctx;

/* Context on the current HIP device with its own stream (default priority).
 * opts may be NULL (FP64, reference stop rule).  Returns NULL on failure. */
suli_ps_ctx *suli_ps_create(const suli_ps_opts *opts);
void suli_ps_destroy(suli_ps_ctx *ctx);
int suli_ps_set_opts(suli_ps_ctx *ctx, const suli_ps_opts *opts);
const char *suli_ps_last_error(const suli_ps_ctx *ctx);
int suli_ps_abi_version(void);

/* graddin (SuLi-IC/9_grad.f90:10-663; SuLi-ICdz/9_grad.f90 when dz/dzz are
 * given): solves -div(dt/rho grad p) = -div u* with the reference's
 * symmetric Jacobi scaling and CG, warm-started from prd1.  prd1 is read
 * (initial guess, whole halo box) and written (whole halo box).
 * obs_ls* are read only when obst_t != 0 && ibm_t == 2.  Host pointers. */
int suli_ps_graddin(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *u,
                    const double *v, const double *w, const double *rho,
                    const double *obs_lsx, const double *obs_lsy,
                    const double *obs_lsz, const double *dz, const double *dzz,
                    double *prd1, suli_ps_stats *stats);

/* posdin, t_press == 1 branch (SuLi-IC/10_pos.f90:17-46,95; SuLi-ICdz/
 * 10_pos.f90:21,41): u, v, w -= dt grad(prd1) / rho_face on the faces the
 * reference updates; prd = prd1 (prd may be NULL).  Host pointers. */
int suli_ps_posdin(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v,
                   double *w, const double *rho, const double *prd1,
                   const double *dz, const double *dzz, double *prd);

/* The same two operations on device-resident arrays (same layouts, device
 * pointers), enqueued on the context's stream.  graddin_dev returns once
 * the solve has finished. */
int suli_ps_graddin_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *u,
                        const double *v, const double *w, const double *rho,
                        const double *obs_lsx, const double *obs_lsy,
                        const double *obs_lsz, const double *dz, const double *dzz,
                        double *prd1, suli_ps_stats *stats);
int suli_ps_posdin_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v,
                       double *w, const double *rho, const double *prd1,
                       const double *dz, const double *dzz, double *prd);

/* suli_ps_graddin_dev without allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (tight serving loops,
 * harnesses that forbid hipMalloc, code that must not synchronise on an
 * allocation).  Same arguments, options, arithmetic and results as
 * suli_ps_graddin_dev, except that the problem's solver state (scaled face
 * coefficients, diagonal, right-hand side, iterate, Krylov vectors, partial
 * sums) lives in one fixed device array of SULI_PS_NOALLOC_MAX_BYTES per GPU,
 * which the library holds from the first context on that GPU until the
 * process ends.  The call allocates and frees nothing, creates no event and
 * captures no graph (it launches every kernel directly); the launch tables
 * and host lists it needs are reserved by suli_ps_create.  A problem whose
 * state does not fit (problem bytes beyond the maximum: about 80 B per cell
 * in FP64, 104 B in mixed precision, so up to about 0.8 M / 0.64 M cells) is
 * rejected with SULI_PS_EINVAL and a message giving both sizes; with
 * opts.autotune it uses a tile shape already tuned for the size, or the
 * default rule, and never tunes.  The array keeps the last call's contents
 * (every value a solve reads is written by its own setup first).  Every
 * call on one GPU shares that GPU's array: calls that overlap in time on
 * one GPU, from any contexts or threads, race and must be serialised by
 * the caller.  After the call the context's last-solve state refers to the
 * array (suli_ps_time_spmv reads it). */
#define SULI_PS_NOALLOC_MAX_BYTES ((size_t)64 << 20)
int suli_ps_graddin_noalloc_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *u,
                                const double *v, const double *w, const double *rho,
                                const double *obs_lsx, const double *obs_lsy,
                                const double *obs_lsz, const double *dz, const double *dzz,
                                double *prd1, suli_ps_stats *stats);

/* graddin / posdin with a solid-cell mask, an extension with no reference
 * counterpart (the reference's IBM only masks the right-hand side,
 * 9_grad.f90:45-88).  solid (nx*ny*nz bytes, x fastest, device; NULL = none)
 * marks cells that are masked out: every face with a solid cell on either
 * side (ghosts by the boundary rule) carries no flux, in the operator, in
 * the divergence and in the correction; solid cells get pressure 0.
 * Otherwise as suli_ps_graddin_dev / _posdin_dev. */
int suli_ps_graddin_masked_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *u,
                               const double *v, const double *w, const double *rho,
                               const double *obs_lsx, const double *obs_lsy,
                               const double *obs_lsz, const double *dz, const double *dzz,
                               const uint8_t *solid, double *prd1, suli_ps_stats *stats);
int suli_ps_posdin_masked_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v,
                              double *w, const double *rho, const double *prd1,
                              const double *dz, const double *dzz, const uint8_t *solid,
                              double *prd);

/* Hydrostatic mode, t_press == 0 (SuLi-IC/1_principal.f90:97-100 calls
 * pressh() then posdin()).
 *
 * pressh (SuLi-IC/9_grad.f90:667-702; SuLi-ICdz/9_grad.f90:665-700 when dz
 * is given): prd1's interior becomes the column integral of gz*rho from the
 * top (dz*gz*rho in the top cell, the trapezoid of the top two cells below
 * it, Simpson's rule over two cells elsewhere); prd1's halo is not written.
 * gz is the reference's module value (9.80665*cos(atan(decliv)),
 * SuLi-IC/2_iniciais.f90:110).  t_press and dt are not read.
 *
 * posdin_hydro, posdin's t_press != 1 branch (SuLi-IC/10_pos.f90:24-37,
 * 48-95; SuLi-ICdz/10_pos.f90 with dz): u, v -= dt grad(prd1)/rho_face as
 * in suli_ps_posdin; w(:,:,1) = 0 and w from continuity up every column;
 * then u, v, w set to 0 on the faces where the face value of hs (the
 * Heaviside field, (nx,ny,nz) like rho) is below 1e-4; prd = prd1 (prd may
 * be NULL).  The reference's posdin reads t_press to pick this branch;
 * here the caller picks the entry point.
 *
 * Host pointers; the _dev forms take device pointers and run on the
 * context's stream. */
int suli_ps_pressh(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *rho,
                   const double *dz, double gz, double *prd1);
int suli_ps_pressh_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *rho,
                       const double *dz, double gz, double *prd1);
int suli_ps_posdin_hydro(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v,
                         double *w, const double *rho, const double *hs,
                         const double *prd1, const double *dz, double *prd);
int suli_ps_posdin_hydro_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v,
                             double *w, const double *rho, const double *hs,
                             const double *prd1, const double *dz, double *prd);

/* Density from the level set on the device: heaviside with t_tens = 0
 * (SuLi-IC/14_level_set.f90:508-594; suli_ps_heaviside_tens_dev adds the
 * t_tens = 1 surface-tension gradients).  For each of the nx*ny*nz cells
 * (rho's layout): hs = 0 below -eps, 1 above eps, else
 * 0.5*(1 + ls/eps + sin(pi*ls/eps)/pi); rho = rho_f1*(1-hs) + rho_f2*hs
 * with rho_f1 from the grid; mu = mu_f1*(1-hs) + mu_f2*hs (the reference's
 * ls_mu).  eps = alpha1*dx1, dx1 = max(dx,dy,dz) (:41).  hs and mu may be
 * NULL.  Device pointers, context stream; returns once written.  With it a
 * caller holding the level set feeds suli_ps_graddin_dev directly. */
int suli_ps_heaviside_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *ls,
                          double eps, double rho_f2, double mu_f1, double mu_f2,
                          double *hs, double *rho, double *mu);

/* heaviside with t_tens = 1 (SuLi-IC/14_level_set.f90:508-594): as
 * suli_ps_heaviside_dev, and the surface-tension gradients hsx, hsy, hsz
 * (nx,ny,nz): inside the band |ls| <= eps, 0.5*g*(1 + cos(pi*ls/eps))/eps
 * with g the level set's WENO5 derivative along that axis, of the two
 * one-sided ones the one of larger magnitude (der_weno, :338-370, :519-544;
 * SuLi-IC/8_derivadas.f90:675-1095); 0 outside.  The line ends are
 * periodic in x (y) when ccx0 (ccy0) is 0, else extrapolated with zero
 * derivative; z always the latter.  Grid: nx..nz (each >= 4), dx, dy, dz
 * (uniform), ccx0, ccy0, rho_f1.  hs and mu may be NULL.  Device pointers,
 * context stream; returns once written. */
int suli_ps_heaviside_tens_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *ls,
                               double eps, double rho_f2, double mu_f1, double mu_f2,
                               double *hs, double *rho, double *mu, double *hsx, double *hsy,
                               double *hsz);

/* exp_press, the IPA scheme's pressure increment at the start of a time step
 * (SuLi-IC/11_diversos.f90:455-495, called at SuLi-IC/1_principal.f90:82
 * when t_press = 2; SuLi-ICdz/11_diversos.f90:455-495 when dz is given):
 * prd0 = prd1 + prd0 over the whole halo box, then with
 * prdaux = prd1 + prd0 and the face densities of graddin (interp*_cf,
 * dz-weighted in z when dz is given)
 *   px = -(prdaux(i) - prdaux(i-1))/dx/rho_x   on the (nx+1,ny,nz) x faces,
 *   py = -(prdaux(j) - prdaux(j-1))/dy/rho_y   on the (nx,ny+1,nz) y faces,
 *   pz = -(prdaux(k) - prdaux(k-1))/dzz(k)/rho_z  (dz of the grid when dzz
 *        is NULL) on the (nx,ny,nz+1) z faces.
 * rho (nx,ny,nz), prd1 and prd0 the halo box; dz, dzz (0:nz+1) together or
 * both NULL.  Grid: nx..nz, dx, dy, dz, ccx0.  Host pointers; the _dev form
 * takes device pointers and runs on the context's stream. */
int suli_ps_exp_press(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *rho,
                      const double *prd1, double *prd0, const double *dz, const double *dzz,
                      double *px, double *py, double *pz);
int suli_ps_exp_press_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *rho,
                          const double *prd1, double *prd0, const double *dz, const double *dzz,
                          double *px, double *py, double *pz);

/* SuLi-IT's non-hydrostatic pressure (Casulli's semi-implicit scheme):
 * graddin, SuLi-IT/9_grad.f90:10-263, the same Jacobi-scaled CG with
 *   sx = tetah dt/dx^2 dzx, sy = tetah dt/dy^2 dzy, sz = tetah dt/dzz on the
 *   faces (no face above the top plane), the continuity right-hand side in
 *   fluxes (u dzx, v dzy, w), and on the top plane the free surface:
 *   d += 1/(gz tetah dt), q -= (eta1 - deltah)/(tetah dt).
 * Grid: nx..nz, dx, dy, dt, ccx0, ccy0 (dz, t_press, obst_t, ibm_t and
 * rho_f1 are not read; top_bc must be 0).  tetah and gz are module
 * `parametros`'s (SuLi-IT/0_modules.f90:150-153).  Layouts: u and dzx
 * (0:nx+2, 0:ny+1, 0:nz+1), v and dzy (0:nx+1, 0:ny+2, 0:nz+1), w and dzz
 * (0:nx+1, 0:ny+1, 0:nz+2), eta1 (0:nx+1, 0:ny+1), deltah (nx, ny), prd1
 * the halo box (in: initial guess; out: the whole box).  The reference stop
 * rule is SuLi-IT's own: |r.r| > 1e-11 tested on the r.r that entered the
 * previous iteration (:165-177); the relative rule and precisions are as
 * for suli_ps_graddin.  The operator has no null space.
 *
 * posdin, SuLi-IT/10_pos.f90:73-132 without its layer-thickness update
 * (dz(:,:,nz) and dz_it() stay with the caller): eta1 += prd1(:,:,nz)/gz over
 * the whole eta1 box, u, v -= tetah dt/dx (dy) times the pressure
 * difference, w -= tetah dt/dzz times it; prd = prd1 (prd may be NULL).
 * The reference updates eta1, then dz and dzz (dz_it), then corrects w with
 * the new dzz: a caller doing the same passes eta1 = NULL (its own eta1
 * update already done) and the updated dzz.
 *
 * Host pointers; the _dev forms take device pointers. */
int suli_ps_it_graddin(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz,
                       const double *u, const double *v, const double *w, const double *dzx,
                       const double *dzy, const double *dzz, const double *eta1,
                       const double *deltah, double *prd1, suli_ps_stats *stats);
int suli_ps_it_graddin_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz,
                           const double *u, const double *v, const double *w, const double *dzx,
                           const double *dzy, const double *dzz, const double *eta1,
                           const double *deltah, double *prd1, suli_ps_stats *stats);
int suli_ps_it_posdin(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz, double *u,
                      double *v, double *w, const double *dzz, double *eta1, const double *prd1,
                      double *prd);
int suli_ps_it_posdin_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz,
                          double *u, double *v, double *w, const double *dzz, double *eta1,
                          const double *prd1, double *prd);

/* SuLi-IT's posdin whole, as the reference runs it (SuLi-IT/10_pos.f90:73-132
 * with dz_it, SuLi-IT/11_diversos.f90:80-141): over the (0:nx+1,0:ny+1) box
 * eta1 += prd1(:,:,nz)/gz and dz(:,:,nz) += eta1 - eta0; then the face
 * thicknesses of the top planes: dzx and dzy on planes nz and nz+1 (face
 * means; end faces periodic when ccx0 / ccy0 is 0, else extrapolated from
 * the two inside), dzz on planes nz, nz+1 (means) and nz+2 (extrapolated);
 * then the correction of suli_ps_it_posdin with the new dzz; prd = prd1
 * (prd may be NULL).  dz in prd1's layout (0:nx+1,0:ny+1,0:nz+1), eta0 as
 * eta1.  Host pointers; the _dev form takes device pointers. */
int suli_ps_it_posdin_layers(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz,
                             double *u, double *v, double *w, double *dz, double *dzx, double *dzy,
                             double *dzz, const double *eta0, double *eta1, const double *prd1,
                             double *prd);
int suli_ps_it_posdin_layers_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz,
                                 double *u, double *v, double *w, double *dz, double *dzx,
                                 double *dzy, double *dzz, const double *eta0, double *eta1,
                                 const double *prd1, double *prd);

/* SuLi restart file (dados/arquivorestart): restart_salva writes the model
 * state as one list-directed record (SuLi-IC/11_diversos.f90:581-599) that
 * restart_ini reads back (:500-577).  This reads the pressure step's arrays
 * out of such a file for a grid of nx*ny*nz cells, in the layouts above
 * (host memory, Fortran order): any non-NULL pointer receives its array,
 * NULL ones are skipped.  Host code only; no context needed.  Returns 0, or
 * SULI_PS_EINVAL with a message in err (when err is non-NULL) for a file
 * that cannot be opened, holds a malformed value, or ends early.  Null
 * values ("r*") leave the destination as it was, as Fortran's read does.
 * The record holds byxf, never allocated by SuLi (0_modules.f90:153); when
 * hs (after it) is asked for, the file is counted once to size it. */
typedef struct suli_ps_restart {
    double *u, *v, *w;          /* face layouts as suli_ps_graddin */
    double *prd0, *prd1, *prd;  /* (0:nx+1, 0:ny+1, 0:nz+1) */
    double *rho, *ls_mu, *ls;   /* (nx, ny, nz) */
    double *hs;                 /* (nx, ny, nz), the Heaviside field */
} suli_ps_restart;
int suli_ps_restart_read(const char *path, int32_t nx, int32_t ny, int32_t nz,
                         suli_ps_restart *out, char *err, size_t errlen);

/* One problem of a batch: its grid, device arrays (layouts as above; solid
 * as for suli_ps_graddin_masked_dev, may be NULL) and, on return, its
 * statistics. */
typedef struct suli_ps_problem {
    suli_ps_grid grid;
    const double *u, *v, *w, *rho;
    const double *obs_lsx, *obs_lsy, *obs_lsz;
    const double *dz, *dzz;
    double *prd1;
    const uint8_t *solid;
    suli_ps_stats stats;
} suli_ps_problem;

/* graddin on `count` independent problems (device pointers, any mix of
 * sizes), solved together: each CG phase is one launch over the tiles of
 * every problem still iterating.  Each problem stops by its own rule and
 * its result is bit-identical to solving it alone with suli_ps_graddin_dev.
 * The reference solves one grid per call (9_grad.f90:10); this is the same
 * operation applied to many.  Returns the largest per-problem status (0 or
 * 1) or an error; stats.solve_ms is the time of the whole batch. */
int suli_ps_graddin_batch_dev(suli_ps_ctx *ctx, suli_ps_problem *probs, int32_t count);

/* ---- Serving: pressure solves submitted as they arrive -------------------
 *
 * A server owns a device stream, a copy stream and two threads.  Requests
 * (host arrays in the layouts above) are admitted in arrival order while
 * their device memory (fields plus solver state) fits the HBM budget; their
 * inputs are copied while other requests iterate, and every admitted
 * request joins the running batch at the next chunk of CG steps (the
 * kernels of suli_ps_graddin_batch_dev).  Each result is bit-identical to
 * solving the request alone.  The reference has no counterpart: it solves
 * the one grid of its time loop (1_principal.f90:102). */
typedef struct suli_ps_server suli_ps_server;

typedef struct suli_ps_server_opts {
    suli_ps_opts solve;     /* options of every request (precision, stop rule, ...) */
    uint64_t hbm_budget;    /* device bytes for resident requests; 0: 80% of free memory */
    int32_t max_active;     /* requests resident at once; <= 0 or > 256: 256 */
    int32_t reserved;
} suli_ps_server_opts;

typedef struct suli_ps_server_info {
    int64_t submitted, completed;
    int32_t queued, resident;      /* waiting for memory / holding device memory */
    uint64_t resident_bytes, peak_resident_bytes, budget;
    int64_t chunks;                /* chunks of CG steps launched */
} suli_ps_server_info;

suli_ps_server *suli_ps_server_create(const suli_ps_server_opts *opts);
/* Finishes every submitted request, then stops the threads. */
void suli_ps_server_destroy(suli_ps_server *srv);
const char *suli_ps_server_last_error(const suli_ps_server *srv);
/* Queues a solve of graddin's problem (host pointers, kept valid by the
 * caller until wait returns; prd1 is read as the initial guess and written
 * with the result).  Returns a ticket >= 0, or < 0 (a request larger than
 * the budget is SULI_PS_ENOMEM).  Thread-safe. */
int64_t suli_ps_server_submit(suli_ps_server *srv, const suli_ps_grid *g, const double *u,
                              const double *v, const double *w, const double *rho,
                              const double *obs_lsx, const double *obs_lsy,
                              const double *obs_lsz, const double *dz, const double *dzz,
                              double *prd1);
/* Waits up to timeout_s seconds (< 0: no limit) for a request; on return 0
 * or 1 its prd1 holds the result and stats (if non-NULL) its statistics,
 * with solve_ms the time from submit to result on the host.  Returns
 * SULI_PS_EAGAIN on timeout.  Each ticket is waited for once. */
int suli_ps_server_wait(suli_ps_server *srv, int64_t ticket, double timeout_s,
                        suli_ps_stats *stats);
int suli_ps_server_get_info(suli_ps_server *srv, suli_ps_server_info *info);

/* Device-memory helpers for callers without their own HIP code. */
void *suli_ps_dev_alloc(suli_ps_ctx *ctx, size_t bytes);
int suli_ps_dev_free(suli_ps_ctx *ctx, void *ptr);
/* Host -> device and device -> host copies on the context's stream, through
 * the context's pinned staging above a few MiB; each returns once the copy is
 * complete.  NULL pointers with a nonzero size are rejected. */
int suli_ps_memcpy_h2d(suli_ps_ctx *ctx, void *dst, const void *src, size_t bytes);
int suli_ps_memcpy_d2h(suli_ps_ctx *ctx, void *dst, const void *src, size_t bytes);
int suli_ps_dev_zero(suli_ps_ctx *ctx, void *ptr, size_t bytes); /* async, ctx stream */
int suli_ps_synchronize(suli_ps_ctx *ctx);

/* Measurement hook: runs the solver's SpMV kernel q = A p `reps` times on
 * the operator of the last graddin call (precision as given) for each of a
 * few tile shapes of its own (depths of 2..24 planes; 1, 2 or 4 cells per
 * lane in FP32) and returns the fastest shape's mean launch duration in ms,
 * timed with HIP events on the context stream. */
int suli_ps_time_spmv(suli_ps_ctx *ctx, int32_t precision, int32_t reps, double *ms);

/* Tile shape of the CG kernels for a grid size, no reference counterpart:
 * times the per-step kernel (FP64 cg_a, or the mixed solver's FP32 step)
 * on a scratch problem of this size over tile depths of 2..48 planes and,
 * for the FP32 step, one, two or four x cells per lane (two and four where
 * nx is a multiple), and keeps the
 * fastest for this context (kz, vx, and its time per step).  Solves of
 * that size then use it when opts.autotune is set.  A result depends on
 * the tiling (the order of its sums), so solves of one size agree bit for
 * bit only under the same choice.  suli_ps_set_tiles fixes a shape
 * instead (vx 1, 2 or 4; 1 for FP64 and where nx is not a multiple).
 *
 * With two or four cells per lane the FP32 step comes in two kinds: its
 * plane walk issued plane by plane, or pipelined (each plane's loads issued
 * while the plane before it is computed: more memory requests in flight per
 * wave, fewer waves per SIMD).  vx | SULI_PS_TILES_PIPELINED names the
 * pipelined kind, here and in what suli_ps_autotune returns. */
#define SULI_PS_TILES_PIPELINED 8
int suli_ps_autotune(suli_ps_ctx *ctx, const suli_ps_grid *g, int32_t precision, int32_t *kz,
                     int32_t *vx, double *step_us);
int suli_ps_set_tiles(suli_ps_ctx *ctx, const suli_ps_grid *g, int32_t precision, int32_t kz,
                      int32_t vx);

/* Measurement hook with no reference counterpart: a plain device copy of
 * `bytes` (16-byte vectors, grid-stride) run `reps` times between two
 * device buffers the library allocates; returns the mean launch time in ms.
 * Its known traffic (bytes read, bytes written) calibrates HBM counters. */
int suli_ps_copy_bench(suli_ps_ctx *ctx, size_t bytes, int32_t reps, double *ms);

/* Measurement hook with no reference counterpart: a plain device read of
 * `bytes` (16-byte vectors, grid-stride, summed per thread) run `reps`
 * times; returns the mean launch time in ms.  With the copy it brackets the
 * HBM bandwidth a kernel mixing reads and writes can reach. */
int suli_ps_read_bench(suli_ps_ctx *ctx, size_t bytes, int32_t reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
