// Internal declarations shared by the kernels (ps_kernels.hip) and the host
// solver (ps_solver.hip).  Not part of the public ABI (include/suli_ps.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sps {

constexpr int kThreads = 256;        // every kernel: 4 wave64s
constexpr int kTileX = 64;           // SpMV tile: one wave per x row
constexpr int kTileY = 4;            // four rows per workgroup
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, round-robin dispatch

// Grid geometry as the kernels see it.  Cells are 0-based (i, j, k) with x
// fastest; a cell vector holds nx*ny*nz values with no halo.  Ghost values
// follow the reference's mirror (Neumann) or periodic rule per direction,
// applied through index clamping/wrapping instead of stored halos.
struct Geo {
    int nx, ny, nz;
    int px, py;               // periodic in x / y (z is always mirrored)
    long long n;              // nx*ny*nz
    int tiles_x, tiles_y, tiles_z, kz;   // SpMV tiling: kz planes per workgroup
};

// Device-resident CG state.  Written by exactly one block at a time (the
// last block of a reduction, found with a ticket), read by all blocks at
// kernel entry.
struct CgScalars {
    double rr;        // r.r the recurrence uses (FP32 residual in mixed mode)
    double pap;       // p.Ap
    double alpha, beta;
    double bb;        // b.b (scaled right-hand side)
    double thresh;    // stop threshold on r.r
    double rr0;       // initial r.r
    double rr_true;   // FP64 r.r of the last recomputed true residual
    int stop;         // 1: the CG kernels skip (converged, capped, or refresh due)
    int final_;       // 1: solve finished; every kernel skips
    int k2n, k3n;     // completed update_r / update_ep launches
    int iters;        // CG iterations (reference `cont`)
    int max_iter;
    int refreshes;    // FP64 residual replacements (mixed)
    int refresh_mark; // k3n at the last replacement
    int status;       // 0 converged, 1 capped
    int mixed;        // 1: convergence is judged on rr_true at replacements
    int pad[2];
    unsigned ticket[8];
};

enum TicketSlot { kTkSpmv = 0, kTkUpdR = 1, kTkUpdEP = 2, kTkTrue = 3, kTkInit = 4, kTkRefresh = 5 };

// Setup inputs in the reference's Fortran layouts (device pointers).
struct SetupIn {
    const double *u, *v, *w, *rho, *prd1;
    const double *obsx, *obsy, *obsz;   // nullptr unless the IBM mask is on
    const double *dz, *dzz;             // nullptr on uniform grids
    double dx, dy, dzu, dt;             // dzu: uniform dz
    double aux1, aux2, aux3;            // dt/dx^2, dt/dy^2, dt/dz^2
    double rho_f1;
    int ipa;                            // faces use rho_f1 (t_press == 2)
};

// Operator storage (scaled face coefficients, x-face array (nx+1)*ny*nz,
// y-face nx*(ny+1)*nz, z-face nx*ny*(nz+1)) plus the diagonal d.
template <typename T>
struct Op {
    const T *sx, *sy, *sz;
};

// ---- launchers (ps_kernels.hip) ----
void launch_face_coefs(const Geo &g, const SetupIn &in, double *sx, double *sy, double *sz, hipStream_t s);
void launch_diag_rhs(const Geo &g, const SetupIn &in, const double *sx, const double *sy, const double *sz,
                     double *d, double *b, hipStream_t s);
void launch_scale_faces(const Geo &g, const double *d, double *sx, double *sy, double *sz, float *sx32,
                        float *sy32, float *sz32, hipStream_t s);
void launch_init_residual(const Geo &g, const SetupIn &in, const double *sx, const double *sy, const double *sz,
                          const double *d, double *b, double *e, double *r64, double *p64, float *r32, float *p32,
                          float *c32, double *partial, double *partial2, int nblk, CgScalars *S, int rel_rule,
                          double tol, double thresh_ref, hipStream_t s);
template <typename T>
void launch_spmv_pap(const Geo &g, const T *p, Op<T> op, T *q, double *partial, CgScalars *S, hipStream_t s);
template <typename T>
void launch_update_r(const Geo &g, T *r, const T *q, double *partial, int nblk, CgScalars *S, hipStream_t s);
template <typename T, typename TE>
void launch_update_ep(const Geo &g, TE *e, T *p, const T *r, int nblk, CgScalars *S, hipStream_t s);
void launch_refresh_e(const Geo &g, double *e, float *c32, int nblk, CgScalars *S, hipStream_t s);
void launch_true_residual(const Geo &g, const double *e, Op<double> op, const double *b, float *r32,
                          double *partial, double *partial2, int final_check, CgScalars *S, hipStream_t s);
void launch_unscale(const Geo &g, const double *e, const double *d, double *prd1, hipStream_t s);
void launch_posdin(const Geo &g, const SetupIn &in, double *u, double *v, double *w, hipStream_t s);

int spmv_blocks(const Geo &g);
Geo make_geo(int nx, int ny, int nz, int px, int py);

}  // namespace sps
