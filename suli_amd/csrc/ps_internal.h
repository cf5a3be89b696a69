// Internal declarations shared by the kernels (ps_setup.hip, ps_cg.hip)
// and the host solver (ps_solver.hip).  Not part of the public ABI
// (include/suli_ps.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sps {

constexpr int kThreads = 256;        // every kernel: 4 wave64s
constexpr int kTileX = 64;           // SpMV tile: one wave per x row
constexpr int kTileY = 4;            // four rows per workgroup
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, round-robin dispatch
constexpr int kVecBlocks = 512;      // workgroups of the 1-D vector kernels
constexpr int kSpmvBlocks = 1024;    // target workgroups of the SpMV
constexpr int kNotDone = 0x7fffffff;

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

// Device-resident CG state.
//
// No kernel communicates between its own workgroups: every reduction
// leaves one partial sum per workgroup, and the kernel that needs the total
// sums those partials itself (each workgroup in the same fixed order, so all
// get the same bits) after the kernel boundary has made them visible on
// every XCD.  Scalars derived from a total are stored by workgroup 0 for
// later kernels; values a kernel reads while its own workgroup 0 may be
// writing are double-buffered by iteration parity (rr) or are benign
// (done_at only ever moves from "running" to the current iteration).
struct CgScalars {
    double rr[2];     // r.r entering iteration n in slot n & 1 (r_{n-1}.r_{n-1})
    double alpha;     // alpha of iteration n (cg_b -> next cg_a, refresh)
    double pap;
    double bb;        // b.b (scaled right-hand side)
    double thresh;    // stop threshold on r.r
    double rr0;       // initial r.r
    double rr_true;   // FP64 r.r of the last recomputed true residual
    double gamma;     // null-space component removed from the initial residual
    double en0;       // e0 . sqrt(d): the initial gauge (null-space coordinate)
    double dd;        // sum of d = sqrt(d) . sqrt(d)
    double gshift;    // constant removed from prd1 to restore the initial gauge
    int done_at;      // iteration whose cg_a met the stop rule; kNotDone while running
    int final_;       // 1: the solve is over (mixed: decided at a replacement)
    int iters;        // CG iterations completed (reference `cont`)
    int max_iter;
    int refreshes;    // FP64 residual replacements (mixed)
    int status;       // 0 converged, 1 capped
    int mixed;        // 1: convergence is judged on rr_true at replacements
    int project;      // 1: remove the null-space component of the residual
    int rel_rule;     // 1: thresh = tol^2 b.b
    double tol;
    double thresh_ref;
};

// Offsets (in doubles) of the per-workgroup partial sums in one buffer.
struct Partials {
    double *spmv;     // [B1]      p.Ap of cg_a
    double *upd;      // [B2]      r.r of cg_b (or one total, rest zero, after a replacement)
    double *tres;     // [3*B1]    true residual r.r (FP64, FP32 copy) and e.sqrt(d)
    double *init;     // [5*B0]    setup: r0.r0, b.b, r0.sqrt(d), d, e0.sqrt(d)
    double *proj;     // [2*B0]    setup after projection: r0.r0, FP32 r0.r0
};

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
// y-face nx*(ny+1)*nz, z-face nx*ny*(nz+1)).
template <typename T>
struct Op {
    const T *sx, *sy, *sz;
};

// ---- setup (ps_setup.hip) ----
void launch_face_coefs(const Geo &g, const SetupIn &in, double *sx, double *sy, double *sz, hipStream_t s);
void launch_diag_rhs(const Geo &g, const SetupIn &in, const double *sx, const double *sy, const double *sz,
                     double *d, double *b, hipStream_t s);
void launch_scale_faces(const Geo &g, const double *d, double *sx, double *sy, double *sz, float *sx32,
                        float *sy32, float *sz32, hipStream_t s);
// Initial residual, null-space projection; leaves r (r64, or r32 and c32 = 0
// for mixed precision) and S ready for iteration 1.
void launch_init(const Geo &g, const SetupIn &in, Op<double> op, const double *d, double *b, double *e,
                 double *r0, double *r64, float *r32, float *c32, const Partials &P, CgScalars *S, hipStream_t s);
// prd1 = e / sqrt(d) - S->gshift on the whole halo box.
void launch_unscale(const Geo &g, const double *e, const double *d, double *prd1, const CgScalars *S,
                    hipStream_t s);
void launch_posdin(const Geo &g, const SetupIn &in, double *u, double *v, double *w, hipStream_t s);

// ---- CG iteration n (ps_cg.hip), n = 1, 2, ... ----
// cg_a(n): deferred e -= alpha_{n-1} p_{n-1}; stop rule on r.r; p_n = r +
// beta p_{n-1} (p_old -> p_new, two buffers); q = A p_n; p.q partials.
// ran (nullable): set to 1 by the launch if it did the SpMV (timing samples).
// do_e = false right after a residual replacement (the deferred update is
// already in e).
template <typename T>
void launch_cg_a(const Geo &g, int n, bool do_e, T *e, const T *p_old, T *p_new, const T *r, Op<T> op, T *q,
                 const Partials &P, CgScalars *S, int *ran, hipStream_t s);
// cg_b(n): alpha_n = r.r / p.q; r -= alpha_n q; r.r partials.
template <typename T>
void launch_cg_b(const Geo &g, int n, T *r, const T *q, const Partials &P, CgScalars *S, hipStream_t s);
// Mixed precision residual replacement after iteration n_last: e += c (with
// the deferred alpha p term), r = b - A e in FP64 handed to the FP32
// recurrence; decide converged / capped / continue.
void launch_refresh(const Geo &g, int n_last, double *e, float *c32, const float *p_last, Op<double> op,
                    const double *b, float *r32, const Partials &P, CgScalars *S, hipStream_t s);
// After the solve: FP64 true residual r.r into S->rr_true and, when the
// null space is projected, the gauge shift that restores e0's coordinate.
void launch_finish(const Geo &g, const double *e, const double *d, Op<double> op, const double *b,
                   const Partials &P, CgScalars *S, hipStream_t s);
// Stand-alone SpMV q = A p (no vector updates), for timing.
template <typename T>
void launch_spmv_only(const Geo &g, const T *p, Op<T> op, T *q, const Partials &P, hipStream_t s);

int spmv_blocks(const Geo &g);
Geo make_geo(int nx, int ny, int nz, int px, int py);

}  // namespace sps
