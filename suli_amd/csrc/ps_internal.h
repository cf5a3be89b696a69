// Internal declarations shared by the kernels (ps_setup.hip, ps_cg.hip)
// and the host side (ps_engine.hip, ps_solver.hip).  Not part of the public
// ABI (include/suli_ps.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/suli_ps.h"

namespace sps {

constexpr int kThreads = 256;        // every kernel: 4 wave64s
constexpr int kTileX = 64;           // 3-D tile: one wave per x row
constexpr int kTileY = 4;            // four rows per workgroup
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, round-robin dispatch
constexpr int kVecBlocks = 512;      // 1-D tiles of a large problem
constexpr int kInitBlocks = 4096;    // workgroups (and partial sums) of the setup reductions
constexpr int kSpmvBlocks = 1024;    // target 3-D tiles of a large problem
constexpr int kNotDone = 0x7fffffff;
// Mixed precision: residual replacements happen at iterations that are
// multiples of kRefreshGrain, the first at kFirstRefresh; each problem
// schedules its next one from its own convergence history, so its results
// never depend on the other problems of a batch.
constexpr int kRefreshGrain = 8;
constexpr int kFirstRefresh = 32;
constexpr int kMaxRefreshGap = 128;
// Mixed precision: the FP32 correction c takes the updates alpha_j p_j of
// kFold iterations at once (at iterations n = 1 mod kFold, from a ring of
// kFold direction vectors), so that a replacement, which falls after an
// iteration that is a multiple of kFold, never splits a group.  A plain
// step moves 8 words per cell; a folding one also reads c and the kFold - 2
// oldest directions and writes c, so the step moves 9 words per cell on
// average.  Measured per mixed solve at 256x128x96 (when the steps still
// stored r): 26.29 ms folding pairs, 26.05 ms folding fours; folding eights
// ran the step kernel out of registers (the fold's loads were issued one at
// a time) and took 33.5 ms.
constexpr int kFold = 4;
static_assert(kRefreshGrain % kFold == 0 && kFirstRefresh % kFold == 0, "replacements must end fold groups");
static_assert((kFold & (kFold - 1)) == 0, "kFold is a power of two");

// Grid geometry as the kernels see it.  Cells are 0-based (i, j, k) with x
// fastest; a cell vector holds nx*ny*nz values with no halo.  Ghost values
// follow the reference's mirror (Neumann) or periodic rule per direction,
// applied through index clamping/wrapping instead of stored halos.
struct Geo {
    int nx, ny, nz;
    int px, py;               // periodic in x / y (z is always mirrored)
    long long n;              // nx*ny*nz
    int tiles_x, tiles_y, tiles_z, kz;   // 3-D tiling: (64 vx) x 4 columns, kz planes
    int vx;                   // x cells per lane in 3-D tiles: 1, or 2 or 4 (FP32, nx a multiple)
    int pipe;                 // the mixed step's plane walk is pipelined (vx 2 or 4)
    int dtop;                 // 1: p = 0 on the top face (top ghost = -interior)
};

// Lane code of a tile shape: the cells per lane, with kPipelined set for the
// mixed step's pipelined walk (SULI_PS_TILES_PIPELINED of the C ABI).
constexpr int kPipelined = SULI_PS_TILES_PIPELINED;
inline int lane_code(const Geo &g) { return g.vx | (g.pipe ? kPipelined : 0); }
// The mixed step's kernel kinds, one launch per kind in a step: one cell per
// lane; two; two, pipelined; four; four, pipelined.
constexpr int kLaneKinds = 5;
inline int lane_kind(const Geo &g) { return g.vx == 1 ? 0 : (g.vx == 2 ? 1 : 3) + g.pipe; }
constexpr int kind_vx(int kind) { return kind == 0 ? 1 : kind < 3 ? 2 : 4; }
constexpr bool kind_pipe(int kind) { return kind == 2 || kind == 4; }

// Device-resident CG state of one problem.
//
// No kernel communicates between its own workgroups: every reduction
// leaves one partial sum per tile, and the kernel that needs the total sums
// those partials itself (each workgroup in the same fixed order, so all get
// the same bits) after the kernel boundary has made them visible on every
// XCD.  Scalars derived from a total are stored by the workgroup that
// handles tile 0 for later kernels; values a kernel reads while that
// workgroup may be writing are double-buffered by iteration parity (rr) or
// are benign (done_at only ever moves from "running" to the current
// iteration).
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
    int skip_e_at;    // global step whose cg_a skips the deferred e update
    int status;       // 0 converged, 1 capped
    int mixed;        // 1: convergence is judged on rr_true at replacements
    int project;      // 1: remove the null-space component of the residual
    int rel_rule;     // 1: thresh = tol^2 b.b
    int next_refresh; // iteration of the next residual replacement (mixed)
    int prev_iters;   // iteration of the previous replacement (0: none yet)
    double tol;
    double thresh_ref;
    double prev_rr;   // true r.r at the previous replacement
    double slope;     // d log(r.r) / d iteration between replacements (< 0)
    double rr_after;  // r.r of the replaced residual (mixed), used by the next step
    double beta_after;  // beta of the step after a replacement (mixed)
    double alph[kFold];   // mixed: alpha_k in slot k % kFold, for the correction c
    double bet[2];    // mixed: beta_n in slot n & 1, the factor that formed p_n
    int lag_stop;     // 1: the FP64 stop test reads the r.r that entered the previous
                      // iteration (SuLi-IT's loop, SuLi-IT/9_grad.f90:165-177)
};

// Per-tile partial sums of one problem.
struct Partials {
    double *spmv;     // [nt3]      p.Ap of cg_a
    double *upd;      // [nt1]      r.r of cg_b (or one total, rest zero, after a replacement)
    double *tres;     // [3*nt3]    true residual r.r (FP64, FP32 copy) and e.sqrt(d)
    double *init;     // [5*ni]     setup: r0.r0, b.b, r0.sqrt(d), d, e0.sqrt(d) (ni = init_blocks)
    double *proj;     // [2*ni]     setup after projection: r0.r0, FP32 r0.r0
    double *m4;       // [2][nt3][4] one-launch FP32 iteration n, slot n & 1:
                      //            p_n.q_n, q_n.r_{n-1}, q_n.q_n, r_{n-1}.r_{n-1}
};

// Device descriptor of one problem of a batch: geometry, tiling, where its
// arrays live, and the global step at which its iteration 1 runs.  Written
// once when the problem is admitted.
struct ProbDev {
    Geo g;
    int n0;           // global step of iteration 1 (iteration n runs at step n0 + n - 1)
    int nt3;          // 3-D tiles (cg_a, true residual)
    int nt1;          // 1-D tiles (cg_b, refresh)
    int mixed;
    // operator (Jacobi-scaled face coefficients), diagonal, effective rhs
    const double *sx, *sy, *sz, *d;
    double *b;
    // FP64 iterate and, for FP64 solves, the CG vectors (p alternates)
    double *e, *r, *q, *p0, *p1;
    // FP32 inner solve (mixed): operator copy, correction c, vectors.  The
    // one-launch iteration alternates q between two buffers (q32/q32b by
    // iteration parity) and keeps p_k in ring slot k % kFold (at p32 +
    // (k % kFold) * pslot) until c has taken alpha_k p_k.  r32 holds the
    // residual the setup and each replacement leave (r_0, r_n at a
    // replacement after an even step n); the steps recover r from the
    // directions and store none, or with store_r store r_k in r32 (k even)
    // or r32b (k odd).
    const float *sx32, *sy32, *sz32;
    float *c32, *r32, *q32, *q32b, *r32b;
    int store_r;
    float *p32;
    long long pslot;
    // constant density (t_press = 2, no solid mask): every face of a plane
    // carries the same scaled coefficient, so the iteration kernels read
    // tx[k], ty[k] (x and y faces of plane k) and tz[f] (z face f) instead
    // of the face arrays; nullptr otherwise
    const double *tx, *ty, *tz;
    const float *tx32, *ty32, *tz32;
    // any other mixed problem: per-plane values of the FP32 face arrays and
    // one bit per row (j, k) whose faces all hold them (word k * ((ny + 31)
    // / 32) + j / 32, bit j % 32); such rows read their coefficients from
    // the tables, which repeat each value (8 times per plane in rx32, 4 in
    // ry32 and per z face in rz32) so that the same vector loads serve
    // both.  The FP32 kernels address the tables whenever they read the face
    // arrays, so these point to valid memory then (rowreg may be null: no
    // row is read from the tables)
    const float *rx32, *ry32, *rz32;
    const unsigned *rowreg;
    Partials P;
    CgScalars *S;
};

// Global step of a launch: *base + off when base is set (kernels replayed
// from a captured HIP graph read the step from device memory), else off.
// next (nullable): the launch stores *base + kRefreshGrain there (from one
// thread) for the next replayed group, which reads that counter instead:
// two counters alternate, so no launch ever writes the one it reads.
struct StepRef {
    const int *base;
    int off;
    int *next;
};

// Entry of a launch table: tile `tile` of problem `prob`.
struct TileRef {
    int prob;
    int tile;
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
    const uint8_t *solid;               // cells (x fastest, no halo) masked out; nullptr: none
    // SuLi-IT's non-hydrostatic system (SuLi-IT/9_grad.f90:44-81) when it
    // is set: coefficients from the layer thicknesses dzx (u layout), dzy
    // (v layout), dz3 (w layout), a free-surface row on the top plane from
    // eta1 (0:nx+1,0:ny+1) and deltah (nx,ny)
    int it;
    const double *dzx, *dzy, *dz3, *eta1, *deltah;
    double tetah, gz;
};

// Operator storage (scaled face coefficients, x-face array (nx+1)*ny*nz,
// y-face nx*(ny+1)*nz, z-face nx*ny*(nz+1)).
template <typename T>
struct Op {
    const T *sx, *sy, *sz;
};

// ---- setup, per problem (ps_setup.hip) ----
void launch_face_coefs(const Geo &g, const SetupIn &in, double *sx, double *sy, double *sz, hipStream_t s);
// Also writes the scaled unknown e = prd1 sqrt(d) of every cell.
void launch_diag_rhs(const Geo &g, const SetupIn &in, const double *sx, const double *sy, const double *sz,
                     double *d, double *b, double *e, hipStream_t s);
void launch_scale_faces(const Geo &g, const double *d, double *sx, double *sy, double *sz, float *sx32,
                        float *sy32, float *sz32, hipStream_t s);
// Per-plane coefficient tables of a constant-density operator, copied from
// its (plane-constant) face arrays; the FP32 ones when sx32 is non-null.
// The FP32 row tables and the bit per row (j, k) whose faces all hold them
// (rowreg: (ny + 31) / 32 words per plane).
void launch_row_tables(const Geo &g, const float *sx32, const float *sy32, const float *sz32, float *rx, float *ry,
                       float *rz, unsigned *rowreg, hipStream_t s);
void launch_face_tables(const Geo &g, const double *sx, const double *sy, const double *sz, const float *sx32,
                        const float *sy32, const float *sz32, double *tx, double *ty, double *tz, float *tx32,
                        float *ty32, float *tz32, hipStream_t s);
// Workgroups of launch_init's reductions (its nb0): one per cell tile of
// 64 x 4 cells, at most kInitBlocks; the setup partials hold 5 and 2 per
// workgroup.
int init_blocks(const Geo &g);
// Initial residual and null-space projection on nb0 workgroups; leaves r
// (r64, or r32 with c32 = 0 for mixed precision) and S ready for iteration 1.
void launch_init(const Geo &g, const SetupIn &in, Op<double> op, const double *d, double *b, double *e,
                 double *r0, double *r64, float *r32, float *c32, const Partials &P, int nb0, CgScalars *S,
                 hipStream_t s);
// prd1 = e / sqrt(d) - S->gshift on the whole halo box (0 in solid cells).
void launch_unscale(const Geo &g, const double *e, const double *d, double *prd1, const CgScalars *S,
                    const uint8_t *solid, hipStream_t s);
void launch_posdin(const Geo &g, const SetupIn &in, double *u, double *v, double *w, hipStream_t s);
// Hydrostatic mode (t_press = 0): pressh's column integral into prd1's
// interior, and posdin's other branch (u, v correction, w from continuity,
// velocities zeroed where hs < 1e-4 on the face).
void launch_pressh(const Geo &g, const double *rho, const double *dz, double dzu, double gz, double *prd1,
                   hipStream_t s);
// SuLi-IT posdin without dz_it (SuLi-IT/10_pos.f90:73-132).
void launch_it_posdin(const Geo &g, double dx, double dy, double dt, double tetah, double gz, const double *prd1,
                      const double *dz3, double *eta1, double *u, double *v, double *w, hipStream_t s);
// SuLi-IT posdin's surface layer before that: eta1 and dz(:,:,nz) from prd1
// and eta0, then dz_it (dzx, dzy, dzz of the top planes).
void launch_it_layers(const Geo &g, double gz, const double *prd1, const double *eta0, double *eta1, double *dz,
                      double *dzx, double *dzy, double *dzz, hipStream_t s);
// rho (and hs, mu when non-null) from the level set, heaviside with t_tens = 0.
void launch_heaviside(long long n, const double *ls, double eps, double rho1, double rho2, double mu1, double mu2,
                      double *hs, double *rho, double *mu, hipStream_t s);
// The same with t_tens = 1: also hsx, hsy, hsz from the level set's WENO5
// derivatives (x, y periodic per g.px, g.py; uniform spacing dx, dy, dz).
void launch_heaviside_tens(const Geo &g, const double *ls, double eps, double rho1, double rho2, double mu1,
                           double mu2, double dx, double dy, double dz, double *hs, double *rho, double *mu,
                           double *hsx, double *hsy, double *hsz, hipStream_t s);
// exp_press (IPA): px, py, pz from prd1 + (prd1 + prd0) and the face
// densities, then prd0 += prd1 over the halo box.
void launch_exp_press(const Geo &g, const double *rho, const double *dz, const double *dzz, double dx, double dy,
                      double dzu, const double *prd1, double *prd0, double *px, double *py, double *pz,
                      hipStream_t s);
void launch_posdin_hydro(const Geo &g, const SetupIn &in, const double *hs, double *u, double *v, double *w,
                         hipStream_t s);

// ---- batched CG (ps_cg.hip) ----
// Global step `step` runs iteration step - n0 + 1 of every problem in the
// table.  Kernels are persistent: `grid` workgroups walk the tile table.
// cg_a: deferred e -= alpha p; stop rule on r.r; p_n = r + beta p_{n-1}
// (p buffers alternate); q = A p_n; p.q partials.  ran (nullable): set to 1
// if problem 0's tile 0 did the SpMV (timing samples).
void launch_cg_a(bool fp32, const ProbDev *P, const TileRef *t3, int nt3, int grid, StepRef step, int *ran,
                 hipStream_t s);
// cg_b: alpha = r.r / p.q; r -= alpha q; r.r partials.
void launch_cg_b(bool fp32, const ProbDev *P, const TileRef *t1, int nt1, int grid, StepRef step, hipStream_t s);
// The FP32 inner iteration of the mixed solver in one launch per step:
// finishes iteration n-1 (r and e updates, formed where needed) and does
// iteration n's direction update and SpMV, with r.r of the finished
// iteration from the exact identity r_{n-1}.r_{n-1} = r_{n-2}.r_{n-2}
// - 2 alpha q.r_{n-2} + alpha^2 q.q anchored on the directly summed
// r_{n-2}.r_{n-2}.  ran: as for cg_a.
// vx: the cells per lane of every problem in t3
void launch_cg_m(const ProbDev *P, const TileRef *t3, int nt3, int grid, StepRef step, int *ran, int vx, bool pipe,
                 hipStream_t s);
// Mixed precision residual replacement after global step `step` for the
// problems of the tables (`probs` lists them, one entry per problem): e += c
// (with the deferred alpha p term), r = b - A e in FP64 handed to the FP32
// recurrence; decide converged / capped / continue.
void launch_refresh(const ProbDev *P, const TileRef *t1, int nt1, const TileRef *t3, int nt3, const int *probs,
                    int nprobs, int grid, StepRef step, hipStream_t s);
// *base = value: sets a device step counter.
void launch_set_step(int *base, int value, hipStream_t s);
// After a problem's solve: FP64 true residual r.r into S->rr_true and the
// gauge shift that restores e0's null-space coordinate.
void launch_finish(const ProbDev *P, const TileRef *t3, int nt3, const int *probs, int nprobs, int grid,
                   hipStream_t s);
// Stand-alone SpMV q = A p of one problem (no vector updates), for timing.
void launch_spmv_only(bool fp32, const ProbDev *P, const TileRef *t3, int nt3, int grid, hipStream_t s);

// Plain copy of n16 16-byte vectors (measurement only).
void launch_copy(const void *src, void *dst, long long n16, hipStream_t s);
// Plain read of n16 16-byte vectors (measurement only); out holds
// kReadGrid * kThreads floats.
constexpr int kReadGrid = 2048;
void launch_read(const void *src, float *out, long long n16, hipStream_t s);

int spmv_tiles(const Geo &g);
int vec_tiles(long long n);
// vx: a lane code (cells per lane, kPipelined for the pipelined walk); a
// width that does not divide nx falls back to one cell per lane
Geo make_geo(int nx, int ny, int nz, int px, int py, int kz = 0, int vx = 1);

}  // namespace sps
