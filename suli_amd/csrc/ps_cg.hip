// CG iteration kernels (the hot path).
//
// SuLi's CG iteration (SuLi-IC/9_grad.f90:398-587) is restated for a
// bandwidth-bound device as two launches per iteration n:
//   cg_a(n)  finishes iteration n-1 (e -= alpha p, deferred), tests the
//            stop rule on r.r, forms p_n = r + beta p_{n-1} (on the fly at
//            every stencil point, into the other p buffer), q = A p_n, and
//            one p.q partial per workgroup;
//   cg_b(n)  alpha = r.r / p.q; r -= alpha q; one r.r partial per workgroup.
// That moves 9 + 3 = 12 words per cell instead of the reference's 17
// (cg_a reads e, p_old, r and the three face-coefficient arrays and writes
// e, p_new and q; cg_b reads r and q and writes r).  The reference's ghost
// copies of p (9_grad.f90:531-579) become clamped or wrapped neighbour
// indices.  Totals of partial sums are formed by the consuming kernel (see
// CgScalars), so no kernel needs a device-wide fence.
//
// T is double for the FP64 solver and float for the inner solve of the
// mixed-precision one (e is then the FP32 correction c); dot products and
// the CG scalars are always FP64.
#include <algorithm>

#include "ps_device.h"

namespace sps {
namespace {

template <typename T>
struct alignas(16) Vec {
    static constexpr int W = 16 / sizeof(T);
    T v[W];
};

// Tile of the 3-D kernels: 64 x 4 columns, kz planes marched in z.
struct Tile {
    int i, j, k0, k1;
};

__device__ __forceinline__ Tile tile_of(const Geo &g)
{
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % g.tiles_x;
    const int rest = tile / g.tiles_x;
    const int ty = rest % g.tiles_y;
    const int tz = rest / g.tiles_y;
    Tile t;
    t.i = tx * kTileX + (threadIdx.x & (kTileX - 1));
    t.j = ty * kTileY + (threadIdx.x / kTileX);
    t.k0 = tz * g.kz;
    t.k1 = min(t.k0 + g.kz, g.nz);
    return t;
}

// cg_a(n).  Each lane marches its (i, j) column over kz planes keeping p_n
// at k-1, k, k+1 (and p_{n-1} at k) in registers; x/y neighbours of p_n are
// formed from r and p_{n-1} loaded from the same or adjacent rows (L1/L2
// hits).  Tiles are ordered x, y, z and given to XCDs in contiguous ranges.
//   (A p)_c = p_c - sx_E p_E - sx_W p_W - sy_N p_N - sy_S p_S - sz_T p_T - sz_B p_B
// (9_grad.f90:408-420; unit diagonal after the symmetric Jacobi scaling).
// FIRST: iteration 1 (p_1 = r_0, nothing to finish).  DO_E: apply the
// deferred e -= alpha p of iteration n-1 (false right after a residual
// replacement, which folded it into e already).  Both are template
// parameters so that the plane loop has no branches: every load of a plane
// is issued before any of them is waited for.
template <typename T, bool FIRST, bool DO_E>
__global__ __launch_bounds__(kThreads) void k_cg_a(Geo g, int n, T *__restrict__ e, const T *__restrict__ p_old,
                                                   T *__restrict__ p_new, const T *__restrict__ r,
                                                   const T *__restrict__ sx, const T *__restrict__ sy,
                                                   const T *__restrict__ sz, T *__restrict__ q,
                                                   const double *__restrict__ rr_part, int n_rr,
                                                   double *__restrict__ pap_part, CgScalars *S, int *ran)
{
    if (n > S->done_at) return;
    constexpr bool first = FIRST;
    constexpr bool do_e = DO_E && !FIRST;
    // r.r entering this iteration: r_0.r_0 from setup, else cg_b(n-1)'s partials
    const double rr_in = first ? S->rr[1] : total_of(rr_part, n_rr);
    const double beta = first ? 0.0 : rr_in / S->rr[(n - 1) & 1];   // 9_grad.f90:510-515
    const T a_prev = (T)S->alpha;
    // stop rule (9_grad.f90:398,582): FP64 tests the recurrence's r.r; the
    // mixed solver decides at its FP64 residual replacements
    const bool conv = !(fabs(rr_in) > S->thresh);
    const bool cap = n - 1 >= S->max_iter;
    const bool stop = cap || (conv && !S->mixed);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->rr[n & 1] = rr_in;
        S->iters = n - 1;
        if (stop) {
            S->done_at = n;
            if (!S->mixed) {
                S->final_ = 1;
                S->status = conv ? 0 : 1;
            }
        }
        if (ran && !stop) *ran = 1;
    }
    const Tile t = tile_of(g);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long plane = (long long)nx * ny;
    double acc = 0.0;
    if (t.i < nx && t.j < ny) {
        long long c = (long long)t.k0 * plane + (long long)t.j * nx + t.i;
        if (stop) {
            // finish iteration n-1 only (e -= alpha p, 9_grad.f90:458-468)
            if (do_e)
                for (int k = t.k0; k < t.k1; ++k, c += plane) e[c] = e[c] - a_prev * p_old[c];
        } else {
            const T b = (T)beta;
            const int iw = nb_lo(t.i, nx, g.px), ie = nb_hi(t.i, nx, g.px);
            const int js = nb_lo(t.j, ny, g.py), jn = nb_hi(t.j, ny, g.py);
            const long long dS = (long long)(js - t.j) * nx, dN = (long long)(jn - t.j) * nx;
            long long fx = ((long long)t.k0 * ny + t.j) * (nx + 1) + t.i;
            long long fy = ((long long)t.k0 * (ny + 1) + t.j) * nx + t.i;
            // p_n = r + beta p_{n-1} (9_grad.f90:518-527), formed where needed
            auto pn = [&](long long x) -> T { return first ? r[x] : r[x] + b * p_old[x]; };
            T poc = first ? T(0) : p_old[c];
            T pc = first ? r[c] : r[c] + b * poc;
            T pb = t.k0 > 0 ? pn(c - plane) : pc;
            T szb = sz[c];
            for (int k = t.k0; k < t.k1; ++k) {
                const long long r0 = c - t.i;
                // top neighbour; at the lid the mirror ghost is the cell itself
                const long long up = (k < nz - 1) ? plane : 0;
                const T pot = first ? T(0) : p_old[c + up];
                const T pt = first ? r[c + up] : r[c + up] + b * pot;
                const T szt = sz[c + plane];
                const T pw = pn(r0 + iw), pe = pn(r0 + ie);
                const T ps = pn(c + dS), pnn = pn(c + dN);
                const T sxw = sx[fx], sxe = sx[fx + 1];
                const T sys = sy[fy], syn = sy[fy + nx];
                const T y = pc - sxe * pe - sxw * pw - syn * pnn - sys * ps - szt * pt - szb * pb;
                q[c] = y;
                p_new[c] = pc;
                if (do_e) e[c] = e[c] - a_prev * poc;
                acc += (double)pc * (double)y;
                pb = pc;
                pc = pt;
                poc = pot;
                szb = szt;
                c += plane;
                fx += (long long)ny * (nx + 1);
                fy += (long long)(ny + 1) * nx;
            }
        }
    }
    if (stop) return;
    const double s = block_sum(acc);
    if (threadIdx.x == 0) pap_part[blockIdx.x] = s;
}

// cg_b(n): alpha = r.r / p.Ap (9_grad.f90:450-455); r -= alpha q
// (:470-480); r.r partials (:482-508).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cg_b(long long len, int n, T *__restrict__ r, const T *__restrict__ q,
                                                   const double *__restrict__ pap_part, int n_pap,
                                                   double *__restrict__ partial, CgScalars *S)
{
    if (n >= S->done_at) return;
    const double pap = total_of(pap_part, n_pap);
    const double alpha = S->rr[n & 1] / pap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        S->pap = pap;
        S->alpha = alpha;
    }
    const T a = (T)alpha;
    constexpr int W = Vec<T>::W;
    const long long nv = len / W;
    const long long stride = (long long)gridDim.x * kThreads;
    double acc = 0.0;
    for (long long v = (long long)blockIdx.x * kThreads + threadIdx.x; v < nv; v += stride) {
        Vec<T> rv = reinterpret_cast<Vec<T> *>(r)[v];
        const Vec<T> qv = reinterpret_cast<const Vec<T> *>(q)[v];
#pragma unroll
        for (int l = 0; l < W; ++l) {
            rv.v[l] = rv.v[l] - a * qv.v[l];
            acc += (double)rv.v[l] * (double)rv.v[l];
        }
        reinterpret_cast<Vec<T> *>(r)[v] = rv;
    }
    if (blockIdx.x == 0) {
        const long long c = nv * W + threadIdx.x;
        if (c < len) {
            const T rv = r[c] - a * q[c];
            r[c] = rv;
            acc += (double)rv * (double)rv;
        }
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Mixed precision, residual replacement step 1: fold the FP32 correction
// into the FP64 iterate, e += c - alpha p (the deferred update of the last
// iteration, when cg_b ran last), c = 0.
__global__ __launch_bounds__(kThreads) void k_refresh_e(long long len, double *__restrict__ e,
                                                        float *__restrict__ c, const float *__restrict__ p_last,
                                                        const CgScalars *S)
{
    if (S->final_) return;
    const bool pending = S->done_at == kNotDone;
    const float a = (float)S->alpha;
    const long long stride = (long long)gridDim.x * kThreads;
    const long long nv = len / 4;
    for (long long v = (long long)blockIdx.x * kThreads + threadIdx.x; v < nv; v += stride) {
        Vec<float> cv = reinterpret_cast<const Vec<float> *>(c)[v];
        if (pending) {
            const Vec<float> pv = reinterpret_cast<const Vec<float> *>(p_last)[v];
#pragma unroll
            for (int l = 0; l < 4; ++l) cv.v[l] = cv.v[l] - a * pv.v[l];
        }
        Vec<double> e0 = reinterpret_cast<Vec<double> *>(e)[2 * v];
        Vec<double> e1 = reinterpret_cast<Vec<double> *>(e)[2 * v + 1];
        e0.v[0] += (double)cv.v[0];
        e0.v[1] += (double)cv.v[1];
        e1.v[0] += (double)cv.v[2];
        e1.v[1] += (double)cv.v[3];
        reinterpret_cast<Vec<double> *>(e)[2 * v] = e0;
        reinterpret_cast<Vec<double> *>(e)[2 * v + 1] = e1;
        reinterpret_cast<Vec<float> *>(c)[v] = Vec<float>{{0.f, 0.f, 0.f, 0.f}};
    }
    if (blockIdx.x == 0) {
        const long long x = nv * 4 + threadIdx.x;
        if (x < len) {
            float cv = c[x];
            if (pending) cv = cv - a * p_last[x];
            e[x] += (double)cv;
            c[x] = 0.0f;
        }
    }
}

// FP64 true residual r = A e - b (mirrored ghosts), same tiling as cg_a.
// Partials (3 slots of gridDim.x): r.r; when r32 is given, r.r of its FP32
// copy handed to the inner recurrence; when d is given, e.sqrt(d).
// check_final: skip once the solve is over.
__global__ __launch_bounds__(kThreads) void k_true_residual(Geo g, const double *__restrict__ e,
                                                            const double *__restrict__ d,
                                                            const double *__restrict__ sx,
                                                            const double *__restrict__ sy,
                                                            const double *__restrict__ sz,
                                                            const double *__restrict__ b, float *__restrict__ r32,
                                                            double *__restrict__ partial, int check_final,
                                                            const CgScalars *S)
{
    if (check_final && S->final_) return;
    const Tile t = tile_of(g);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    double acc = 0.0, acc32 = 0.0, acc_en = 0.0;
    if (t.i < nx && t.j < ny) {
        const long long plane = (long long)nx * ny;
        const int iw = nb_lo(t.i, nx, g.px), ie = nb_hi(t.i, nx, g.px);
        const int js = nb_lo(t.j, ny, g.py), jn = nb_hi(t.j, ny, g.py);
        const long long dS = (long long)(js - t.j) * nx, dN = (long long)(jn - t.j) * nx;
        long long c = (long long)t.k0 * plane + (long long)t.j * nx + t.i;
        long long fx = ((long long)t.k0 * ny + t.j) * (nx + 1) + t.i;
        long long fy = ((long long)t.k0 * (ny + 1) + t.j) * nx + t.i;
        for (int k = t.k0; k < t.k1; ++k) {
            const long long r0 = c - t.i;
            const double ec = e[c];
            const double et = (k < nz - 1) ? e[c + plane] : ec;
            const double eb = (k > 0) ? e[c - plane] : ec;
            const double r = ec - sx[fx + 1] * e[r0 + ie] - sx[fx] * e[r0 + iw] - sy[fy + nx] * e[c + dN] -
                             sy[fy] * e[c + dS] - sz[c + plane] * et - sz[c] * eb - b[c];
            acc += r * r;
            if (r32) {
                const float rf = (float)r;
                r32[c] = rf;
                acc32 += (double)rf * (double)rf;
            }
            if (d) acc_en += ec * sqrt(d[c]);
            c += plane;
            fx += (long long)ny * (nx + 1);
            fy += (long long)(ny + 1) * nx;
        }
    }
    const double s = block_sum(acc);
    const double s32 = block_sum(acc32);
    const double sen = block_sum(acc_en);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s;
        partial[gridDim.x + blockIdx.x] = s32;
        partial[2 * gridDim.x + blockIdx.x] = sen;
    }
}

// One workgroup after the residual replacement following iteration n_last:
// converged (on the FP64 true residual), capped, or continue with the
// replaced residual r (its r.r becomes the "entering" r.r of the next cg_a,
// through cg_b's partial slots).
__global__ __launch_bounds__(kThreads) void k_refresh_decide(int n_last, const double *__restrict__ partial, int nb,
                                                             double *__restrict__ upd_part, int n_upd, CgScalars *S)
{
    if (S->final_) return;
    const double rrt = total_of(partial, nb);
    const double rrs = total_of(partial + nb, nb);
    const bool pending = S->done_at == kNotDone;
    const int iters = pending ? n_last : S->iters;
    const bool conv = !(fabs(rrt) > S->thresh);
    const bool cap = iters >= S->max_iter;
    if (!conv && !cap)
        for (int t = threadIdx.x; t < n_upd; t += kThreads) upd_part[t] = t == 0 ? rrs : 0.0;
    if (threadIdx.x == 0) {
        S->rr_true = rrt;
        S->refreshes += 1;
        S->iters = iters;
        if (conv || cap) {
            S->final_ = 1;
            S->status = conv ? 0 : 1;
            // launches after n_last skip (cg_a tests done_at only)
            if (pending) S->done_at = n_last;
        }
        // otherwise continue: the deferred alpha p of iteration n_last is
        // already in e, so the host launches cg_a(n_last + 1) without it
    }
}

// After the solve: record the true r.r and the gauge shift.
__global__ __launch_bounds__(kThreads) void k_finish(const double *__restrict__ partial, int nb, CgScalars *S)
{
    const double rrt = total_of(partial, nb);
    const double en = total_of(partial + 2 * nb, nb);
    if (threadIdx.x == 0) {
        S->rr_true = rrt;
        // e's coordinate along sqrt(d) moved by rounding (and by FP32
        // corrections): shift prd1 by the matching constant
        S->gshift = S->project ? (en - S->en0) / S->dd : 0.0;
    }
}

// Plain SpMV q = A p (timing only).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_spmv(Geo g, const T *__restrict__ p, const T *__restrict__ sx,
                                                   const T *__restrict__ sy, const T *__restrict__ sz,
                                                   T *__restrict__ q, double *__restrict__ partial)
{
    const Tile t = tile_of(g);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    double acc = 0.0;
    if (t.i < nx && t.j < ny) {
        const long long plane = (long long)nx * ny;
        const int iw = nb_lo(t.i, nx, g.px), ie = nb_hi(t.i, nx, g.px);
        const int js = nb_lo(t.j, ny, g.py), jn = nb_hi(t.j, ny, g.py);
        const long long dS = (long long)(js - t.j) * nx, dN = (long long)(jn - t.j) * nx;
        long long c = (long long)t.k0 * plane + (long long)t.j * nx + t.i;
        long long fx = ((long long)t.k0 * ny + t.j) * (nx + 1) + t.i;
        long long fy = ((long long)t.k0 * (ny + 1) + t.j) * nx + t.i;
        T pc = p[c];
        T pb = t.k0 > 0 ? p[c - plane] : pc;
        T szb = sz[c];
        for (int k = t.k0; k < t.k1; ++k) {
            const long long r0 = c - t.i;
            const T pt = (k < nz - 1) ? p[c + plane] : pc;
            const T szt = sz[c + plane];
            const T y = pc - sx[fx + 1] * p[r0 + ie] - sx[fx] * p[r0 + iw] - sy[fy + nx] * p[c + dN] -
                        sy[fy] * p[c + dS] - szt * pt - szb * pb;
            q[c] = y;
            acc += (double)pc * (double)y;
            pb = pc;
            pc = pt;
            szb = szt;
            c += plane;
            fx += (long long)ny * (nx + 1);
            fy += (long long)(ny + 1) * nx;
        }
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

int vec_blocks(long long n)
{
    long long b = (n / 4 + kThreads - 1) / kThreads;
    return (int)std::max<long long>(1, std::min<long long>(kVecBlocks, b));
}

}  // namespace

Geo make_geo(int nx, int ny, int nz, int px, int py)
{
    Geo g{};
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.px = px;
    g.py = py;
    g.n = (long long)nx * ny * nz;
    g.tiles_x = (nx + kTileX - 1) / kTileX;
    g.tiles_y = (ny + kTileY - 1) / kTileY;
    // about kSpmvBlocks workgroups (4 per CU), at least 4 planes each
    const long long cols = (long long)g.tiles_x * g.tiles_y;
    int kz = (int)((cols * nz + kSpmvBlocks - 1) / kSpmvBlocks);
    if (kz < 4) kz = 4;
    if (kz > nz) kz = nz;
    g.kz = kz;
    g.tiles_z = (nz + kz - 1) / kz;
    return g;
}

int spmv_blocks(const Geo &g) { return g.tiles_x * g.tiles_y * g.tiles_z; }

template <typename T>
void launch_cg_a(const Geo &g, int n, bool do_e, T *e, const T *p_old, T *p_new, const T *r, Op<T> op, T *q,
                 const Partials &P, CgScalars *S, int *ran, hipStream_t s)
{
    const dim3 grid(spmv_blocks(g)), block(kThreads);
    if (n == 1)
        hipLaunchKernelGGL((k_cg_a<T, true, false>), grid, block, 0, s, g, n, e, p_old, p_new, r, op.sx, op.sy, op.sz,
                           q, P.upd, vec_blocks(g.n), P.spmv, S, ran);
    else if (do_e)
        hipLaunchKernelGGL((k_cg_a<T, false, true>), grid, block, 0, s, g, n, e, p_old, p_new, r, op.sx, op.sy,
                           op.sz, q, P.upd, vec_blocks(g.n), P.spmv, S, ran);
    else
        hipLaunchKernelGGL((k_cg_a<T, false, false>), grid, block, 0, s, g, n, e, p_old, p_new, r, op.sx, op.sy,
                           op.sz, q, P.upd, vec_blocks(g.n), P.spmv, S, ran);
}

template <typename T>
void launch_cg_b(const Geo &g, int n, T *r, const T *q, const Partials &P, CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_cg_b<T>, dim3(vec_blocks(g.n)), dim3(kThreads), 0, s, g.n, n, r, q, P.spmv,
                       spmv_blocks(g), P.upd, S);
}

void launch_refresh(const Geo &g, int n_last, double *e, float *c32, const float *p_last, Op<double> op,
                    const double *b, float *r32, const Partials &P, CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_refresh_e, dim3(vec_blocks(g.n)), dim3(kThreads), 0, s, g.n, e, c32, p_last, S);
    hipLaunchKernelGGL(k_true_residual, dim3(spmv_blocks(g)), dim3(kThreads), 0, s, g, e, (const double *)nullptr,
                       op.sx, op.sy, op.sz, b, r32, P.tres, 1, S);
    hipLaunchKernelGGL(k_refresh_decide, dim3(1), dim3(kThreads), 0, s, n_last, P.tres, spmv_blocks(g), P.upd,
                       vec_blocks(g.n), S);
}

void launch_finish(const Geo &g, const double *e, const double *d, Op<double> op, const double *b,
                   const Partials &P, CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_true_residual, dim3(spmv_blocks(g)), dim3(kThreads), 0, s, g, e, d, op.sx, op.sy, op.sz,
                       b, (float *)nullptr, P.tres, 0, S);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kThreads), 0, s, P.tres, spmv_blocks(g), S);
}

template <typename T>
void launch_spmv_only(const Geo &g, const T *p, Op<T> op, T *q, const Partials &P, hipStream_t s)
{
    hipLaunchKernelGGL(k_spmv<T>, dim3(spmv_blocks(g)), dim3(kThreads), 0, s, g, p, op.sx, op.sy, op.sz, q, P.spmv);
}

template void launch_cg_a<double>(const Geo &, int, bool, double *, const double *, double *, const double *, Op<double>,
                                  double *, const Partials &, CgScalars *, int *, hipStream_t);
template void launch_cg_a<float>(const Geo &, int, bool, float *, const float *, float *, const float *, Op<float>,
                                 float *, const Partials &, CgScalars *, int *, hipStream_t);
template void launch_cg_b<double>(const Geo &, int, double *, const double *, const Partials &, CgScalars *,
                                  hipStream_t);
template void launch_cg_b<float>(const Geo &, int, float *, const float *, const Partials &, CgScalars *,
                                 hipStream_t);
template void launch_spmv_only<double>(const Geo &, const double *, Op<double>, double *, const Partials &,
                                       hipStream_t);
template void launch_spmv_only<float>(const Geo &, const float *, Op<float>, float *, const Partials &,
                                      hipStream_t);

}  // namespace sps
