// CG iteration kernels (the hot path).
//
// One CG iteration n is three launches (SuLi-IC/9_grad.f90:398-587
// restated for a bandwidth-bound device):
//   spmv_pap   q = A p, one p.q partial per workgroup
//   update_r   alpha = r.r / p.q; r -= alpha q, one r.r partial per workgroup
//   update_ep  beta = r.r(new) / r.r(old); e -= alpha p; p = r + beta p;
//              workgroup 0 records r.r and applies the stop rule
// moving 5 + 3 + 5 = 13 words per cell instead of the reference's 17.  The
// reference's ghost copies of p (9_grad.f90:531-579) are replaced by clamped
// or wrapped neighbour indices inside the SpMV.  Totals of partial sums are
// formed by the consuming kernel (see CgScalars), so no kernel needs a
// device-wide fence.
//
// T is double for the FP64 solver and float for the inner solve of the
// mixed-precision one; dot products and the CG scalars are always FP64.
#include <algorithm>

#include "ps_device.h"

namespace sps {
namespace {

template <typename T>
struct alignas(16) Vec {
    static constexpr int W = 16 / sizeof(T);
    T v[W];
};

// SpMV over a 64 x 4 column tile that marches kz planes in z.  Each lane
// keeps p(k-1), p(k), p(k+1) and the shared z-face coefficient in registers;
// x/y neighbours come from the same or adjacent rows (L1/L2 hits).  Tiles
// are ordered x, y, z and assigned to XCDs in contiguous ranges.
//   (A p)_c = p_c - sx_E p_E - sx_W p_W - sy_N p_N - sy_S p_S - sz_T p_T - sz_B p_B
// (9_grad.f90:408-420; unit diagonal after the symmetric Jacobi scaling).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_spmv_pap(Geo g, int n, const T *__restrict__ p,
                                                       const T *__restrict__ sx, const T *__restrict__ sy,
                                                       const T *__restrict__ sz, T *__restrict__ q,
                                                       double *__restrict__ partial, const CgScalars *S, int *ran)
{
    if (n > S->done_at) return;
    if (ran && blockIdx.x == 0 && threadIdx.x == 0) *ran = 1;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % g.tiles_x;
    const int rest = tile / g.tiles_x;
    const int ty = rest % g.tiles_y;
    const int tz = rest / g.tiles_y;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int i = tx * kTileX + (threadIdx.x & (kTileX - 1));
    const int j = ty * kTileY + (threadIdx.x / kTileX);
    const int k0 = tz * g.kz;
    const int k1 = min(k0 + g.kz, nz);
    double acc = 0.0;
    if (i < nx && j < ny) {
        const long long plane = (long long)nx * ny;
        const int iw = nb_lo(i, nx, g.px), ie = nb_hi(i, nx, g.px);
        const int js = nb_lo(j, ny, g.py), jn = nb_hi(j, ny, g.py);
        const long long dS = (long long)(js - j) * nx, dN = (long long)(jn - j) * nx;
        long long c = (long long)k0 * plane + (long long)j * nx + i;
        long long fx = ((long long)k0 * ny + j) * (nx + 1) + i;
        long long fy = ((long long)k0 * (ny + 1) + j) * nx + i;
        T pc = p[c];
        T pb = k0 > 0 ? p[c - plane] : pc;
        T szb = sz[c];
        for (int k = k0; k < k1; ++k) {
            const long long r0 = c - i;
            const T pt = (k < nz - 1) ? p[c + plane] : pc;
            const T szt = sz[c + plane];
            const T pw = p[r0 + iw], pe = p[r0 + ie];
            const T ps = p[c + dS], pn = p[c + dN];
            const T sxw = sx[fx], sxe = sx[fx + 1];
            const T sys = sy[fy], syn = sy[fy + nx];
            const T y = pc - sxe * pe - sxw * pw - syn * pn - sys * ps - szt * pt - szb * pb;
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

// alpha = r.r / p.Ap (9_grad.f90:450-455); r -= alpha q (:470-480); r.r
// partials (:482-508).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_update_r(long long len, int n, T *__restrict__ r,
                                                       const T *__restrict__ q, const double *__restrict__ pap_part,
                                                       int n_pap, double *__restrict__ partial, CgScalars *S)
{
    if (n > S->done_at) return;
    const double pap = total_of(pap_part, n_pap);
    const double alpha = S->rr[(n - 1) & 1] / pap;
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

// beta = r.r(new) / r.r(old) (9_grad.f90:510-515); e -= alpha p (:458-468);
// p = r + beta p (:518-527).  Workgroup 0 records r.r for iteration n and
// applies the stop rule (9_grad.f90:398,582): the iteration that meets it
// still updates e here.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_update_ep(long long len, int n, T *__restrict__ e, T *__restrict__ p,
                                                        const T *__restrict__ r, const double *__restrict__ rr_part,
                                                        int n_rr, CgScalars *S)
{
    if (n > S->done_at) return;
    const double rrn = total_of(rr_part, n_rr);
    const double rro = S->rr[(n - 1) & 1];
    const double beta = rrn / rro;
    const T a = (T)S->alpha;
    const T b = (T)beta;
    constexpr int W = Vec<T>::W;
    const long long nv = len / W;
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long v = (long long)blockIdx.x * kThreads + threadIdx.x; v < nv; v += stride) {
        Vec<T> ev = reinterpret_cast<Vec<T> *>(e)[v];
        Vec<T> pv = reinterpret_cast<Vec<T> *>(p)[v];
        const Vec<T> rv = reinterpret_cast<const Vec<T> *>(r)[v];
#pragma unroll
        for (int l = 0; l < W; ++l) {
            ev.v[l] = ev.v[l] - a * pv.v[l];
            pv.v[l] = rv.v[l] + b * pv.v[l];
        }
        reinterpret_cast<Vec<T> *>(e)[v] = ev;
        reinterpret_cast<Vec<T> *>(p)[v] = pv;
    }
    if (blockIdx.x == 0) {
        const long long c = nv * W + threadIdx.x;
        if (c < len) {
            e[c] = e[c] - a * p[c];
            p[c] = r[c] + b * p[c];
        }
        if (threadIdx.x == 0) {
            S->rr[n & 1] = rrn;
            S->beta = beta;
            S->iters += 1;
            S->since_refresh += 1;
            const bool conv = !(fabs(rrn) > S->thresh);
            const bool cap = S->iters >= S->max_iter;
            if (conv || cap) {
                S->done_at = n;
                if (!S->mixed) {
                    // FP64: the recurrence's r.r is the reference's stop test
                    S->final_ = 1;
                    S->status = conv ? 0 : 1;
                }
            }
        }
    }
}

// Mixed precision: fold the FP32 correction into the FP64 iterate,
// e += c, c = 0 (skipped when no iteration ran since the last replacement).
__global__ __launch_bounds__(kThreads) void k_refresh_e(long long len, double *__restrict__ e,
                                                        float *__restrict__ c, const CgScalars *S)
{
    if (S->final_ || S->since_refresh == 0) return;
    const long long nv = len / 4;
    const long long stride = (long long)gridDim.x * kThreads;
    for (long long v = (long long)blockIdx.x * kThreads + threadIdx.x; v < nv; v += stride) {
        const Vec<float> cv = reinterpret_cast<const Vec<float> *>(c)[v];
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
        const long long t = nv * 4 + threadIdx.x;
        if (t < len) {
            e[t] += (double)c[t];
            c[t] = 0.0f;
        }
    }
}

// FP64 true residual r = A e - b (mirrored ghosts), same tiling as the SpMV;
// partials of r.r and, when r32 is given, of its FP32 copy handed to the
// inner recurrence.  check_skip: skip like the other replacement kernels.
__global__ __launch_bounds__(kThreads) void k_true_residual(Geo g, const double *__restrict__ e,
                                                            const double *__restrict__ sx,
                                                            const double *__restrict__ sy,
                                                            const double *__restrict__ sz,
                                                            const double *__restrict__ b, float *__restrict__ r32,
                                                            double *__restrict__ partial, int check_skip,
                                                            const CgScalars *S)
{
    if (check_skip && (S->final_ || S->since_refresh == 0)) return;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int tx = tile % g.tiles_x;
    const int rest = tile / g.tiles_x;
    const int ty = rest % g.tiles_y;
    const int tz = rest / g.tiles_y;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int i = tx * kTileX + (threadIdx.x & (kTileX - 1));
    const int j = ty * kTileY + (threadIdx.x / kTileX);
    const int k0 = tz * g.kz;
    const int k1 = min(k0 + g.kz, nz);
    double acc = 0.0, acc32 = 0.0;
    if (i < nx && j < ny) {
        const long long plane = (long long)nx * ny;
        const int iw = nb_lo(i, nx, g.px), ie = nb_hi(i, nx, g.px);
        const int js = nb_lo(j, ny, g.py), jn = nb_hi(j, ny, g.py);
        const long long dS = (long long)(js - j) * nx, dN = (long long)(jn - j) * nx;
        long long c = (long long)k0 * plane + (long long)j * nx + i;
        long long fx = ((long long)k0 * ny + j) * (nx + 1) + i;
        long long fy = ((long long)k0 * (ny + 1) + j) * nx + i;
        for (int k = k0; k < k1; ++k) {
            const long long r0 = c - i;
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
            c += plane;
            fx += (long long)ny * (nx + 1);
            fy += (long long)(ny + 1) * nx;
        }
    }
    const double s = block_sum(acc);
    const double s32 = block_sum(acc32);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s;
        partial[gridDim.x + blockIdx.x] = s32;
    }
}

// One workgroup after the residual replacement: converged (on the FP64 true
// residual), capped, or continue with the replaced residual.
__global__ __launch_bounds__(kThreads) void k_refresh_decide(const double *__restrict__ partial, int nb,
                                                             CgScalars *S)
{
    if (S->final_ || S->since_refresh == 0) return;
    const double rrt = total_of(partial, nb);
    const double rrs = total_of(partial + nb, nb);
    if (threadIdx.x == 0) {
        S->rr_true = rrt;
        S->refreshes += 1;
        S->since_refresh = 0;
        if (!(fabs(rrt) > S->thresh)) {
            S->final_ = 1;
            S->status = 0;
        } else if (S->iters >= S->max_iter) {
            S->final_ = 1;
            S->status = 1;
        } else {
            // resume: the next iteration reads r.r from either parity slot
            S->rr[0] = rrs;
            S->rr[1] = rrs;
            S->done_at = kNotDone;
        }
    }
}

__global__ __launch_bounds__(kThreads) void k_record_true(const double *__restrict__ partial, int nb, CgScalars *S)
{
    const double rrt = total_of(partial, nb);
    if (threadIdx.x == 0) S->rr_true = rrt;
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
void launch_spmv_pap(const Geo &g, int n, const T *p, Op<T> op, T *q, const Partials &P, CgScalars *S, int *ran,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_spmv_pap<T>, dim3(spmv_blocks(g)), dim3(kThreads), 0, s, g, n, p, op.sx, op.sy, op.sz, q,
                       P.spmv, S, ran);
}

template <typename T>
void launch_update_r(const Geo &g, int n, T *r, const T *q, const Partials &P, CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_update_r<T>, dim3(vec_blocks(g.n)), dim3(kThreads), 0, s, g.n, n, r, q, P.spmv,
                       spmv_blocks(g), P.upd, S);
}

template <typename T>
void launch_update_ep(const Geo &g, int n, T *e, T *p, const T *r, const Partials &P, CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_update_ep<T>, dim3(vec_blocks(g.n)), dim3(kThreads), 0, s, g.n, n, e, p, r, P.upd,
                       vec_blocks(g.n), S);
}

void launch_refresh(const Geo &g, double *e, float *c32, Op<double> op, const double *b, float *r32,
                    const Partials &P, CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_refresh_e, dim3(vec_blocks(g.n)), dim3(kThreads), 0, s, g.n, e, c32, S);
    hipLaunchKernelGGL(k_true_residual, dim3(spmv_blocks(g)), dim3(kThreads), 0, s, g, e, op.sx, op.sy, op.sz, b,
                       r32, P.tres, 1, S);
    hipLaunchKernelGGL(k_refresh_decide, dim3(1), dim3(kThreads), 0, s, P.tres, spmv_blocks(g), S);
}

void launch_record_true(const Geo &g, const double *e, Op<double> op, const double *b, const Partials &P,
                        CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_true_residual, dim3(spmv_blocks(g)), dim3(kThreads), 0, s, g, e, op.sx, op.sy, op.sz, b,
                       (float *)nullptr, P.tres, 0, S);
    hipLaunchKernelGGL(k_record_true, dim3(1), dim3(kThreads), 0, s, P.tres, spmv_blocks(g), S);
}

template void launch_spmv_pap<double>(const Geo &, int, const double *, Op<double>, double *, const Partials &,
                                      CgScalars *, int *, hipStream_t);
template void launch_spmv_pap<float>(const Geo &, int, const float *, Op<float>, float *, const Partials &,
                                     CgScalars *, int *, hipStream_t);
template void launch_update_r<double>(const Geo &, int, double *, const double *, const Partials &, CgScalars *,
                                      hipStream_t);
template void launch_update_r<float>(const Geo &, int, float *, const float *, const Partials &, CgScalars *,
                                     hipStream_t);
template void launch_update_ep<double>(const Geo &, int, double *, double *, const double *, const Partials &,
                                       CgScalars *, hipStream_t);
template void launch_update_ep<float>(const Geo &, int, float *, float *, const float *, const Partials &,
                                      CgScalars *, hipStream_t);

}  // namespace sps
