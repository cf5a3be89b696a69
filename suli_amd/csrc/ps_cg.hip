// CG iteration kernels (the hot path).
//
// One CG iteration is three launches (SuLi-IC/9_grad.f90:398-587 restated
// for a bandwidth-bound device):
//   spmv_pap   q = A p and p.q; the last workgroup forms alpha = r.r / p.q
//   update_r   r -= alpha q and r.r; the last workgroup forms beta, tests the
//              stop rule
//   update_ep  e -= alpha p, p = r + beta p
// Moving 5 + 3 + 5 = 13 words per cell instead of the reference's 17.  The
// reference's ghost copies of p (9_grad.f90:531-579) are replaced by clamped
// or wrapped neighbour indices inside the SpMV.
//
// T is double for the FP64 solver and float for the inner solve of the
// mixed-precision one; dot products and the CG scalars are always FP64.
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
__global__ __launch_bounds__(kThreads) void k_spmv_pap(Geo g, const T *__restrict__ p, const T *__restrict__ sx,
                                                       const T *__restrict__ sy, const T *__restrict__ sz,
                                                       T *__restrict__ q, double *__restrict__ partial, CgScalars *S)
{
    if (S->stop | S->final_) return;
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
        const long long row = (long long)k0 * plane + (long long)j * nx;
        const long long dS = (long long)(js - j) * nx, dN = (long long)(jn - j) * nx;
        long long c = row + i;
        long long fx = ((long long)k0 * ny + j) * (nx + 1) + i;
        long long fy = ((long long)k0 * (ny + 1) + j) * nx + i;
        long long fz = c;
        T pc = p[c];
        T pb = k0 > 0 ? p[c - plane] : pc;
        T szb = sz[fz];
        for (int k = k0; k < k1; ++k) {
            const long long r0 = c - i;
            const T pt = (k < nz - 1) ? p[c + plane] : pc;
            const T szt = sz[fz + plane];
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
            fz += plane;
            fx += (long long)ny * (nx + 1);
            fy += (long long)(ny + 1) * nx;
        }
    }
    const double s = block_sum(acc);
    if (publish_and_ticket(partial, s, nullptr, 0.0, &S->ticket[kTkSpmv])) {
        const double pap = sum_partials(partial, gridDim.x);
        if (threadIdx.x == 0) {
            S->pap = pap;
            S->alpha = S->rr / pap;   // 9_grad.f90:450-455
            finish_ticket(&S->ticket[kTkSpmv]);
        }
    }
}

// r -= alpha q; r.r; beta and the stop rule (9_grad.f90:470-515,582).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_update_r(long long n, T *__restrict__ r, const T *__restrict__ q,
                                                       double *__restrict__ partial, CgScalars *S)
{
    if (S->stop | S->final_) return;
    const T a = (T)S->alpha;
    constexpr int W = Vec<T>::W;
    const long long nv = n / W;
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
        if (c < n) {
            const T rv = r[c] - a * q[c];
            r[c] = rv;
            acc += (double)rv * (double)rv;
        }
    }
    const double s = block_sum(acc);
    if (publish_and_ticket(partial, s, nullptr, 0.0, &S->ticket[kTkUpdR])) {
        const double rrn = sum_partials(partial, gridDim.x);
        if (threadIdx.x == 0) {
            S->beta = rrn / S->rr;
            S->rr = rrn;
            S->k2n += 1;
            S->iters += 1;
            const bool conv = !(fabs(rrn) > S->thresh);
            const bool cap = S->iters >= S->max_iter;
            if (S->mixed) {
                // the FP64 residual recomputation decides (k_true_residual)
                if (conv || cap) S->stop = 1;
            } else if (conv || cap) {
                S->stop = 1;
                S->final_ = 1;
                S->status = conv ? 0 : 1;
            }
            finish_ticket(&S->ticket[kTkUpdR]);
        }
    }
}

// e -= alpha p; p = r + beta p (9_grad.f90:458-468, 518-527).  Runs once for
// every completed update_r, including the one that met the stop rule.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_update_ep(long long n, T *__restrict__ e, T *__restrict__ p,
                                                        const T *__restrict__ r, CgScalars *S)
{
    if (S->k2n <= S->k3n) return;
    const T a = (T)S->alpha;
    const T b = (T)S->beta;
    constexpr int W = Vec<T>::W;
    const long long nv = n / W;
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
        if (c < n) {
            e[c] = e[c] - a * p[c];
            p[c] = r[c] + b * p[c];
        }
    }
    if (publish_and_ticket(nullptr, 0.0, nullptr, 0.0, &S->ticket[kTkUpdEP])) {
        if (threadIdx.x == 0) {
            S->k3n += 1;
            finish_ticket(&S->ticket[kTkUpdEP]);
        }
    }
}

// Mixed precision: fold the FP32 correction into the FP64 iterate,
// e += c, c = 0.
__global__ __launch_bounds__(kThreads) void k_refresh_e(long long n, double *__restrict__ e, float *__restrict__ c,
                                                        CgScalars *S)
{
    if (S->final_ || S->k3n == S->refresh_mark) return;
    const long long nv = n / 4;
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
        if (t < n) {
            e[t] += (double)c[t];
            c[t] = 0.0f;
        }
    }
}

// FP64 true residual r = A e - b (mirrored ghosts), same tiling as the SpMV.
// Mixed precision: r is handed to the FP32 recurrence (residual
// replacement) and r.r in FP64 decides convergence.  final_check: only
// record r.r (used after an FP64 solve to report the true residual).
__global__ __launch_bounds__(kThreads) void k_true_residual(Geo g, const double *__restrict__ e,
                                                            const double *__restrict__ sx,
                                                            const double *__restrict__ sy,
                                                            const double *__restrict__ sz,
                                                            const double *__restrict__ b, float *__restrict__ r32,
                                                            double *__restrict__ partial, int final_check,
                                                            CgScalars *S)
{
    if (!final_check && S->final_) return;
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
    if (publish_and_ticket(partial, s, partial + gridDim.x, s32, &S->ticket[kTkTrue])) {
        const double rrt = sum_partials(partial, gridDim.x);
        const double rrs = sum_partials(partial + gridDim.x, gridDim.x);
        if (threadIdx.x == 0) {
            S->rr_true = rrt;
            if (!final_check) {
                S->refreshes += 1;
                S->refresh_mark = S->k3n;
                if (!(fabs(rrt) > S->thresh)) {
                    S->stop = 1;
                    S->final_ = 1;
                    S->status = 0;
                } else if (S->iters >= S->max_iter) {
                    S->stop = 1;
                    S->final_ = 1;
                    S->status = 1;
                } else {
                    S->rr = rrs;
                    S->stop = 0;
                }
            }
            finish_ticket(&S->ticket[kTkTrue]);
        }
    }
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
    // aim for >= 4 workgroups per CU (256 CUs) with at least 4 planes each
    const long long cols = (long long)g.tiles_x * g.tiles_y;
    int kz = (int)((cols * nz + 2047) / 2048);
    if (kz < 4) kz = 4;
    if (kz > nz) kz = nz;
    g.kz = kz;
    g.tiles_z = (nz + kz - 1) / kz;
    return g;
}

int spmv_blocks(const Geo &g) { return g.tiles_x * g.tiles_y * g.tiles_z; }

template <typename T>
void launch_spmv_pap(const Geo &g, const T *p, Op<T> op, T *q, double *partial, CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_spmv_pap<T>, dim3(spmv_blocks(g)), dim3(kThreads), 0, s, g, p, op.sx, op.sy, op.sz, q,
                       partial, S);
}

template <typename T>
void launch_update_r(const Geo &g, T *r, const T *q, double *partial, int nblk, CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_update_r<T>, dim3(nblk), dim3(kThreads), 0, s, g.n, r, q, partial, S);
}

template <typename T, typename TE>
void launch_update_ep(const Geo &g, TE *e, T *p, const T *r, int nblk, CgScalars *S, hipStream_t s)
{
    static_assert(sizeof(T) == sizeof(TE), "e and p share a precision");
    hipLaunchKernelGGL(k_update_ep<T>, dim3(nblk), dim3(kThreads), 0, s, g.n, (T *)e, p, r, S);
}

void launch_refresh_e(const Geo &g, double *e, float *c32, int nblk, CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_refresh_e, dim3(nblk), dim3(kThreads), 0, s, g.n, e, c32, S);
}

void launch_true_residual(const Geo &g, const double *e, Op<double> op, const double *b, float *r32,
                          double *partial, double * /*partial2*/, int final_check, CgScalars *S, hipStream_t s)
{
    hipLaunchKernelGGL(k_true_residual, dim3(spmv_blocks(g)), dim3(kThreads), 0, s, g, e, op.sx, op.sy, op.sz, b,
                       r32, partial, final_check, S);
}

template void launch_spmv_pap<double>(const Geo &, const double *, Op<double>, double *, double *, CgScalars *,
                                      hipStream_t);
template void launch_spmv_pap<float>(const Geo &, const float *, Op<float>, float *, double *, CgScalars *,
                                     hipStream_t);
template void launch_update_r<double>(const Geo &, double *, const double *, double *, int, CgScalars *,
                                      hipStream_t);
template void launch_update_r<float>(const Geo &, float *, const float *, double *, int, CgScalars *, hipStream_t);
template void launch_update_ep<double, double>(const Geo &, double *, double *, const double *, int, CgScalars *,
                                               hipStream_t);
template void launch_update_ep<float, float>(const Geo &, float *, float *, const float *, int, CgScalars *,
                                             hipStream_t);

}  // namespace sps
