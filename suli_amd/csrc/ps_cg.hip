// CG iteration kernels (the hot path), batched over problems.
//
// SuLi's CG iteration (SuLi-IC/9_grad.f90:398-587) is restated for a
// bandwidth-bound device as two launches per global step; a step runs one
// iteration n of every problem in the launch table:
//   cg_a  finishes iteration n-1 (e -= alpha p, deferred), tests the stop
//         rule on r.r, forms p_n = r + beta p_{n-1} (on the fly at every
//         stencil point, into the other p buffer), q = A p_n, and one p.q
//         partial per tile;
//   cg_b  alpha = r.r / p.q; r -= alpha q; one r.r partial per tile.
// That moves 9 + 3 = 12 words per cell instead of the reference's 17
// (cg_a reads e, p_old, r and the three face-coefficient arrays and writes
// e, p_new and q; cg_b reads r and q and writes r).  The reference's ghost
// copies of p (9_grad.f90:531-579) become clamped or wrapped neighbour
// indices.  Totals of partial sums are formed by the consuming kernel (see
// CgScalars), so no kernel needs a device-wide fence.
//
// Kernels are persistent: a fixed number of workgroups walks the launch
// table of (problem, tile) entries.  Each workgroup takes a contiguous run
// of entries and workgroups on one XCD take neighbouring runs, so tiles
// that share halo planes share an L2.  A problem's arithmetic depends only
// on its own tiling, never on which workgroup handles a tile or what else
// is in the batch: results are identical alone and in a mix.
//
// Precision: double for the FP64 solver; float for the inner solve of the
// mixed-precision one (e is then the FP32 correction c).  Dot products and
// the CG scalars are always FP64.
#include <algorithm>

#include "ps_device.h"

namespace sps {
namespace {

// 16 bytes of T as one vector (one 16-byte load or store)
template <typename T>
using Vec = T __attribute__((ext_vector_type(16 / sizeof(T))));

__device__ __forceinline__ int step_of(StepRef s)
{
    const int step = s.base ? *s.base + s.off : s.off;
    if (s.next && blockIdx.x == 0 && threadIdx.x == 0) *s.next = *s.base + kRefreshGrain;
    return step;
}

// Contiguous run [begin, end) of the launch table for this workgroup.
struct Run {
    int begin, end;
};

__device__ __forceinline__ Run my_run(int ntab)
{
    const int G = gridDim.x;
    const long long L = xcd_tile(blockIdx.x, G);
    Run r;
    r.begin = (int)(L * ntab / G);
    r.end = (int)((L + 1) * ntab / G);
    return r;
}

// 3-D tile `tile` of a problem: (64 vx) x 4 columns, planes [k0, k1); a
// lane owns x cells i .. i + vx - 1.  i0, the tile's first x, is the same
// in every lane (derived from the tile index alone, so the compiler keeps it
// and what follows from it in scalar registers).
struct Tile {
    int i, j, k0, k1, i0;
};

__device__ __forceinline__ Tile tile3(const Geo &g, int tile)
{
    const int tx = tile % g.tiles_x;
    const int rest = tile / g.tiles_x;
    const int ty = rest % g.tiles_y;
    const int tz = rest / g.tiles_y;
    Tile t;
    t.i0 = tx * kTileX * g.vx;
    t.i = (tx * kTileX + (threadIdx.x & (kTileX - 1))) * g.vx;
    t.j = ty * kTileY + (threadIdx.x / kTileX);
    t.k0 = tz * g.kz;
    t.k1 = min(t.k0 + g.kz, g.nz);
    return t;
}

// Face coefficients of the scaled operator as the iteration kernels read
// them: the three face arrays (index f of the face), or, for a constant-
// density operator, the per-plane tables (every x and y face of plane k
// holds tx[k], ty[k]; z face kf holds tz[kf]), which store the same values.
//
// FP32 face arrays come with row tables: per-plane values (repeated: 8 at
// rx + 8k, 4 at ry + 4k and rz + 4kf) and a bit per row (j, k), set when
// every face of the row holds them.  The kernels point the row's
// coefficient loads at the tables then: the same instructions and values,
// without the face arrays' traffic.
template <typename T>
struct FaceArrays {
    gptr<const T> sx, sy, sz;
    gptr<const T> rx = nullptr, ry = nullptr, rz = nullptr;
    cptr<unsigned> rowreg = nullptr;
    int wpp = 0;   // words of rowreg per plane
    __device__ __forceinline__ T x(long long f, int) const { return sx[f]; }
    __device__ __forceinline__ T y(long long f, int) const { return sy[f]; }
    __device__ __forceinline__ T z(long long f, int) const { return sz[f]; }
    // Bit k - k0 set when every face of row j of plane k holds the plane's
    // table values, for the planes k0 .. k1-1 of a tile (up to 64; later
    // planes read the arrays).  j is the wave's row, the same in all its
    // lanes: lane l fetches plane k0 + l's bit, the ballot gathers them.
    __device__ __forceinline__ unsigned long long rows(int j, int k0, int k1) const
    {
        if (!rowreg) return 0ull;
        const int ju = __builtin_amdgcn_readfirstlane(j);
        const int k = k0 + (int)(threadIdx.x & 63);
        const bool bit = k < k1 && ((rowreg[(long long)k * wpp + ju / 32] >> (ju % 32)) & 1u);
        return __ballot(bit);
    }
    // where a row's coefficients are read: the tables (reg) or the arrays
    __device__ __forceinline__ gptr<const T> xp(bool reg, long long f, int k) const { return reg ? rx + 8 * k : sx + f; }
    __device__ __forceinline__ gptr<const T> yp(bool reg, long long f, int k) const { return reg ? ry + 4 * k : sy + f; }
    __device__ __forceinline__ gptr<const T> zp(bool reg, long long f, int kf) const { return reg ? rz + 4 * kf : sz + f; }
    __device__ __forceinline__ T x_at(bool reg, long long f, int k, int m) const { return xp(reg, f, k)[m]; }
    __device__ __forceinline__ T y_at(bool reg, long long f, int k) const { return *yp(reg, f, k); }
    __device__ __forceinline__ T z_at(bool reg, long long f, int kf) const { return *zp(reg, f, kf); }
};
template <typename T>
struct PlaneTables {
    cptr<T> tx, ty, tz;
    __device__ __forceinline__ unsigned long long rows(int, int, int) const { return ~0ull; }
    __device__ __forceinline__ T x_at(bool, long long, int k, int) const { return tx[k]; }
    __device__ __forceinline__ T y_at(bool, long long, int k) const { return ty[k]; }
    __device__ __forceinline__ T z_at(bool, long long, int kf) const { return tz[kf]; }
    __device__ __forceinline__ T x(long long, int k) const { return tx[k]; }
    __device__ __forceinline__ T y(long long, int k) const { return ty[k]; }
    __device__ __forceinline__ T z(long long, int kf) const { return tz[kf]; }
};

// Plane k's bit in a tile's row mask (planes from k0)
__device__ __forceinline__ bool row_bit(unsigned long long rows, int k, int k0)
{
    const int d = k - k0;
    return d < 64 && ((rows >> d) & 1ull);
}

// Per-workgroup cache of a problem's partial-sum total (the FP64 kernels:
// one wave sums, the others wait at the workgroup barrier).
struct TotalCache {
    int prob = -1;
    double value = 0.0;
    __device__ __forceinline__ double get(int p, const double *partial, int n)
    {
        if (p != prob) {
            value = block_total_of(partial, n);
            prob = p;
        }
        return value;
    }
};

// One 3-D tile of cg_a.  Each lane marches its (i, j) column over the
// tile's planes keeping p_n at k-1, k, k+1 (and p_{n-1} at k) in registers;
// x/y neighbours of p_n are formed from r and p_{n-1} loaded from the same
// or adjacent rows (L1/L2 hits).
//   (A p)_c = p_c - sx_E p_E - sx_W p_W - sy_N p_N - sy_S p_S - sz_T p_T - sz_B p_B
// (9_grad.f90:408-420; unit diagonal after the symmetric Jacobi scaling).
// FIRST (iteration 1: p_1 = r_0, nothing to finish) and DO_E (apply the
// deferred e -= alpha p) are template parameters so that the plane loop has
// no branches: every load of a plane is issued before any is waited for.
template <typename T, bool FIRST, bool DO_E, class CO>
__device__ __forceinline__ double cg_a_tile(const Geo &g, const Tile &t, gptr<T> __restrict__ e,
                                            gptr<const T> __restrict__ p_old, gptr<T> __restrict__ p_new,
                                            gptr<const T> __restrict__ r, const CO &co, gptr<T> __restrict__ q, T b,
                                            T a_prev)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long plane = (long long)nx * ny;
    double acc = 0.0;
    if (t.i >= nx || t.j >= ny) return acc;
    long long c = (long long)t.k0 * plane + (long long)t.j * nx + t.i;
    const int iw = nb_lo(t.i, nx, g.px), ie = nb_hi(t.i, nx, g.px);
    const int js = nb_lo(t.j, ny, g.py), jn = nb_hi(t.j, ny, g.py);
    const long long dS = (long long)(js - t.j) * nx, dN = (long long)(jn - t.j) * nx;
    long long fx = ((long long)t.k0 * ny + t.j) * (nx + 1) + t.i;
    long long fy = ((long long)t.k0 * (ny + 1) + t.j) * nx + t.i;
    // p_n = r + beta p_{n-1} (9_grad.f90:518-527), formed where needed
    auto pn = [&](long long x) -> T { return FIRST ? r[x] : r[x] + b * p_old[x]; };
    T poc = FIRST ? T(0) : p_old[c];
    T pc = FIRST ? r[c] : r[c] + b * poc;
    T pb = t.k0 > 0 ? pn(c - plane) : pc;
    T szb = co.z(c, t.k0);
    for (int k = t.k0; k < t.k1; ++k) {
        const long long r0 = c - t.i;
        // top neighbour; at the lid the mirror ghost is the cell itself
        const long long up = (k < nz - 1) ? plane : 0;
        const T pot = FIRST ? T(0) : p_old[c + up];
        const T pt = FIRST ? r[c + up] : r[c + up] + b * pot;
        const T szt = co.z(c + plane, k + 1);
        const T pw = pn(r0 + iw), pe = pn(r0 + ie);
        const T ps = pn(c + dS), pnn = pn(c + dN);
        const T sxw = co.x(fx, k), sxe = co.x(fx + 1, k);
        const T sys = co.y(fy, k), syn = co.y(fy + nx, k);
        const T ptu = (g.dtop && k == nz - 1) ? -pc : pt;   // Dirichlet lid: ghost = -cell
        const T y = pc - sxe * pe - sxw * pw - syn * pnn - sys * ps - szt * ptu - szb * pb;
        q[c] = y;
        p_new[c] = pc;
        if (DO_E) e[c] = e[c] - a_prev * poc;   // 9_grad.f90:458-468
        acc += (double)pc * (double)y;
        pb = pc;
        pc = pt;
        poc = pot;
        szb = szt;
        c += plane;
        fx += (long long)ny * (nx + 1);
        fy += (long long)(ny + 1) * nx;
    }
    return acc;
}

template <typename T>
struct Vecs {
    gptr<T> e, r, q, p0, p1;
    gptr<const T> sx, sy, sz;
    cptr<T> tx, ty, tz;         // null unless the operator has plane tables
    gptr<const T> rx, ry, rz;   // row tables (FP32), null unless rowreg is set
    cptr<unsigned> rowreg;
    int wpp;
};

__device__ __forceinline__ Vecs<double> vecs(const ProbDev &pd, double)
{
    return {glob(pd.e),  glob(pd.r),  glob(pd.q),  glob(pd.p0), glob(pd.p1), glob(pd.sx),
            glob(pd.sy), glob(pd.sz), konst(pd.tx), konst(pd.ty), konst(pd.tz), nullptr,
            nullptr,     nullptr,     nullptr,     0};
}
__device__ __forceinline__ Vecs<float> vecs(const ProbDev &pd, float)
{
    return {glob(pd.c32),  glob(pd.r32),  glob(pd.q32),  glob(pd.p32),  glob(pd.p32 + pd.pslot),
            glob(pd.sx32), glob(pd.sy32), glob(pd.sz32), konst(pd.tx32), konst(pd.ty32),
            konst(pd.tz32), glob(pd.rx32), glob(pd.ry32), glob(pd.rz32), konst(pd.rowreg),
            (pd.g.ny + 31) / 32};
}

// Calls f with the problem's coefficient source (tables or face arrays).
template <typename T, class F>
__device__ __forceinline__ auto with_coefs(const Vecs<T> &V, F &&f)
{
    if (V.tx) return f(PlaneTables<T>{V.tx, V.ty, V.tz});
    return f(FaceArrays<T>{V.sx, V.sy, V.sz, V.rx, V.ry, V.rz, V.rowreg, V.wpp});
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_cg_a(const ProbDev *__restrict__ P, const TileRef *__restrict__ tab,
                                                   int ntab, StepRef sref, int *ran)
{
    const Run run = my_run(ntab);
    const int step = step_of(sref);
    TotalCache rr_tot;
    for (int t = run.begin; t < run.end; ++t) {
        const TileRef tr = tab[t];
        const ProbDev &pd = P[tr.prob];
        const gptr<CgScalars> S = glob(pd.S);
        const int n = step - pd.n0 + 1;
        // the problem's scalars, then the previous kernel's partials, both
        // requested before the scalars decide whether the problem runs, so
        // that the step starts after one memory round trip (as in cg_m)
        const int done_at = S->done_at, skip_e_at = S->skip_e_at, max_iter = S->max_iter;
        const int lag_stop = S->lag_stop, mixed = S->mixed;
        // (n - 1) & 1 is (sref.off - 1) & 1, known without the replayed step
        // counter (see cg_m)
        const double rr1 = S->rr[1], rr_prev = S->rr[(sref.off - 1) & 1], thresh = S->thresh;
        const double alpha_prev = S->alpha;
        const double rr_sum = rr_tot.get(tr.prob, pd.P.upd, pd.nt1);   // unused at n = 1
        if (n < 1 || n > done_at) continue;
        const bool first = n == 1;
        const bool do_e = !first && step != skip_e_at;
        // r.r entering this iteration: r_0.r_0 from setup, else cg_b's partials
        const double rr_in = first ? rr1 : rr_sum;
        const double beta = first ? 0.0 : rr_in / rr_prev;   // 9_grad.f90:510-515
        // stop rule (9_grad.f90:398,582): FP64 tests the recurrence's r.r;
        // the mixed solver decides at its FP64 residual replacements
        const double rr_test = (lag_stop && !first) ? rr_prev : rr_in;
        const bool conv = !(fabs(rr_test) > thresh);
        const bool cap = n - 1 >= max_iter;
        const bool stop = cap || (conv && !mixed);
        if (tr.tile == 0 && threadIdx.x == 0) {
            S->rr[n & 1] = rr_in;
            S->iters = n - 1;
            if (stop) {
                S->done_at = n;
                if (!mixed) {
                    S->final_ = 1;
                    S->status = conv ? 0 : 1;
                }
            }
            if (ran && tr.prob == 0 && !stop) *ran = 1;
        }
        const Vecs<T> V = vecs(pd, T());
        const gptr<T> p_old = (n & 1) ? V.p0 : V.p1;
        const gptr<T> p_new = (n & 1) ? V.p1 : V.p0;
        const T a_prev = (T)alpha_prev;
        const Tile tl = tile3(pd.g, tr.tile);
        if (stop) {
            // finish iteration n-1 only (e -= alpha p, 9_grad.f90:458-468)
            if (do_e && tl.i < pd.g.nx && tl.j < pd.g.ny) {
                const long long plane = (long long)pd.g.nx * pd.g.ny;
                long long c = (long long)tl.k0 * plane + (long long)tl.j * pd.g.nx + tl.i;
                for (int k = tl.k0; k < tl.k1; ++k, c += plane) V.e[c] = V.e[c] - a_prev * p_old[c];
            }
            continue;
        }
        const T b = (T)beta;
        const double acc = with_coefs(V, [&](const auto &co) {
            if (first) return cg_a_tile<T, true, false>(pd.g, tl, V.e, p_old, p_new, V.r, co, V.q, b, a_prev);
            if (do_e) return cg_a_tile<T, false, true>(pd.g, tl, V.e, p_old, p_new, V.r, co, V.q, b, a_prev);
            return cg_a_tile<T, false, false>(pd.g, tl, V.e, p_old, p_new, V.r, co, V.q, b, a_prev);
        });
        const double s = block_sum(acc);
        if (threadIdx.x == 0) pd.P.spmv[tr.tile] = s;
    }
}

// One 3-D tile of the one-launch FP32 iteration n (cg_m).  With UPD the
// residual of iteration n-1 is formed where needed,
//   r_{n-1} = r_{n-2} - alpha_{n-1} q_{n-1}            (9_grad.f90:470-480)
// with r_{n-2} = p_{n-1} - beta_{n-1} p_{n-2} (REC: CG's direction update
// run backwards, so that no step stores r) or read from r_old; without UPD
// (iteration 1, or the step after a residual replacement) r_old is r_{n-1}
// itself.  Then as cg_a: p_n = r_{n-1} + beta p_{n-1}, q_n = A p_n, and
// with FOLD, c -= alpha_j p_j for the kFold iterations j = n-kFold .. n-1
// on the tile's own cells.  Writes p_n, q_n; partials p_n.q_n,
// q_n.r_{n-1}, q_n.q_n, r_{n-1}.r_{n-1}.
//
// The fold's operands: p_{n-kFold} is in the ring slot that p_n replaces
// (read through p_new before the store), p_{n-1} is p_old, and the slots in
// between are listed in Fold; alphas oldest first.
struct Fold {
    gptr<const float> p[kFold - 2];   // p_{n-kFold+1} .. p_{n-2}
    float a[kFold];              // alpha_{n-kFold} .. alpha_{n-1}
};

template <bool FIRST, bool UPD, bool REC, bool FOLD, class CO>
__device__ __forceinline__ void cg_m_tile(const Geo &g, const Tile &t, gptr<float> __restrict__ c32,
                                          gptr<const float> __restrict__ r_old, gptr<const float> __restrict__ q_old,
                                          gptr<const float> __restrict__ p_old, gptr<const float> __restrict__ p_prev,
                                          gptr<float> __restrict__ p_new, gptr<float> __restrict__ q_new,
                                          gptr<float> __restrict__ r_w, const CO &co, float a, float b, float bp,
                                          const Fold &fd, double acc[4])
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long plane = (long long)nx * ny;
    if (t.i >= nx || t.j >= ny) return;
    long long c = (long long)t.k0 * plane + (long long)t.j * nx + t.i;
    const int iw = nb_lo(t.i, nx, g.px), ie = nb_hi(t.i, nx, g.px);
    const int js = nb_lo(t.j, ny, g.py), jn = nb_hi(t.j, ny, g.py);
    const long long dS = (long long)(js - t.j) * nx, dN = (long long)(jn - t.j) * nx;
    long long fx = ((long long)t.k0 * ny + t.j) * (nx + 1) + t.i;
    long long fy = ((long long)t.k0 * (ny + 1) + t.j) * nx + t.i;
    // r_{n-2} at x from p_{n-1} (po) and p_{n-2} (pp) there, or as stored
    auto rn = [&](long long x, float po, float pp) -> float {
        const float r = REC ? po - bp * pp : r_old[x];
        return UPD ? r - a * q_old[x] : r;
    };
    auto pn = [&](long long x) -> float {
        const float po = FIRST ? 0.0f : p_old[x];
        const float r = rn(x, po, REC ? p_prev[x] : 0.0f);
        return FIRST ? r : r + b * po;
    };
    float poc = FIRST ? 0.0f : p_old[c];
    float ppc = REC ? p_prev[c] : 0.0f;
    float rc = rn(c, poc, ppc);
    float pc = FIRST ? rc : rc + b * poc;
    float pb = t.k0 > 0 ? pn(c - plane) : pc;
    const unsigned long long rows = co.rows(t.j, t.k0, t.k1);
    float szb = co.z_at(row_bit(rows, t.k0, t.k0), c, t.k0);
    for (int k = t.k0; k < t.k1; ++k) {
        const long long r0 = c - t.i;
        const long long up = (k < nz - 1) ? plane : 0;
        const float pot = FIRST ? 0.0f : p_old[c + up];
        const float ppt = REC ? p_prev[c + up] : 0.0f;
        const float rt = rn(c + up, pot, ppt);
        const float pt = FIRST ? rt : rt + b * pot;
        const float pw = pn(r0 + iw), pe = pn(r0 + ie);
        const float ps = pn(c + dS), pnn = pn(c + dN);
        const bool reg = row_bit(rows, k, t.k0);
        const float szt = co.z_at(reg, c + plane, k + 1);
        const float sxw = co.x_at(reg, fx, k, 0), sxe = co.x_at(reg, fx, k, 1);
        const float sys = co.y_at(reg, fy, k), syn = co.y_at(reg, fy + nx, k);
        const float ptu = (g.dtop && k == nz - 1) ? -pc : pt;   // Dirichlet lid
        const float y = pc - sxe * pe - sxw * pw - syn * pnn - sys * ps - szt * ptu - szb * pb;
        // p_{n-kFold}, in the slot p_n is about to replace (this lane's cell)
        const float p_first = FOLD ? p_new[c] : 0.0f;
        q_new[c] = y;
        p_new[c] = pc;
        if (r_w) r_w[c] = rc;
        if (FOLD) {
            // 9_grad.f90:458-468 for kFold iterations (p_{n-2} is at hand
            // when r was recovered from it)
            float sum = fd.a[0] * p_first;
#pragma unroll
            for (int m = 0; m < kFold - 2; ++m)
                sum += fd.a[m + 1] * ((REC && m == kFold - 3) ? ppc : fd.p[m][c]);
            sum += fd.a[kFold - 1] * poc;
            c32[c] = c32[c] - sum;
        }
        acc[0] += (double)pc * (double)y;
        acc[1] += (double)y * (double)rc;
        acc[2] += (double)y * (double)y;
        acc[3] += (double)rc * (double)rc;
        pb = pc;
        pc = pt;
        rc = rt;
        poc = pot;
        ppc = ppt;
        szb = szt;
        c += plane;
        fx += (long long)ny * (nx + 1);
        fy += (long long)(ny + 1) * nx;
    }
}

// V x cells per lane (FP32, V = 2 or 4, nx a multiple of V): the lane's
// cells i .. i+V-1 are one 8- or 16-byte vector in every cell-layout array
// (all 256-byte aligned, i a multiple of V).
template <int V>
__device__ __forceinline__ void ldv(gptr<const float> p, long long x, float (&o)[V])
{
    typedef float vt __attribute__((ext_vector_type(V)));
    const vt t = *reinterpret_cast<gptr<const vt>>(p + x);
#pragma unroll
    for (int m = 0; m < V; ++m) o[m] = t[m];
}
template <int V>
__device__ __forceinline__ void stv(gptr<float> p, long long x, const float (&v)[V])
{
    typedef float vt __attribute__((ext_vector_type(V)));
    vt t;
#pragma unroll
    for (int m = 0; m < V; ++m) t[m] = v[m];
    *reinterpret_cast<gptr<vt>>(p + x) = t;
}
// V consecutive floats at a 4-byte-aligned address (x faces: a row holds
// nx + 1 of them) as one vector load.
template <int V>
__device__ __forceinline__ void ldvu(gptr<const float> p, long long x, float (&o)[V])
{
    typedef float vt __attribute__((ext_vector_type(V), aligned(4)));
    const vt t = *reinterpret_cast<gptr<const vt>>(p + x);
#pragma unroll
    for (int m = 0; m < V; ++m) o[m] = t[m];
}

// Wavefront-wide lane shifts by one (DPP wave_shr / wave_shl): lane l gets
// lane l-1's (from_lower) or lane l+1's (from_upper) value; lane 0, resp.
// lane 63, which has no such neighbour, keeps `edge`.  Every lane of the
// wave executes them.
__device__ __forceinline__ float from_lower(float edge, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_upper(float edge, float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_value(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Plane tables: the V + 1 x faces of a lane's V cells (faces i .. i+V), the
// V y or z faces at one offset, all the plane's value.
template <int V>
__device__ __forceinline__ void xv(const PlaneTables<float> &co, long long, int k, float (&o)[V + 1])
{
    const float v = co.tx[k];
#pragma unroll
    for (int m = 0; m <= V; ++m) o[m] = v;
}
template <int V>
__device__ __forceinline__ void yv(const PlaneTables<float> &co, long long, int k, float (&o)[V])
{
    const float v = co.ty[k];
#pragma unroll
    for (int m = 0; m < V; ++m) o[m] = v;
}
template <int V>
__device__ __forceinline__ void zv(const PlaneTables<float> &co, long long, int kf, float (&o)[V])
{
    const float v = co.tz[kf];
#pragma unroll
    for (int m = 0; m < V; ++m) o[m] = v;
}

// The coefficients of one plane k of a row: x faces (V + 1), y faces j and
// j + 1, and z face kf (k + 1 or k, at offset fz): the row tables when the
// row's bit is set, else the face arrays.
template <int V>
__device__ __forceinline__ void coefs_v(const FaceArrays<float> &co, bool reg, long long fx, long long fy,
                                        long long fz, int kf, int nx, int k, float (&sxv)[V + 1], float (&sys)[V],
                                        float (&syn)[V], float (&sz)[V])
{
    const gptr<const float> px = co.xp(reg, fx, k);
    float t[V];
    ldvu<V>(px, 0, t);
#pragma unroll
    for (int m = 0; m < V; ++m) sxv[m] = t[m];
    sxv[V] = px[V];
    ldv<V>(co.yp(reg, fy, k), 0, sys);
    ldv<V>(co.yp(reg, fy + nx, k), 0, syn);
    ldv<V>(co.zp(reg, fz, kf), 0, sz);
}
template <int V>
__device__ __forceinline__ void coefs_v(const PlaneTables<float> &co, bool, long long fx, long long fy,
                                        long long fz, int kf, int nx, int k, float (&sxv)[V + 1], float (&sys)[V],
                                        float (&syn)[V], float (&sz)[V])
{
    xv<V>(co, fx, k, sxv);
    yv<V>(co, fy, k, sys);
    yv<V>(co, fy + nx, k, syn);
    zv<V>(co, fz, kf, sz);
}
// z faces of a plane's row at the tile's first plane
template <int V>
__device__ __forceinline__ void zb_v(const FaceArrays<float> &co, bool reg, long long f, int kf, float (&o)[V])
{
    ldv<V>(co.zp(reg, f, kf), 0, o);
}
template <int V>
__device__ __forceinline__ void zb_v(const PlaneTables<float> &co, bool, long long f, int kf, float (&o)[V])
{
    zv<V>(co, f, kf, o);
}
// Where a wave of V-cell lanes sits in its row: the wave covers the
// segment of 64 V cells from x = i0 (one segment when nx <= 64 V).  Lanes
// past the row's end (nx is a multiple of V) read the row's last V cells
// instead, store nothing and add nothing.  The x neighbours of a lane's
// cells come from the adjacent lanes; beyond the segment's ends they are the
// mirror ghost (the end cell itself), the periodic partner (another lane of
// a one-segment row), or a cell of the next segment (loaded).
struct RowWave {
    int lane, last;      // this lane; the segment's last lane inside the row
    bool active;         // the lane's cells are inside the row
    int ci;              // the lane's first cell (clamped for inactive lanes)
    bool load_w, load_e; // the cell beyond the segment's west / east end is loaded
    int xw, xe;          // ... at x = xw / xe of the same row
    bool wrap;           // periodic one-segment row: the partner is lane `last` / lane 0
};

template <int V>
__device__ __forceinline__ RowWave row_wave(const Geo &g, const Tile &t)
{
    RowWave w;
    const int nx = g.nx;
    w.lane = threadIdx.x & (kTileX - 1);
    const int i0 = t.i0;
    const int seg = kTileX * V;
    w.last = min(kTileX - 1, (nx - i0) / V - 1);
    w.active = t.i < nx;
    w.ci = w.active ? t.i : nx - V;
    const bool one = i0 == 0 && nx <= seg;
    w.load_w = i0 > 0 || (g.px && !one);
    w.load_e = i0 + seg < nx || (g.px && !one);
    w.xw = i0 > 0 ? i0 - 1 : nx - 1;
    w.xe = i0 + seg < nx ? i0 + seg : 0;
    w.wrap = g.px && one;
    return w;
}

// cg_m_tile with V cells per lane in row waves: the same arithmetic per cell
// (and the same term order).  The x neighbours of p_n are exchanged between
// lanes (DPP) instead of being formed again from three loads each, and the
// x faces of a lane come as one vector.
//
// PIPE: the walk is software-pipelined: the loads of plane k + 1
// (PlaneLoads) are issued before plane k is computed, so that a wave has two
// planes' memory round trips in flight instead of one (the last plane issues
// its own loads once more, in bounds and unused, so that the loop has no
// branch on it).  The two load sets take about 60 more registers at four
// cells per lane, which leaves two waves per SIMD instead of three; without
// PIPE each plane's loads are issued just before it is computed.
//
// EDGE: the row has more than one segment, so the cells beyond a segment's
// ends may have to be loaded; without it (nx <= 64 V) the loop has no
// branch on them.
static_assert(kFold == 4, "PlaneLoads holds the fold's three older directions");
template <int V>
struct PlaneLoads {
    // p_{n-1}, then p_{n-2} (REC) or r_{n-2}, then q_{n-1} (UPD): the plane
    // above (t), the south (s) and north (n) rows
    float to[V], tb[V], tq[V], so[V], sb[V], sq[V], no[V], nb[V], nq[V];
    float sxv[V + 1], sys[V], syn[V];   // x and y faces
    float szt[V], szb[V];                // z faces k + 1 and k (szb: PIPE only)
    float wo, wb, wq, eo, eb, eq;        // cells beyond the segment's ends (EDGE)
    float f0[V], f1[V], f2[V], cv[V];    // FOLD: p_{n-4}, p_{n-3}, p_{n-2} (unless REC) at the cell; c
};

// v copied into a register of its own here.  A value that a plane's loads
// brought and a later plane still needs is copied out before the next loads
// into the same set are issued: left in place, it would keep those loads
// from reusing its register, and the loop would then have to copy the set
// at its back edge, waiting for the loads in flight.
__device__ __forceinline__ float copy_now(float v)
{
    float r;
    asm volatile("v_mov_b32 %0, %1" : "=v"(r) : "v"(v) : "memory");
    return r;
}

template <int V, bool PIPE, bool FIRST, bool UPD, bool REC, bool FOLD, bool EDGE, class CO>
__device__ __forceinline__ void cg_m_tilev(const Geo &g, const Tile &t, gptr<float> __restrict__ c32,
                                           gptr<const float> __restrict__ r_old, gptr<const float> __restrict__ q_old,
                                           gptr<const float> __restrict__ p_old, gptr<const float> __restrict__ p_prev,
                                           gptr<float> __restrict__ p_new, gptr<float> __restrict__ q_new,
                                           gptr<float> __restrict__ r_w, const CO &co, float a, float b, float bp,
                                           const Fold &fd, double acc[4])
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long plane = (long long)nx * ny;
    if (t.j >= ny) return;   // the whole wave
    const RowWave w = row_wave<V>(g, t);
    const long long dfx = (long long)ny * (nx + 1), dfy = (long long)(ny + 1) * nx;   // one plane up
    long long c = (long long)t.k0 * plane + (long long)t.j * nx + w.ci;
    const int js = nb_lo(t.j, ny, g.py), jn = nb_hi(t.j, ny, g.py);
    const long long dS = (long long)(js - t.j) * nx, dN = (long long)(jn - t.j) * nx;
    long long fx = ((long long)t.k0 * ny + t.j) * (nx + 1) + w.ci;
    long long fy = ((long long)t.k0 * (ny + 1) + t.j) * nx + w.ci;
    const gptr<const float> rb = REC ? p_prev : r_old;   // the second operand of r_{n-2}
    // the operands of p_n at x (V cells, or one)
    auto raw = [&](long long x, float (&o)[V], float (&rr)[V], float (&q)[V]) {
        if (!FIRST) ldv<V>(p_old, x, o);
        ldv<V>(rb, x, rr);
        if (UPD) ldv<V>(q_old, x, q);
    };
    auto raw1 = [&](long long x, float &o, float &rr, float &q) {
        if (!FIRST) o = p_old[x];
        rr = rb[x];
        if (UPD) q = q_old[x];
    };
    // r_{n-1} = r_{n-2} - alpha q_{n-1} with r_{n-2} = p_{n-1} - beta_{n-1}
    // p_{n-2} (REC) or as stored; p_n = r_{n-1} + beta_n p_{n-1}
    auto r_of = [&](float o, float rr, float q) -> float {
        float r = REC ? o - bp * rr : rr;
        if (UPD) r = r - a * q;
        return r;
    };
    auto p_of = [&](float o, float r) -> float { return FIRST ? r : r + b * o; };
    const unsigned long long rows = co.rows(t.j, t.k0, t.k1);
    // the fold's operands at cc (plane by plane: loaded where they are used)
    auto fold_loads = [&](long long cc, PlaneLoads<V> &L) {
        ldv<V>(p_new, cc, L.f0);   // p_{n-kFold}, before p_n replaces it
        ldv<V>(fd.p[0], cc, L.f1);
        if (!REC) ldv<V>(fd.p[1], cc, L.f2);
        ldv<V>(c32, cc, L.cv);
    };
    auto issue = [&](int k, long long cc, long long fxx, long long fyy, PlaneLoads<V> &L) {
        const long long up = k < nz - 1 ? plane : 0;   // the top neighbour (mirror ghost at the top)
        raw(cc + up, L.to, L.tb, L.tq);
        raw(cc + dS, L.so, L.sb, L.sq);
        raw(cc + dN, L.no, L.nb, L.nq);
        // z faces k + 1 and, pipelined, k (a row's bit covers both).  Plane
        // by plane, face k stays from the plane below; pipelined, a value
        // kept from one plane's loads for the next would hold a register
        // the next loads need, so it is loaded again (an L1 hit)
        const bool reg = row_bit(rows, k, t.k0);
        coefs_v<V>(co, reg, fxx, fyy, cc + plane, k + 1, nx, k, L.sxv, L.sys, L.syn, L.szt);
        if (PIPE) zb_v<V>(co, reg, cc, k, L.szb);
        if (EDGE && PIPE) {
            // loaded whether used or not (the row's first cell then, the
            // same address in every lane), so that the loop has no branch
            // around loads in flight; plane by plane they are loaded where
            // they are used
            const long long r0 = cc - w.ci;
            raw1(w.load_w ? r0 + w.xw : r0, L.wo, L.wb, L.wq);
            raw1(w.load_e ? r0 + w.xe : r0, L.eo, L.eb, L.eq);
        }
        if (FOLD && PIPE) fold_loads(cc, L);
    };
    // r_{n-1} and p_n at the plane at hand (c), p_n at the one below (pb),
    // the z faces below (szb); for the fold, p_{n-1} and p_{n-2} (REC) at c
    // (copy_now)
    float rc[V], pc[V], pb[V], szb[V], poc[V], ppc[V];
    {
        float o[V], rr[V], q[V];
        raw(c, o, rr, q);
#pragma unroll
        for (int m = 0; m < V; ++m) {
            rc[m] = r_of(o[m], rr[m], q[m]);
            pc[m] = p_of(o[m], rc[m]);
            if (FOLD) poc[m] = PIPE ? copy_now(o[m]) : o[m];
            if (FOLD && REC) ppc[m] = PIPE ? copy_now(rr[m]) : rr[m];
        }
    }
    if (t.k0 > 0) {
        float o[V], rr[V], q[V];
        raw(c - plane, o, rr, q);
#pragma unroll
        for (int m = 0; m < V; ++m) pb[m] = p_of(o[m], r_of(o[m], rr[m], q[m]));
    } else {
#pragma unroll
        for (int m = 0; m < V; ++m) pb[m] = pc[m];   // mirror ghost
    }
    if (!PIPE) zb_v<V>(co, row_bit(rows, t.k0, t.k0), c, t.k0, szb);
    // one plane: compute it from `cur` (PIPE: while the next plane's loads
    // go into `nxt`, two sets used in turn, so that no copy waits for loads
    // in flight)
    auto plane_step = [&](int k, PlaneLoads<V> &cur, PlaneLoads<V> &nxt) {
        const bool more = k + 1 < t.k1;
        if (PIPE)
            issue(more ? k + 1 : k, more ? c + plane : c, more ? fx + dfx : fx, more ? fy + dfy : fy, nxt);
        else
            issue(k, c, fx, fy, cur);
        float rt[V], pt[V], ps[V], pnn[V];
#pragma unroll
        for (int m = 0; m < V; ++m) {
            rt[m] = r_of(cur.to[m], cur.tb[m], cur.tq[m]);
            pt[m] = p_of(cur.to[m], rt[m]);
            ps[m] = p_of(cur.so[m], r_of(cur.so[m], cur.sb[m], cur.sq[m]));
            pnn[m] = p_of(cur.no[m], r_of(cur.no[m], cur.nb[m], cur.nq[m]));
        }
        // x neighbours of the lane's end cells: adjacent lanes, or beyond the segment
        const float ew_row = w.wrap ? lane_value(pc[V - 1], w.last) : pc[0];
        const float ee_row = w.wrap ? lane_value(pc[0], 0) : pc[V - 1];
        // p_n at a cell beyond the segment's ends
        auto edge = [&](float o, float rr, float q) { return p_of(o, r_of(o, rr, q)); };
        auto edge_at = [&](long long x) {
            float o = 0.0f, rr, q = 0.0f;
            raw1(x, o, rr, q);
            return edge(o, rr, q);
        };
        const long long r0 = c - w.ci;
        float ew = ew_row, ee = ee_row;
        if (EDGE && w.load_w) ew = PIPE ? edge(cur.wo, cur.wb, cur.wq) : edge_at(r0 + w.xw);
        if (EDGE && w.load_e) ee = PIPE ? edge(cur.eo, cur.eb, cur.eq) : edge_at(r0 + w.xe);
        const float pw = from_lower(ew, pc[V - 1]);
        const float pe_dpp = from_upper(ee, pc[0]);
        const float pe = w.lane == w.last ? ee : pe_dpp;
        const bool lid = g.dtop && k == nz - 1;   // Dirichlet lid
        float y[V];
#pragma unroll
        for (int m = 0; m < V; ++m) {
            const float east = m == V - 1 ? pe : pc[m + 1];
            const float west = m == 0 ? pw : pc[m - 1];
            const float top = lid ? -pc[m] : pt[m];
            y[m] = pc[m] - cur.sxv[m + 1] * east - cur.sxv[m] * west - cur.syn[m] * pnn[m] - cur.sys[m] * ps[m] -
                   cur.szt[m] * top - (PIPE ? cur.szb[m] : szb[m]) * pb[m];
        }
        if (FOLD && !PIPE) fold_loads(c, cur);   // before p_n replaces p_{n-kFold}
        if (w.active) {
            stv<V>(q_new, c, y);
            stv<V>(p_new, c, pc);
            if (r_w) stv<V>(r_w, c, rc);
            if (FOLD) {
                // 9_grad.f90:458-468 for kFold iterations
                float cv[V];
#pragma unroll
                for (int m = 0; m < V; ++m) {
                    float sum = fd.a[0] * cur.f0[m];
                    sum += fd.a[1] * cur.f1[m];
                    sum += fd.a[2] * (REC ? ppc[m] : cur.f2[m]);
                    cv[m] = cur.cv[m] - (sum + fd.a[kFold - 1] * poc[m]);
                }
                stv<V>(c32, c, cv);
            }
#pragma unroll
            for (int m = 0; m < V; ++m) acc[0] += (double)pc[m] * (double)y[m];
#pragma unroll
            for (int m = 0; m < V; ++m) acc[1] += (double)y[m] * (double)rc[m];
#pragma unroll
            for (int m = 0; m < V; ++m) acc[2] += (double)y[m] * (double)y[m];
#pragma unroll
            for (int m = 0; m < V; ++m) acc[3] += (double)rc[m] * (double)rc[m];
        }
#pragma unroll
        for (int m = 0; m < V; ++m) {
            pb[m] = pc[m];
            pc[m] = pt[m];
            rc[m] = rt[m];
            if (!PIPE) szb[m] = cur.szt[m];
            if (FOLD) poc[m] = PIPE ? copy_now(cur.to[m]) : cur.to[m];
            if (FOLD && REC) ppc[m] = PIPE ? copy_now(cur.tb[m]) : cur.tb[m];
        }
        c += plane;
        fx += dfx;
        fy += dfy;
    };
    // (both exits leave the loop, so that the back edge carries one version
    // of each set: a merge of two would copy registers whose loads are in
    // flight, and wait for them)
    PlaneLoads<V> L0, L1;
    if (PIPE) {
        issue(t.k0, c, fx, fy, L0);
        for (int k = t.k0;; k += 2) {
            plane_step(k, L0, L1);
            if (k + 1 >= t.k1) break;
            plane_step(k + 1, L1, L0);
            if (k + 2 >= t.k1) break;
        }
    } else {
        for (int k = t.k0; k < t.k1; ++k) plane_step(k, L0, L0);
    }
}

__device__ __forceinline__ gptr<float> qbuf(const ProbDev &pd, int parity) { return glob(parity ? pd.q32b : pd.q32); }
// ring slot of p_k
__device__ __forceinline__ gptr<float> pbuf(const ProbDev &pd, int k)
{
    return glob(pd.p32 + (k & (kFold - 1)) * pd.pslot);
}
__device__ __forceinline__ float alph_of(gptr<const CgScalars> S, int k) { return (float)S->alph[k & (kFold - 1)]; }
// stored residual r_k (store_r): r32 for even k (where the setup leaves r_0
// and a replacement, always after an even step, its residual), r32b for odd
__device__ __forceinline__ gptr<float> rbuf(const ProbDev &pd, int k) { return glob((k & 1) ? pd.r32b : pd.r32); }

// One tile of the mixed step n with V cells per lane (V = 1: cg_m_tile).
// REC: r_{n-2} is recovered from the directions, r_{n-2} = p_{n-1} -
// beta_{n-1} p_{n-2} (bp = beta_{n-1}, the FP32 factor that formed p_{n-1}),
// so that no step stores r; otherwise (the first step, and the two steps
// after a residual replacement) it is read from r32, where the setup and the
// replacement leave it.
// FOLD: c takes the updates of iterations n-kFold .. n-1 this step (their
// alphas but alpha_{n-1} = a were stored by the previous steps).  The fold's
// operands are gathered here, inside the FOLD instance only, so that the
// other instances carry no extra live registers.
template <int V, bool PIPE, bool FIRST, bool UPD, bool REC, bool FOLD, class CO>
__device__ __forceinline__ void cg_m_tile_any(const ProbDev &pd, int n, const Tile &tl, gptr<const float> r_o,
                                              gptr<const float> q_o, gptr<const float> p_o, gptr<const float> p_oo,
                                              gptr<float> p_w, gptr<float> q_w, gptr<float> r_w, const CO &co,
                                              float a, float b, float bp, double acc[4])
{
    Fold fd;
    if constexpr (FOLD) {
#pragma unroll
        for (int m = 0; m < kFold - 1; ++m) fd.a[m] = alph_of(glob(pd.S), n - kFold + m);
        fd.a[kFold - 1] = a;
#pragma unroll
        for (int m = 0; m < kFold - 2; ++m) fd.p[m] = pbuf(pd, n - kFold + 1 + m);
    }
#define CG_M_ARGS pd.g, tl, glob(pd.c32), r_o, q_o, p_o, p_oo, p_w, q_w, r_w, co, a, b, bp, fd, acc
    if constexpr (V == 1)
        cg_m_tile<FIRST, UPD, REC, FOLD>(CG_M_ARGS);
    else if (pd.g.nx > kTileX * V)
        cg_m_tilev<V, PIPE, FIRST, UPD, REC, FOLD, true>(CG_M_ARGS);
    else
        cg_m_tilev<V, PIPE, FIRST, UPD, REC, FOLD, false>(CG_M_ARGS);
#undef CG_M_ARGS
}

// FOLD: this launch is one of the steps at which c takes a group (the
// instances of the other steps are not compiled into it).
template <int V, bool PIPE, bool FOLD, class CO>
__device__ __forceinline__ void cg_m_step(bool first, bool upd, bool rec, const ProbDev &pd, int n, const Tile &tl,
                                          gptr<const float> r_o, gptr<const float> q_o, gptr<const float> p_o,
                                          gptr<const float> p_oo, gptr<float> p_w, gptr<float> q_w, gptr<float> r_w,
                                          const CO &co, float a, float b, float bp, double acc[4])
{
#define CG_M(F, U, R, D) cg_m_tile_any<V, PIPE, F, U, R, D>(pd, n, tl, r_o, q_o, p_o, p_oo, p_w, q_w, r_w, co, a, b, bp, acc)
    if (first)
        CG_M(true, false, false, false);
    else if (!upd)
        CG_M(false, false, false, false);
    else if (rec)
        CG_M(false, true, true, FOLD);
    else
        CG_M(false, true, false, FOLD);
#undef CG_M
}

// The mixed solver's FP32 iteration, one launch per global step (register
// budget held to four waves per SIMD).  Scalars
// from the totals of the previous step's partials (slot (n-1) & 1):
//   alpha_{n-1} = rho_{n-2} / p_{n-1}.q_{n-1}               (9_grad.f90:450-455)
//   rho_{n-1}   = rho_{n-2} - 2 alpha q.r + alpha^2 q.q      (r.r of r_{n-1})
//   beta_n      = rho_{n-1} / rho_{n-2}                      (:510-515)
// rho_{n-2} is summed directly by the previous step, so the identity only
// spans one iteration.  After a residual replacement rho_{n-1} is the
// replaced residual's r.r and nothing of iteration n-1 is left to finish.
// The stop test is the FP64 one at replacements; here only the cap.
//
// FOLD: the step folds a group into c (n = 1 mod kFold).  Problems join on
// the kRefreshGrain grain (n0 - 1 is a multiple of it, and of kFold), so n
// and the global step agree mod kFold: every problem of a launch folds or
// none does, and the host launches the instance that does.
//
// VX, PIPE: cells per lane of every problem in the table and whether its
// walk is pipelined (the engine launches one instance per kind).  Waves per
// SIMD that the registers allow: four (one cell per lane, or two without
// PIPE), three (four cells, or two with PIPE), two (four cells with PIPE).
constexpr int cg_m_waves(int vx, bool pipe) { return vx == 1 ? 4 : vx == 2 ? (pipe ? 3 : 4) : (pipe ? 2 : 3); }
template <int VX, bool PIPE, bool FOLD>
__global__ __launch_bounds__(kThreads, cg_m_waves(VX, PIPE))
void k_cg_m(const ProbDev *__restrict__ P, const TileRef *__restrict__ tab, int ntab, StepRef sref, int *ran)
{
    const Run run = my_run(ntab);
    const int step = step_of(sref);
    int cached = -1;
    double T[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = run.begin; t < run.end; ++t) {
        const TileRef tr = tab[t];
        const ProbDev &pd = P[tr.prob];
        const gptr<CgScalars> S = glob(pd.S);
        const int n = step - pd.n0 + 1;
        // The problem's scalars and the previous step's partials were both
        // written by the previous step.  They are requested together, and
        // the partials are summed before the scalars decide whether the
        // problem runs at all, so that a step starts after one memory round
        // trip instead of two (a finished problem's tiles sum them for
        // nothing).
        // The partials' slot is the parity of n - 1, which is that of
        // sref.off - 1: n0 - 1 is a multiple of kRefreshGrain (problems are
        // admitted on the grain) and a replayed group's base step is one too,
        // so the slot is known without waiting for the replayed step counter.
        const int done_at = S->done_at, skip_e_at = S->skip_e_at, max_iter = S->max_iter;
        const double rr1 = S->rr[1], rr_after = S->rr_after, beta_after = S->beta_after;
        const double beta_prev = S->bet[(n - 1) & 1];
        if (tr.prob != cached) {
            total4_of(pd.P.m4 + (size_t)((sref.off - 1) & 1) * 4 * pd.nt3, pd.nt3, T);
            cached = tr.prob;
        }
        if (n < 1 || n > done_at) continue;
        const bool first = n == 1;
        const bool upd = !first && step != skip_e_at;
        double alpha = 0.0, rho, beta = 0.0;
        if (first) {
            rho = rr1;
        } else if (upd) {
            alpha = T[3] / T[0];
            rho = T[3] - 2.0 * alpha * T[1] + alpha * alpha * T[2];
            beta = rho / T[3];
        } else {
            rho = rr_after;
            beta = beta_after;
        }
        const bool cap = n - 1 >= max_iter;
        // c takes the updates of iterations n-kFold .. n-1 together at
        // n = 1 mod kFold (replacements fall after multiples of kFold, so a
        // group never spans one); their alphas but the last were stored by
        // the previous steps
        const bool fold = FOLD && upd;
        // p_{n-1} came from r_{n-2} and beta_{n-1} p_{n-2} unless step n-1
        // was the first or followed a replacement
        // (with stored residuals r_{n-2} is read instead)
        const bool store = pd.store_r != 0;
        const bool rec = !store && upd && n != 2 && step != skip_e_at + 1;
        const float a = (float)alpha;
        if (tr.tile == 0 && threadIdx.x == 0) {
            S->rr[n & 1] = rho;
            S->bet[n & 1] = beta;
            S->iters = n - 1;
            S->alpha = alpha;
            if (upd) S->alph[(n - 1) & (kFold - 1)] = alpha;
            if (cap) S->done_at = n;
            if (ran && tr.prob == 0 && !cap) *ran = 1;
        }
        const int o = (n - 1) & 1, w = n & 1;
        const Tile tl = tile3(pd.g, tr.tile);
        if (cap) {
            // finish what c has not taken yet, iterations lo .. n-1 since the
            // last fold (the replacement that follows recomputes the residual)
            const int lo = (n - 2) / kFold * kFold + 1;
            const gptr<const float> p_old = pbuf(pd, n - 1);
            const gptr<float> c32 = glob(pd.c32);
            for (int sub = 0; upd && sub < pd.g.vx; ++sub) {
                const int ci = tl.i + sub;
                if (ci >= pd.g.nx || tl.j >= pd.g.ny) break;
                const long long plane = (long long)pd.g.nx * pd.g.ny;
                long long c = (long long)tl.k0 * plane + (long long)tl.j * pd.g.nx + ci;
                for (int k = tl.k0; k < tl.k1; ++k, c += plane) {
                    float sum = 0.0f;
                    for (int j = lo; j < n - 1; ++j) sum += alph_of(S, j) * pbuf(pd, j)[c];
                    c32[c] = c32[c] - (sum + a * p_old[c]);
                }
            }
            continue;
        }
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        const float b = (float)beta;
        const float bp = (float)beta_prev;
        const gptr<const float> q_o = qbuf(pd, o), p_o = pbuf(pd, n - 1), p_oo = pbuf(pd, n - 2);
        const gptr<float> p_w = pbuf(pd, n), q_w = qbuf(pd, w);
        // the residual read: r32, where the setup and the replacements leave
        // r_{n-1} for a step without UPD (and r_{n-2} for the step after it);
        // with stored residuals r_k of buffer k & 1, this step storing
        // r_{n-1} into the other one
        const gptr<const float> r_o = store ? rbuf(pd, upd ? n - 2 : n - 1) : glob((const float *)pd.r32);
        const gptr<float> r_w = store && upd ? rbuf(pd, n - 1) : nullptr;
        (void)fold;
        with_coefs(vecs(pd, 0.0f), [&](const auto &co) {
            cg_m_step<VX, PIPE, FOLD>(first, upd, rec, pd, n, tl, r_o, q_o, p_o, p_oo, p_w, q_w, r_w, co, a, b, bp,
                                      acc);
            return 0;
        });
        block_sum4(acc);
        if (threadIdx.x == 0) {
            double *m = pd.P.m4 + (size_t)w * 4 * pd.nt3 + 4 * (size_t)tr.tile;
            m[0] = acc[0];
            m[1] = acc[1];
            m[2] = acc[2];
            m[3] = acc[3];
        }
    }
}

// cg_b: alpha = r.r / p.Ap (9_grad.f90:450-455); r -= alpha q (:470-480);
// r.r partials (:482-508).  1-D tile k of a problem with nt1 tiles handles
// the 16-byte vectors k*256 + lane + m*nt1*256.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_cg_b(const ProbDev *__restrict__ P, const TileRef *__restrict__ tab,
                                                   int ntab, StepRef sref)
{
    const Run run = my_run(ntab);
    const int step = step_of(sref);
    TotalCache pap_tot;
    for (int t = run.begin; t < run.end; ++t) {
        const TileRef tr = tab[t];
        const ProbDev &pd = P[tr.prob];
        const gptr<CgScalars> S = glob(pd.S);
        const int n = step - pd.n0 + 1;
        // scalars and partials requested together (as in cg_a)
        const int done_at = S->done_at;
        const double rr_n = S->rr[sref.off & 1];   // n & 1 (see cg_m)
        const double pap = pap_tot.get(tr.prob, pd.P.spmv, pd.nt3);
        if (n < 1 || n >= done_at) continue;
        const double alpha = rr_n / pap;
        if (tr.tile == 0 && threadIdx.x == 0) {
            S->pap = pap;
            S->alpha = alpha;
        }
        const Vecs<T> V = vecs(pd, T());
        const gptr<T> r = V.r;
        const gptr<const T> q = V.q;
        const T a = (T)alpha;
        constexpr int W = 16 / sizeof(T);
        const long long len = pd.g.n;
        const long long nv = len / W;
        const long long stride = (long long)pd.nt1 * kThreads;
        double acc = 0.0;
        for (long long v = (long long)tr.tile * kThreads + threadIdx.x; v < nv; v += stride) {
            Vec<T> rv = reinterpret_cast<gptr<Vec<T>>>(r)[v];
            const Vec<T> qv = reinterpret_cast<gptr<const Vec<T>>>(q)[v];
#pragma unroll
            for (int l = 0; l < W; ++l) {
                rv[l] = rv[l] - a * qv[l];
                acc += (double)rv[l] * (double)rv[l];
            }
            reinterpret_cast<gptr<Vec<T>>>(r)[v] = rv;
        }
        if (tr.tile == 0) {
            const long long c = nv * W + threadIdx.x;
            if (c < len) {
                const T rv = r[c] - a * q[c];
                r[c] = rv;
                acc += (double)rv * (double)rv;
            }
        }
        const double s = block_sum(acc);
        if (threadIdx.x == 0) pd.P.upd[tr.tile] = s;
    }
}

// Mixed precision: does the residual replacement after global step `step`
// concern this problem?  Its scheduled iteration, or a cap met since the
// last one (the stop path of cg_a left the correction to fold).
__device__ __forceinline__ bool refresh_due(const ProbDev &pd, int step)
{
    const gptr<const CgScalars> S = glob((const CgScalars *)pd.S);
    const int n = step - pd.n0 + 1;
    return n >= 1 && !S->final_ && (n == S->next_refresh || S->done_at != kNotDone);
}

// Mixed precision, residual replacement step 1: fold the FP32 correction
// into the FP64 iterate, e += c - sum of alpha_j p_j over j = n-kFold+1 .. n
// (the updates of the last kFold iterations, which c takes as a group at the
// step after the group; alpha_n from step n's partials: r_{n-1}.r_{n-1} /
// p_n.q_n), c = 0.
__global__ __launch_bounds__(kThreads) void k_refresh_e(const ProbDev *__restrict__ P,
                                                        const TileRef *__restrict__ tab, int ntab, StepRef sref)
{
    const Run run = my_run(ntab);
    const int step = step_of(sref);
    int cached = -1;
    double T[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = run.begin; t < run.end; ++t) {
        const TileRef tr = tab[t];
        const ProbDev &pd = P[tr.prob];
        const gptr<const CgScalars> S = glob((const CgScalars *)pd.S);
        const int n = step - pd.n0 + 1;
        if (!refresh_due(pd, step)) continue;
        const bool pending = S->done_at == kNotDone;
        if (tr.prob != cached) {
            total4_of(pd.P.m4 + (size_t)(n & 1) * 4 * pd.nt3, pd.nt3, T);
            cached = tr.prob;
        }
        // c has not taken iterations n-kFold+1 .. n yet (n is a multiple of
        // kFold at a replacement): alpha_n from step n's partials, the others
        // stored (p_n was formed by step n; no later step has run)
        float fa[kFold];
        gptr<const float> fp[kFold];
#pragma unroll
        for (int m = 0; m < kFold - 1; ++m) {
            fa[m] = alph_of(S, n - kFold + 1 + m);
            fp[m] = pbuf(pd, n - kFold + 1 + m);
        }
        fa[kFold - 1] = (float)(T[3] / T[0]);
        fp[kFold - 1] = pbuf(pd, n);
        const gptr<double> e = glob(pd.e);
        const gptr<float> c = glob(pd.c32);
        const long long len = pd.g.n;
        const long long nv = len / 4;
        const long long stride = (long long)pd.nt1 * kThreads;
        for (long long v = (long long)tr.tile * kThreads + threadIdx.x; v < nv; v += stride) {
            Vec<float> cv = reinterpret_cast<gptr<const Vec<float>>>(c)[v];
            if (pending) {
                float sum[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int m = 0; m < kFold; ++m) {
                    const Vec<float> pv = reinterpret_cast<gptr<const Vec<float>>>(fp[m])[v];
#pragma unroll
                    for (int l = 0; l < 4; ++l) sum[l] += fa[m] * pv[l];
                }
#pragma unroll
                for (int l = 0; l < 4; ++l) cv[l] = cv[l] - sum[l];
            }
            Vec<double> e0 = reinterpret_cast<gptr<Vec<double>>>(e)[2 * v];
            Vec<double> e1 = reinterpret_cast<gptr<Vec<double>>>(e)[2 * v + 1];
            e0[0] += (double)cv[0];
            e0[1] += (double)cv[1];
            e1[0] += (double)cv[2];
            e1[1] += (double)cv[3];
            reinterpret_cast<gptr<Vec<double>>>(e)[2 * v] = e0;
            reinterpret_cast<gptr<Vec<double>>>(e)[2 * v + 1] = e1;
            reinterpret_cast<gptr<Vec<float>>>(c)[v] = Vec<float>{0.f, 0.f, 0.f, 0.f};
        }
        if (tr.tile == 0) {
            const long long x = nv * 4 + threadIdx.x;
            if (x < len) {
                float cv = c[x];
                if (pending) {
                    float sum = 0.0f;
                    for (int m = 0; m < kFold; ++m) sum += fa[m] * fp[m][x];
                    cv = cv - sum;
                }
                e[x] += (double)cv;
                c[x] = 0.0f;
            }
        }
    }
}

// FP64 true residual r = A e - b (mirrored ghosts) over 3-D tiles.
// Partials (3 slots of nt3): r.r; with WRITE32, r.r of its FP32 copy handed
// to the inner recurrence; with GAUGE, e.sqrt(d).
template <bool WRITE32, bool GAUGE>
__global__ __launch_bounds__(kThreads) void k_true_residual(const ProbDev *__restrict__ P,
                                                            const TileRef *__restrict__ tab, int ntab, StepRef sref)
{
    const Run run = my_run(ntab);
    const int step = step_of(sref);
    for (int tt = run.begin; tt < run.end; ++tt) {
        const TileRef tr = tab[tt];
        const ProbDev &pd = P[tr.prob];
        if (WRITE32 && !refresh_due(pd, step)) continue;
        const Geo &g = pd.g;
        const Tile t = tile3(g, tr.tile);
        const int nx = g.nx, ny = g.ny, nz = g.nz;
        const gptr<const double> __restrict__ e = glob((const double *)pd.e);
        const gptr<const double> __restrict__ b = glob(pd.b);
        const gptr<const double> __restrict__ d = glob(pd.d);
        // the replaced residual goes where the next cg_m reads r_{n}
        const gptr<float> __restrict__ r32 = WRITE32 ? glob(pd.r32) : nullptr;
        double acc = 0.0, acc32 = 0.0, acc_en = 0.0;
        // a tile of vx cells per lane is walked as vx one-cell-per-lane
        // sub-tiles, so that every load is coalesced
        const int lane = threadIdx.x & (kTileX - 1);
        const int i0 = t.i0;
        with_coefs(vecs(pd, 0.0), [&](const auto &co) {
        for (int sub = 0; sub < g.vx; ++sub) {
            const int ci = i0 + sub * kTileX + lane;
            if (ci >= nx || t.j >= ny) continue;
            const long long plane = (long long)nx * ny;
            const int iw = nb_lo(ci, nx, g.px), ie = nb_hi(ci, nx, g.px);
            const int js = nb_lo(t.j, ny, g.py), jn = nb_hi(t.j, ny, g.py);
            const long long dS = (long long)(js - t.j) * nx, dN = (long long)(jn - t.j) * nx;
            long long c = (long long)t.k0 * plane + (long long)t.j * nx + ci;
            long long fx = ((long long)t.k0 * ny + t.j) * (nx + 1) + ci;
            long long fy = ((long long)t.k0 * (ny + 1) + t.j) * nx + ci;
            // e of the cell and of the cell below carried up the column
            // (the bottom ghost mirrors the first plane)
            double ec = e[c];
            double eb = t.k0 > 0 ? e[c - plane] : ec;
            double szb = co.z(c, t.k0);
            const bool dtop = g.dtop != 0;
            for (int k = t.k0; k < t.k1; ++k) {
                const long long r0 = c - ci;
                const long long up = (k < nz - 1) ? plane : 0;
                const double eu = e[c + up];
                const double et = (dtop && k == nz - 1) ? -ec : eu;   // Dirichlet lid
                const double szt = co.z(c + plane, k + 1);
                const double r = ec - co.x(fx + 1, k) * e[r0 + ie] - co.x(fx, k) * e[r0 + iw] -
                                 co.y(fy + nx, k) * e[c + dN] - co.y(fy, k) * e[c + dS] - szt * et -
                                 szb * eb - b[c];
                acc += r * r;
                if (WRITE32) {
                    const float rf = (float)r;
                    r32[c] = rf;
                    acc32 += (double)rf * (double)rf;
                }
                if (GAUGE) acc_en += ec * sqrt(d[c]);
                eb = ec;
                ec = eu;
                szb = szt;
                c += plane;
                fx += (long long)ny * (nx + 1);
                fy += (long long)(ny + 1) * nx;
            }
        }
        return 0;
        });
        const double s = block_sum(acc);
        const double s32 = block_sum(acc32);
        const double sen = block_sum(acc_en);
        if (threadIdx.x == 0) {
            pd.P.tres[tr.tile] = s;
            pd.P.tres[pd.nt3 + tr.tile] = s32;
            pd.P.tres[2 * pd.nt3 + tr.tile] = sen;
        }
    }
}

// Iterations until the next residual replacement, from the problem's own
// history: enough to reach the threshold at the observed rate, but no more
// than lets r.r of the FP32 recurrence fall three decades; a multiple of
// kRefreshGrain.  Measured on the two-phase tank and the constant-density
// case at 256x128x96: three decades took 3 % fewer iterations than six
// (960 -> 928, 1216 -> 1176), chiefly because the last replacement lands
// nearer the tolerance; two to four decades did about as well, while a gap
// capped at 48 or a first replacement at 16 cost iterations.
__device__ __forceinline__ int refresh_gap(double slope, double rr, double thresh)
{
    double k = (double)kFirstRefresh;
    if (slope < 0.0 && rr > 0.0 && thresh > 0.0) {
        const double remaining = log(thresh / rr) / slope;
        const double span = log(1e-3) / slope;
        k = fmin(remaining + 2.0, span);
    }
    k = fmax((double)kRefreshGrain, fmin((double)kMaxRefreshGap, ceil(k)));
    const int ki = (int)k;
    return (ki + kRefreshGrain - 1) / kRefreshGrain * kRefreshGrain;
}

// One workgroup per problem after the residual replacement that follows
// global step `step`: converged (on the FP64 true residual), capped, or
// continue with the replaced residual r (its r.r becomes the "entering" r.r
// of the next cg_a, through cg_b's partial slots; that cg_a skips the
// deferred e update, which the replacement already folded into e).
__global__ __launch_bounds__(kThreads) void k_refresh_decide(const ProbDev *__restrict__ P,
                                                             const int *__restrict__ probs, StepRef sref)
{
    const int step = step_of(sref);
    const ProbDev &pd = P[probs[blockIdx.x]];
    CgScalars *S = pd.S;
    if (!refresh_due(pd, step)) return;
    const int n_last = step - pd.n0 + 1;
    const double rrt = total_of(pd.P.tres, pd.nt3);
    const double rrs = total_of(pd.P.tres + pd.nt3, pd.nt3);
    double T[4];
    total4_of(pd.P.m4 + (size_t)(n_last & 1) * 4 * pd.nt3, pd.nt3, T);
    const bool pending = S->done_at == kNotDone;
    const int iters = pending ? n_last : S->iters;
    const bool conv = !(fabs(rrt) > S->thresh);
    const bool cap = iters >= S->max_iter;
    if (threadIdx.x == 0) {
        S->rr_true = rrt;
        S->refreshes += 1;
        S->iters = iters;
        if (conv || cap) {
            S->final_ = 1;
            S->status = conv ? 0 : 1;
            // later launches skip (cg_a tests done_at)
            if (pending) S->done_at = n_last;
        } else {
            S->skip_e_at = step + 1;
            S->rr_after = rrs;
            // The direction p_n carries on with beta = r'.r' / r_{n-1}.r_{n-1}
            // (the replaced residual in CG's beta).  After a cycle that cut
            // the true r.r by less than 100x, beta = r'.r' / r_n.r_n
            // instead (r_n.r_n of the FP32 recurrence, from the identity):
            // the plain continuation can stall there at ~1e-9 relative.
            const double rho_nm1 = T[3];
            const double alpha_n = T[3] / T[0];
            const double rho_n = T[3] - 2.0 * alpha_n * T[1] + alpha_n * alpha_n * T[2];
            const bool stalled = S->prev_rr > 0.0 && rrt > 1e-2 * S->prev_rr;
            S->beta_after = rrs / (stalled ? rho_n : rho_nm1);
            if (S->prev_iters > 0 && S->prev_rr > 0.0 && rrt < S->prev_rr)
                S->slope = log(rrt / S->prev_rr) / (double)(n_last - S->prev_iters);
            S->prev_iters = n_last;
            S->prev_rr = rrt;
            S->next_refresh = n_last + refresh_gap(S->slope, rrt, S->thresh);
        }
    }
}

// After the solve: record the true r.r and the gauge shift (one workgroup
// per problem).
__global__ __launch_bounds__(kThreads) void k_finish(const ProbDev *__restrict__ P, const int *__restrict__ probs)
{
    const ProbDev &pd = P[probs[blockIdx.x]];
    CgScalars *S = pd.S;
    const double rrt = total_of(pd.P.tres, pd.nt3);
    const double en = total_of(pd.P.tres + 2 * pd.nt3, pd.nt3);
    if (threadIdx.x == 0) {
        S->rr_true = rrt;
        // e's coordinate along sqrt(d) moved by rounding (and by FP32
        // corrections): shift prd1 by the matching constant
        S->gshift = S->project ? (en - S->en0) / S->dd : 0.0;
    }
}

// Plain copy (calibration of HBM counters; a copy's bandwidth).
__global__ __launch_bounds__(kThreads) void k_copy(const float4 *__restrict__ x, float4 *__restrict__ y, long long n)
{
    for (long long v = (long long)blockIdx.x * kThreads + threadIdx.x; v < n; v += (long long)gridDim.x * kThreads)
        y[v] = x[v];
}

// Plain read (a read-only stream's bandwidth): each thread's sum of its
// vectors goes to out[thread], so that the loads are kept.
__global__ __launch_bounds__(kThreads) void k_read(const float4 *__restrict__ x, float *__restrict__ out, long long n)
{
    float s = 0.0f;
    for (long long v = (long long)blockIdx.x * kThreads + threadIdx.x; v < n; v += (long long)gridDim.x * kThreads) {
        const float4 t = x[v];
        s += (t.x + t.y) + (t.z + t.w);
    }
    out[(size_t)blockIdx.x * kThreads + threadIdx.x] = s;
}

__global__ void k_set_step(int *base, int value)
{
    if (threadIdx.x == 0) *base = value;
}

// Plain FP32 SpMV of one tile with V cells per lane in row waves (timing
// only): x neighbours from the adjacent lanes, as in cg_m_tilev.
template <int V, bool EDGE, class CO>
__device__ __forceinline__ double spmv_tilev(const Geo &g, const Tile &t, gptr<const float> __restrict__ p,
                                             gptr<float> __restrict__ q, const CO &co)
{
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    double acc = 0.0;
    if (t.j >= ny) return acc;
    const RowWave w = row_wave<V>(g, t);
    const long long plane = (long long)nx * ny;
    const int js = nb_lo(t.j, ny, g.py), jn = nb_hi(t.j, ny, g.py);
    const long long dS = (long long)(js - t.j) * nx, dN = (long long)(jn - t.j) * nx;
    long long c = (long long)t.k0 * plane + (long long)t.j * nx + w.ci;
    long long fx = ((long long)t.k0 * ny + t.j) * (nx + 1) + w.ci;
    long long fy = ((long long)t.k0 * (ny + 1) + t.j) * nx + w.ci;
    float pc[V], pb[V], szb[V];
    ldv<V>(p, c, pc);
    if (t.k0 > 0)
        ldv<V>(p, c - plane, pb);
    else
#pragma unroll
        for (int m = 0; m < V; ++m) pb[m] = pc[m];
    const unsigned long long rows = co.rows(t.j, t.k0, t.k1);
    zb_v<V>(co, row_bit(rows, t.k0, t.k0), c, t.k0, szb);
    for (int k = t.k0; k < t.k1; ++k) {
        const long long r0 = c - w.ci;
        const long long up = (k < nz - 1) ? plane : 0;
        float pt[V], szt[V], ps[V], pnn[V], sys[V], syn[V], sxv[V + 1], y[V];
        ldv<V>(p, c + up, pt);
        ldv<V>(p, c + dS, ps);
        ldv<V>(p, c + dN, pnn);
        coefs_v<V>(co, row_bit(rows, k, t.k0), fx, fy, c + plane, k + 1, nx, k, sxv, sys, syn, szt);
        const float ew_row = w.wrap ? lane_value(pc[V - 1], w.last) : pc[0];
        const float ee_row = w.wrap ? lane_value(pc[0], 0) : pc[V - 1];
        const float ew = (EDGE && w.load_w) ? p[r0 + w.xw] : ew_row;
        const float ee = (EDGE && w.load_e) ? p[r0 + w.xe] : ee_row;
        const float pw = from_lower(ew, pc[V - 1]);
        const float pe_dpp = from_upper(ee, pc[0]);
        const float pe = w.lane == w.last ? ee : pe_dpp;
        const bool lid = g.dtop && k == nz - 1;
#pragma unroll
        for (int m = 0; m < V; ++m) {
            const float east = m == V - 1 ? pe : pc[m + 1];
            const float west = m == 0 ? pw : pc[m - 1];
            const float top = lid ? -pc[m] : pt[m];
            y[m] = pc[m] - sxv[m + 1] * east - sxv[m] * west - syn[m] * pnn[m] - sys[m] * ps[m] - szt[m] * top -
                   szb[m] * pb[m];
        }
        if (w.active) {
            stv<V>(q, c, y);
#pragma unroll
            for (int m = 0; m < V; ++m) acc += (double)pc[m] * (double)y[m];
        }
#pragma unroll
        for (int m = 0; m < V; ++m) {
            pb[m] = pc[m];
            pc[m] = pt[m];
            szb[m] = szt[m];
        }
        c += plane;
        fx += (long long)ny * (nx + 1);
        fy += (long long)(ny + 1) * nx;
    }
    return acc;
}

// Plain SpMV q = A p of the FP64 or FP32 operator (timing only).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_spmv(const ProbDev *__restrict__ P, const TileRef *__restrict__ tab,
                                                   int ntab)
{
    const Run run = my_run(ntab);
    for (int tt = run.begin; tt < run.end; ++tt) {
        const TileRef tr = tab[tt];
        const ProbDev &pd = P[tr.prob];
        const Geo &g = pd.g;
        const Vecs<T> V = vecs(pd, T());
        const gptr<const T> __restrict__ p = V.p0;
        const gptr<T> __restrict__ q = V.q;
        const Tile t = tile3(g, tr.tile);
        const int nx = g.nx, ny = g.ny, nz = g.nz;
        double acc = 0.0;
        with_coefs(V, [&](const auto &co) {
        if constexpr (sizeof(T) == 4) {
            if (g.vx == 4) {
                acc = g.nx > kTileX * 4 ? spmv_tilev<4, true>(g, t, p, q, co) : spmv_tilev<4, false>(g, t, p, q, co);
                return 0;
            }
            if (g.vx == 2) {
                acc = g.nx > kTileX * 2 ? spmv_tilev<2, true>(g, t, p, q, co) : spmv_tilev<2, false>(g, t, p, q, co);
                return 0;
            }
        }
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
            T szb = co.z(c, t.k0);
            for (int k = t.k0; k < t.k1; ++k) {
                const long long r0 = c - t.i;
                const long long up = (k < nz - 1) ? plane : 0;
                const T pt = p[c + up];
                const T szt = co.z(c + plane, k + 1);
                const T ptu = (g.dtop && k == nz - 1) ? -pc : pt;
                const T y = pc - co.x(fx + 1, k) * p[r0 + ie] - co.x(fx, k) * p[r0 + iw] -
                            co.y(fy + nx, k) * p[c + dN] - co.y(fy, k) * p[c + dS] - szt * ptu - szb * pb;
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
        return 0;
        });
        const double s = block_sum(acc);
        if (threadIdx.x == 0) pd.P.spmv[tr.tile] = s;
    }
}

}  // namespace

Geo make_geo(int nx, int ny, int nz, int px, int py, int kz, int vx)
{
    Geo g{};
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.px = px;
    g.py = py;
    g.n = (long long)nx * ny * nz;
    const int w = vx & ~kPipelined;
    g.vx = ((w == 2 || w == 4) && nx % w == 0) ? w : 1;
    g.pipe = g.vx > 1 && (vx & kPipelined) ? 1 : 0;
    g.tiles_x = (nx + kTileX * g.vx - 1) / (kTileX * g.vx);
    g.tiles_y = (ny + kTileY - 1) / kTileY;
    if (kz <= 0) {
        // about kSpmvBlocks tiles (4 per CU) for a large grid, at least 4
        // planes each
        const long long cols = (long long)g.tiles_x * g.tiles_y;
        kz = (int)((cols * nz + kSpmvBlocks - 1) / kSpmvBlocks);
        if (kz < 4) kz = 4;
    }
    if (kz > nz) kz = nz;
    g.kz = kz;
    g.tiles_z = (nz + kz - 1) / kz;
    return g;
}

int spmv_tiles(const Geo &g) { return g.tiles_x * g.tiles_y * g.tiles_z; }

int vec_tiles(long long n)
{
    long long b = (n / 4 + kThreads - 1) / kThreads;
    return (int)std::max<long long>(1, std::min<long long>(kVecBlocks, b));
}

void launch_cg_a(bool fp32, const ProbDev *P, const TileRef *t3, int nt3, int grid, StepRef step, int *ran,
                 hipStream_t s)
{
    if (fp32)
        hipLaunchKernelGGL(k_cg_a<float>, dim3(grid), dim3(kThreads), 0, s, P, t3, nt3, step, ran);
    else
        hipLaunchKernelGGL(k_cg_a<double>, dim3(grid), dim3(kThreads), 0, s, P, t3, nt3, step, ran);
}

void launch_cg_m(const ProbDev *P, const TileRef *t3, int nt3, int grid, StepRef step, int *ran, int vx,
                 bool pipe, hipStream_t s)
{
    // the global step is step.off (direct launch) or base + off with a base
    // on the refresh grain (replayed group): its phase mod kFold is off's
    const bool fold = (step.off & (kFold - 1)) == 1;
#define LAUNCH_CG_M(V, PIPE)                                                                                     \
    do {                                                                                                         \
        if (fold)                                                                                                \
            hipLaunchKernelGGL((k_cg_m<V, PIPE, true>), dim3(grid), dim3(kThreads), 0, s, P, t3, nt3, step, ran);  \
        else                                                                                                     \
            hipLaunchKernelGGL((k_cg_m<V, PIPE, false>), dim3(grid), dim3(kThreads), 0, s, P, t3, nt3, step, ran); \
    } while (0)
    if (vx == 4 && pipe)
        LAUNCH_CG_M(4, true);
    else if (vx == 4)
        LAUNCH_CG_M(4, false);
    else if (vx == 2 && pipe)
        LAUNCH_CG_M(2, true);
    else if (vx == 2)
        LAUNCH_CG_M(2, false);
    else
        LAUNCH_CG_M(1, false);
#undef LAUNCH_CG_M
}

void launch_cg_b(bool fp32, const ProbDev *P, const TileRef *t1, int nt1, int grid, StepRef step, hipStream_t s)
{
    if (fp32)
        hipLaunchKernelGGL(k_cg_b<float>, dim3(grid), dim3(kThreads), 0, s, P, t1, nt1, step);
    else
        hipLaunchKernelGGL(k_cg_b<double>, dim3(grid), dim3(kThreads), 0, s, P, t1, nt1, step);
}

void launch_refresh(const ProbDev *P, const TileRef *t1, int nt1, const TileRef *t3, int nt3, const int *probs,
                    int nprobs, int grid, StepRef step, hipStream_t s)
{
    hipLaunchKernelGGL(k_refresh_e, dim3(std::min(grid, nt1)), dim3(kThreads), 0, s, P, t1, nt1, step);
    hipLaunchKernelGGL((k_true_residual<true, false>), dim3(std::min(grid, nt3)), dim3(kThreads), 0, s, P, t3, nt3,
                       step);
    hipLaunchKernelGGL(k_refresh_decide, dim3(nprobs), dim3(kThreads), 0, s, P, probs, step);
}

void launch_finish(const ProbDev *P, const TileRef *t3, int nt3, const int *probs, int nprobs, int grid,
                   hipStream_t s)
{
    hipLaunchKernelGGL((k_true_residual<false, true>), dim3(std::min(grid, nt3)), dim3(kThreads), 0, s, P, t3, nt3,
                       StepRef{nullptr, 0, nullptr});
    hipLaunchKernelGGL(k_finish, dim3(nprobs), dim3(kThreads), 0, s, P, probs);
}

// Grids: 1024 workgroups copy fastest on this GPU (5.8 TB/s against 5.1
// with 2048), 2048 read fastest.
void launch_read(const void *src, float *out, long long n16, hipStream_t s)
{
    hipLaunchKernelGGL(k_read, dim3(kReadGrid), dim3(kThreads), 0, s, static_cast<const float4 *>(src), out, n16);
}

void launch_copy(const void *src, void *dst, long long n16, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy, dim3(1024), dim3(kThreads), 0, s, static_cast<const float4 *>(src),
                       static_cast<float4 *>(dst), n16);
}

void launch_set_step(int *base, int value, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, s, base, value);
}

void launch_spmv_only(bool fp32, const ProbDev *P, const TileRef *t3, int nt3, int grid, hipStream_t s)
{
    if (fp32)
        hipLaunchKernelGGL(k_spmv<float>, dim3(std::min(grid, nt3)), dim3(kThreads), 0, s, P, t3, nt3);
    else
        hipLaunchKernelGGL(k_spmv<double>, dim3(std::min(grid, nt3)), dim3(kThreads), 0, s, P, t3, nt3);
}

}  // namespace sps
