// Host side of the pressure solver: device workspace, the CG drivers
// (FP64 and mixed-precision residual replacement) and the C ABI declared
// in include/suli_ps.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/suli_ps.h"
#include "ps_internal.h"

using namespace sps;

namespace {

constexpr int kMaxTimed = 256;  // SpMV launches timed per solve

struct Workspace {
    int nx = 0, ny = 0, nz = 0;
    size_t bytes = 0;
    std::vector<void *> allocs;
    // operator (FP64 scaled coefficients, diagonal, effective rhs)
    double *sx = nullptr, *sy = nullptr, *sz = nullptr, *d = nullptr, *b = nullptr;
    // FP64 CG vectors (p and p2: the search direction, alternating by iteration)
    double *e = nullptr, *r = nullptr, *p = nullptr, *p2 = nullptr, *q = nullptr;
    // FP32 inner-solve operator and vectors (mixed precision)
    float *sx32 = nullptr, *sy32 = nullptr, *sz32 = nullptr;
    float *c32 = nullptr, *r32 = nullptr, *p32 = nullptr, *p32b = nullptr, *q32 = nullptr;
    double *partial = nullptr;   // one buffer, carved into Partials
    Partials P{};
    CgScalars *S = nullptr;
    CgScalars *Stime = nullptr;  // scratch state for suli_ps_time_spmv
    int *ran = nullptr;          // timed SpMV launches that did the work
    // staging for the host-pointer API (reference layouts)
    double *u = nullptr, *v = nullptr, *w = nullptr, *rho = nullptr, *prd1 = nullptr;
    double *obsx = nullptr, *obsy = nullptr, *obsz = nullptr, *dz = nullptr, *dzz = nullptr;
    bool have_mixed = false;
    bool have_operator = false;
    Geo geo{};
};

}  // namespace

struct suli_ps_ctx {
    suli_ps_opts opts{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    Workspace ws;
    CgScalars *host_status = nullptr;  // pinned, 2 slots
    hipEvent_t ev_status[2] = {nullptr, nullptr};
    hipEvent_t ev_solve[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> ev_spmv;   // timing pairs
};

namespace {

int fail(suli_ps_ctx *ctx, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) ctx->err = buf;
    return code;
}

#define HIPCHK(ctx, call)                                                                                    \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return fail((ctx), SULI_PS_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                                           \
    } while (0)

suli_ps_opts default_opts()
{
    suli_ps_opts o{};
    o.precision = SULI_PS_FP64;
    o.stop_rule = SULI_PS_STOP_REFERENCE;
    o.tol = 1e-8;
    o.max_iter = 10000;
    o.chunk = 0;
    o.verbose = 1;
    o.timing = 0;
    return o;
}

void free_ws(Workspace &ws)
{
    for (void *p : ws.allocs) (void)hipFree(p);
    ws = Workspace{};
}

template <typename T>
int dalloc(suli_ps_ctx *ctx, T **p, size_t count)
{
    void *q = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(ctx, SULI_PS_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    ctx->ws.allocs.push_back(q);
    ctx->ws.bytes += bytes;
    *p = static_cast<T *>(q);
    return 0;
}

// Allocates the device workspace for a grid (kept while the shape is unchanged).
int ensure_ws(suli_ps_ctx *ctx, int nx, int ny, int nz, bool mixed)
{
    Workspace &ws = ctx->ws;
    if (ws.nx == nx && ws.ny == ny && ws.nz == nz && (!mixed || ws.have_mixed)) return 0;
    if (ws.nx != nx || ws.ny != ny || ws.nz != nz) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        free_ws(ws);
    }
    const size_t n = (size_t)nx * ny * nz;
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz, nfz = (size_t)nx * ny * (nz + 1);
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    int rc = 0;
    if (!ws.sx) {
        ws.nx = nx;
        ws.ny = ny;
        ws.nz = nz;
        ws.geo = make_geo(nx, ny, nz, 0, 0);
        const size_t b1 = (size_t)spmv_blocks(ws.geo), b2 = (size_t)kVecBlocks;
        // spmv[b1] upd[b2] tres[3 b1] init[5 b2] proj[2 b2]
        const size_t npart = b1 + b2 + 3 * b1 + 5 * b2 + 2 * b2;
        if ((rc = dalloc(ctx, &ws.sx, nfx)) || (rc = dalloc(ctx, &ws.sy, nfy)) || (rc = dalloc(ctx, &ws.sz, nfz)) ||
            (rc = dalloc(ctx, &ws.d, n)) || (rc = dalloc(ctx, &ws.b, n)) || (rc = dalloc(ctx, &ws.e, n)) ||
            (rc = dalloc(ctx, &ws.r, n)) || (rc = dalloc(ctx, &ws.p, n)) || (rc = dalloc(ctx, &ws.p2, n)) ||
            (rc = dalloc(ctx, &ws.q, n)) ||
            (rc = dalloc(ctx, &ws.partial, npart)) || (rc = dalloc(ctx, &ws.S, 1)) ||
            (rc = dalloc(ctx, &ws.ran, (size_t)kMaxTimed)) ||
            (rc = dalloc(ctx, &ws.Stime, 1)) || (rc = dalloc(ctx, &ws.u, (size_t)(nx + 3) * (ny + 2) * (nz + 2))) ||
            (rc = dalloc(ctx, &ws.v, (size_t)(nx + 2) * (ny + 3) * (nz + 2))) ||
            (rc = dalloc(ctx, &ws.w, (size_t)(nx + 2) * (ny + 2) * (nz + 3))) || (rc = dalloc(ctx, &ws.rho, n)) ||
            (rc = dalloc(ctx, &ws.prd1, nbox)) || (rc = dalloc(ctx, &ws.obsx, nfx)) ||
            (rc = dalloc(ctx, &ws.obsy, nfy)) || (rc = dalloc(ctx, &ws.obsz, nfz)) ||
            (rc = dalloc(ctx, &ws.dz, (size_t)nz + 2)) || (rc = dalloc(ctx, &ws.dzz, (size_t)nz + 2))) {
            free_ws(ws);
            return rc;
        }
        ws.P.spmv = ws.partial;
        ws.P.upd = ws.P.spmv + b1;
        ws.P.tres = ws.P.upd + b2;
        ws.P.init = ws.P.tres + 3 * b1;
        ws.P.proj = ws.P.init + 5 * b2;
        HIPCHK(ctx, hipMemsetAsync(ws.S, 0, sizeof(CgScalars), ctx->stream));
        HIPCHK(ctx, hipMemsetAsync(ws.Stime, 0, sizeof(CgScalars), ctx->stream));
    }
    if (mixed && !ws.have_mixed) {
        if ((rc = dalloc(ctx, &ws.sx32, nfx)) || (rc = dalloc(ctx, &ws.sy32, nfy)) ||
            (rc = dalloc(ctx, &ws.sz32, nfz)) || (rc = dalloc(ctx, &ws.c32, n)) || (rc = dalloc(ctx, &ws.r32, n)) ||
            (rc = dalloc(ctx, &ws.p32, n)) || (rc = dalloc(ctx, &ws.p32b, n)) || (rc = dalloc(ctx, &ws.q32, n))) {
            free_ws(ws);
            return rc;
        }
        ws.have_mixed = true;
    }
    return 0;
}

int check_grid(suli_ps_ctx *ctx, const suli_ps_grid *g)
{
    if (!g) return fail(ctx, SULI_PS_EINVAL, "grid is NULL");
    if (g->nx < 2 || g->ny < 2 || g->nz < 2)
        return fail(ctx, SULI_PS_EINVAL, "grid %dx%dx%d: every dimension must be >= 2", g->nx, g->ny, g->nz);
    if ((long long)g->nx * g->ny * g->nz > (1LL << 31) - 1)
        return fail(ctx, SULI_PS_EINVAL, "grid %dx%dx%d is too large", g->nx, g->ny, g->nz);
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dt > 0))
        return fail(ctx, SULI_PS_EINVAL, "dx, dy and dt must be positive");
    if (g->t_press != 1 && g->t_press != 2)
        return fail(ctx, SULI_PS_EINVAL, "t_press = %d: only 1 (projection) and 2 (IPA) solve for pressure",
                    g->t_press);
    if (g->t_press == 2 && !(g->rho_f1 > 0)) return fail(ctx, SULI_PS_EINVAL, "IPA needs rho_f1 > 0");
    return 0;
}

SetupIn make_setup(const suli_ps_grid *g, const double *u, const double *v, const double *w, const double *rho,
                   const double *prd1, const double *ox, const double *oy, const double *oz, const double *dz,
                   const double *dzz)
{
    SetupIn in{};
    in.u = u;
    in.v = v;
    in.w = w;
    in.rho = rho;
    in.prd1 = prd1;
    const bool mask = g->obst_t != 0 && g->ibm_t == 2;
    in.obsx = mask ? ox : nullptr;
    in.obsy = mask ? oy : nullptr;
    in.obsz = mask ? oz : nullptr;
    in.dz = dz;
    in.dzz = dzz;
    in.dx = g->dx;
    in.dy = g->dy;
    in.dzu = g->dz;
    in.dt = g->dt;
    // aux = dt / (delta*delta), 9_grad.f90:93-100
    in.aux1 = g->dt / (g->dx * g->dx);
    in.aux2 = g->dt / (g->dy * g->dy);
    in.aux3 = g->dt / (g->dz * g->dz);
    in.rho_f1 = g->rho_f1;
    in.ipa = g->t_press == 2;
    return in;
}

int auto_chunk(const Geo &g)
{
    // small grids iterate fast: poll less often relative to the work
    if (g.n <= 200000) return 64;
    return 32;
}

int run_graddin_dev(suli_ps_ctx *ctx, const suli_ps_grid *gr, const double *u, const double *v, const double *w,
                    const double *rho, const double *ox, const double *oy, const double *oz, const double *dz,
                    const double *dzz, double *prd1, suli_ps_stats *stats)
{
    int rc = check_grid(ctx, gr);
    if (rc) return rc;
    if (!u || !v || !w || !prd1 || (gr->t_press == 1 && !rho))
        return fail(ctx, SULI_PS_EINVAL, "u, v, w, rho and prd1 must be non-NULL");
    if ((dz == nullptr) != (dzz == nullptr)) return fail(ctx, SULI_PS_EINVAL, "dz and dzz go together");
    const bool mask = gr->obst_t != 0 && gr->ibm_t == 2;
    if (mask && (!ox || !oy || !oz)) return fail(ctx, SULI_PS_EINVAL, "IBM mask on (obst_t, ibm_t=2) needs obs_ls*");
    if (!dz && !(gr->dz > 0)) return fail(ctx, SULI_PS_EINVAL, "dz must be positive");
    const suli_ps_opts &o = ctx->opts;
    const bool mixed = o.precision == SULI_PS_MIXED;
    if ((rc = ensure_ws(ctx, gr->nx, gr->ny, gr->nz, mixed))) return rc;
    Workspace &ws = ctx->ws;
    hipStream_t s = ctx->stream;
    Geo g = make_geo(gr->nx, gr->ny, gr->nz, gr->ccx0 == 0, gr->ccy0 == 0);
    ws.geo = g;
    const SetupIn in = make_setup(gr, u, v, w, rho, prd1, ox, oy, oz, dz, dzz);

    HIPCHK(ctx, hipEventRecord(ctx->ev_solve[0], s));
    CgScalars s0{};
    s0.max_iter = o.max_iter > 0 ? o.max_iter : 10000;
    s0.mixed = mixed ? 1 : 0;
    s0.project = o.keep_null ? 0 : 1;
    s0.rel_rule = o.stop_rule == SULI_PS_STOP_RELATIVE;
    s0.tol = o.tol;
    // threshold of the reference stop rule: 0.0001/(nx*ny*nz) (9_grad.f90:393)
    s0.thresh_ref = 0.0001 / (double)(gr->nx * gr->ny * gr->nz);
    s0.done_at = kNotDone;
    HIPCHK(ctx, hipMemcpyAsync(ws.S, &s0, sizeof s0, hipMemcpyHostToDevice, s));
    if (o.timing) HIPCHK(ctx, hipMemsetAsync(ws.ran, 0, kMaxTimed * sizeof(int), s));

    launch_face_coefs(g, in, ws.sx, ws.sy, ws.sz, s);
    launch_diag_rhs(g, in, ws.sx, ws.sy, ws.sz, ws.d, ws.b, s);
    launch_scale_faces(g, ws.d, ws.sx, ws.sy, ws.sz, mixed ? ws.sx32 : nullptr, mixed ? ws.sy32 : nullptr,
                       mixed ? ws.sz32 : nullptr, s);
    const Op<double> op64{ws.sx, ws.sy, ws.sz};
    const Op<float> op32{ws.sx32, ws.sy32, ws.sz32};
    launch_init(g, in, op64, ws.d, ws.b, ws.e, ws.r, mixed ? nullptr : ws.r, ws.r32, ws.c32, ws.P, ws.S, s);
    HIPCHK(ctx, hipGetLastError());
    ws.have_operator = true;

    const int max_iter = s0.max_iter;
    // SpMV timing: events around every 16th cg_a launch; ws.ran[slot] says
    // whether that launch did the SpMV (it skips after the stop rule)
    const int tstride = 16;
    int ntimed = 0;
    int n = 0;  // CG iteration number of the launch sequence (1-based)
    double *p64[2] = {ws.p, ws.p2};
    float *p32[2] = {ws.p32, ws.p32b};

    auto enqueue_chunk = [&](int chunk_idx, int K) -> int {
        for (int it = 0; it < K; ++it) {
            ++n;
            const bool timed = o.timing && (n % tstride == 2) && ntimed < kMaxTimed;
            int *ran = timed ? ws.ran + ntimed : nullptr;
            // after a residual replacement (every mixed chunk but the first
            // starts after one) the deferred e update is already in e
            const bool do_e = !(mixed && it == 0 && chunk_idx > 0);
            if (timed) HIPCHK(ctx, hipEventRecord(ctx->ev_spmv[2 * ntimed], s));
            if (mixed)
                launch_cg_a<float>(g, n, do_e, ws.c32, p32[(n - 1) & 1], p32[n & 1], ws.r32, op32, ws.q32, ws.P,
                                   ws.S, ran, s);
            else
                launch_cg_a<double>(g, n, do_e, ws.e, p64[(n - 1) & 1], p64[n & 1], ws.r, op64, ws.q, ws.P, ws.S,
                                    ran, s);
            if (timed) {
                HIPCHK(ctx, hipEventRecord(ctx->ev_spmv[2 * ntimed + 1], s));
                ++ntimed;
            }
            if (mixed)
                launch_cg_b<float>(g, n, ws.r32, ws.q32, ws.P, ws.S, s);
            else
                launch_cg_b<double>(g, n, ws.r, ws.q, ws.P, ws.S, s);
        }
        if (mixed) launch_refresh(g, n, ws.e, ws.c32, p32[n & 1], op64, ws.b, ws.r32, ws.P, ws.S, s);
        HIPCHK(ctx, hipGetLastError());
        const int slot = chunk_idx & 1;
        HIPCHK(ctx, hipMemcpyAsync(&ctx->host_status[slot], ws.S, sizeof(CgScalars), hipMemcpyDeviceToHost, s));
        HIPCHK(ctx, hipEventRecord(ctx->ev_status[slot], s));
        return 0;
    };

    // Chunk sizes.  FP64: fixed; the stop rule is tested on the device every
    // iteration and later launches return at once.  Mixed: every chunk ends
    // with an FP64 residual replacement, which is also where convergence is
    // decided, so chunk lengths follow the observed convergence rate: long
    // enough to amortise the replacement, short enough not to overshoot the
    // tolerance and to keep the FP32 recurrence within ~3 decades of the
    // residual it started from.
    const int K_fixed = o.chunk > 0 ? o.chunk : auto_chunk(g);
    int known_iters = 0;          // from the latest status read
    double known_rr = 0.0, known_thresh = 0.0, rate = 0.0;
    std::vector<int> chunk_len;
    auto next_chunk = [&](int c) -> int {
        if (!mixed || o.chunk > 0) return K_fixed;
        if (!(rate < 0.0) || !(known_rr > 0.0)) return 32;
        int in_flight = 0;  // iterations enqueued beyond the known status
        for (int i = 0; i < c; ++i) in_flight += chunk_len[i];
        in_flight -= known_iters;
        const double remaining =
            known_thresh > 0 ? std::log(known_thresh / known_rr) / rate - in_flight : 32.0;
        const double fp32_span = std::log(1e-6) / rate;  // rr drop of 1e6 at most
        double k = std::min(remaining + 2.0, fp32_span);
        return (int)std::max(8.0, std::min(128.0, std::ceil(k)));
    };

    // Two chunks in flight: the host reads chunk c-1's state while chunk c
    // runs.
    bool finished = false;
    const int max_chunks = max_iter / 8 + 16;
    chunk_len.push_back(next_chunk(0));
    if ((rc = enqueue_chunk(0, chunk_len[0]))) return rc;
    for (int c = 1; c <= max_chunks; ++c) {
        chunk_len.push_back(next_chunk(c));
        if ((rc = enqueue_chunk(c, chunk_len[c]))) return rc;
        const int prev = (c - 1) & 1;
        HIPCHK(ctx, hipEventSynchronize(ctx->ev_status[prev]));
        const CgScalars &st = ctx->host_status[prev];
        if (st.final_) {
            finished = true;
            break;
        }
        if (mixed && st.iters > known_iters && st.rr_true > 0.0) {
            if (known_rr > 0.0 && st.rr_true < known_rr)
                rate = std::log(st.rr_true / known_rr) / (st.iters - known_iters);
            known_iters = st.iters;
            known_rr = st.rr_true;
        }
        known_thresh = st.thresh;
    }
    if (!finished) {
        HIPCHK(ctx, hipStreamSynchronize(s));
        return fail(ctx, SULI_PS_EHIP, "solver did not finish within %d chunks", max_chunks);
    }
    launch_finish(g, ws.e, ws.d, op64, ws.b, ws.P, ws.S, s);
    launch_unscale(g, ws.e, ws.d, prd1, ws.S, s);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev_solve[1], s));
    CgScalars fin{};
    HIPCHK(ctx, hipMemcpyAsync(&fin, ws.S, sizeof fin, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));

    const int status = fin.status == 0 ? SULI_PS_OK : SULI_PS_NOT_CONVERGED;
    // r.r of the recurrence after the last iteration (slot of iteration done_at)
    const double rr_last = fin.rr[(fin.done_at == kNotDone ? 0 : fin.done_at) & 1];
    if (stats) {
        std::memset(stats, 0, sizeof *stats);
        stats->status = status;
        stats->iters = fin.iters;
        stats->refreshes = fin.refreshes;
        stats->rr0 = fin.rr0;
        stats->bb = fin.bb;
        stats->rr = mixed ? fin.rr_true : rr_last;
        stats->thresh = fin.thresh;
        stats->rel_true = fin.bb > 0 ? std::sqrt(fin.rr_true / fin.bb) : 0.0;
        float ms = 0.f;
        HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev_solve[0], ctx->ev_solve[1]));
        stats->solve_ms = ms;
        double tsum = 0.0;
        int tcount = 0;
        int ran[kMaxTimed] = {0};
        if (ntimed > 0) HIPCHK(ctx, hipMemcpy(ran, ws.ran, ntimed * sizeof(int), hipMemcpyDeviceToHost));
        for (int t = 0; t < ntimed; ++t) {
            if (!ran[t]) continue;
            float dt = 0.f;
            HIPCHK(ctx, hipEventElapsedTime(&dt, ctx->ev_spmv[2 * t], ctx->ev_spmv[2 * t + 1]));
            tsum += dt;
            ++tcount;
        }
        stats->spmv_timed = tcount;
        stats->spmv_ms = tcount ? tsum / tcount : 0.0;
    }
    if (status == SULI_PS_NOT_CONVERGED && o.verbose) {
        // the reference's warning (9_grad.f90:590)
        std::printf(" pulou pressão; erro = %.15E\n", mixed ? fin.rr_true : rr_last);
        std::fflush(stdout);
    }
    return status;
}

int run_posdin_dev(suli_ps_ctx *ctx, const suli_ps_grid *gr, double *u, double *v, double *w, const double *rho,
                   const double *prd1, const double *dz, const double *dzz, double *prd)
{
    int rc = check_grid(ctx, gr);
    if (rc) return rc;
    if (!u || !v || !w || !rho || !prd1) return fail(ctx, SULI_PS_EINVAL, "u, v, w, rho and prd1 must be non-NULL");
    Geo g = make_geo(gr->nx, gr->ny, gr->nz, gr->ccx0 == 0, gr->ccy0 == 0);
    SetupIn in = make_setup(gr, nullptr, nullptr, nullptr, rho, prd1, nullptr, nullptr, nullptr, dz, dzz);
    launch_posdin(g, in, u, v, w, ctx->stream);
    HIPCHK(ctx, hipGetLastError());
    if (prd && prd != prd1) {
        const size_t nbox = (size_t)(gr->nx + 2) * (gr->ny + 2) * (gr->nz + 2);
        HIPCHK(ctx, hipMemcpyAsync(prd, prd1, nbox * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // namespace

extern "C" {

int suli_ps_abi_version(void) { return SULI_PS_ABI_VERSION; }

suli_ps_ctx *suli_ps_create(const suli_ps_opts *opts)
{
    suli_ps_ctx *ctx = new suli_ps_ctx();
    ctx->opts = opts ? *opts : default_opts();
    if (ctx->opts.max_iter <= 0) ctx->opts.max_iter = 10000;
    bool ok = hipGetDevice(&ctx->device) == hipSuccess &&
              hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess &&
              hipHostMalloc((void **)&ctx->host_status, 2 * sizeof(CgScalars), hipHostMallocDefault) == hipSuccess;
    for (int i = 0; ok && i < 2; ++i)
        ok = hipEventCreateWithFlags(&ctx->ev_status[i], hipEventDisableTiming) == hipSuccess &&
             hipEventCreate(&ctx->ev_solve[i]) == hipSuccess;
    if (ok) {
        ctx->ev_spmv.resize(512);
        for (auto &e : ctx->ev_spmv) ok = ok && hipEventCreate(&e) == hipSuccess;
    }
    if (!ok) {
        suli_ps_destroy(ctx);
        return nullptr;
    }
    return ctx;
}

void suli_ps_destroy(suli_ps_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_ws(ctx->ws);
    for (auto &e : ctx->ev_spmv)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; ++i) {
        if (ctx->ev_status[i]) (void)hipEventDestroy(ctx->ev_status[i]);
        if (ctx->ev_solve[i]) (void)hipEventDestroy(ctx->ev_solve[i]);
    }
    if (ctx->host_status) (void)hipHostFree(ctx->host_status);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int suli_ps_set_opts(suli_ps_ctx *ctx, const suli_ps_opts *opts)
{
    if (!ctx || !opts) return SULI_PS_EINVAL;
    if (opts->precision != SULI_PS_FP64 && opts->precision != SULI_PS_MIXED)
        return fail(ctx, SULI_PS_EINVAL, "unknown precision %d", opts->precision);
    if (opts->stop_rule != SULI_PS_STOP_REFERENCE && opts->stop_rule != SULI_PS_STOP_RELATIVE)
        return fail(ctx, SULI_PS_EINVAL, "unknown stop rule %d", opts->stop_rule);
    if (opts->stop_rule == SULI_PS_STOP_RELATIVE && !(opts->tol > 0))
        return fail(ctx, SULI_PS_EINVAL, "relative stop rule needs tol > 0");
    ctx->opts = *opts;
    if (ctx->opts.max_iter <= 0) ctx->opts.max_iter = 10000;
    return 0;
}

const char *suli_ps_last_error(const suli_ps_ctx *ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int suli_ps_graddin_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *u, const double *v, const double *w,
                        const double *rho, const double *obs_lsx, const double *obs_lsy, const double *obs_lsz,
                        const double *dz, const double *dzz, double *prd1, suli_ps_stats *stats)
{
    if (!ctx) return SULI_PS_EINVAL;
    return run_graddin_dev(ctx, g, u, v, w, rho, obs_lsx, obs_lsy, obs_lsz, dz, dzz, prd1, stats);
}

int suli_ps_posdin_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v, double *w, const double *rho,
                       const double *prd1, const double *dz, const double *dzz, double *prd)
{
    if (!ctx) return SULI_PS_EINVAL;
    return run_posdin_dev(ctx, g, u, v, w, rho, prd1, dz, dzz, prd);
}

int suli_ps_graddin(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *u, const double *v, const double *w,
                    const double *rho, const double *obs_lsx, const double *obs_lsy, const double *obs_lsz,
                    const double *dz, const double *dzz, double *prd1, suli_ps_stats *stats)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!u || !v || !w || !prd1 || (g->t_press == 1 && !rho))
        return fail(ctx, SULI_PS_EINVAL, "u, v, w, rho and prd1 must be non-NULL");
    const bool mask = g->obst_t != 0 && g->ibm_t == 2;
    if (mask && (!obs_lsx || !obs_lsy || !obs_lsz))
        return fail(ctx, SULI_PS_EINVAL, "IBM mask on (obst_t, ibm_t=2) needs obs_ls*");
    if ((dz == nullptr) != (dzz == nullptr)) return fail(ctx, SULI_PS_EINVAL, "dz and dzz go together");
    if ((rc = ensure_ws(ctx, g->nx, g->ny, g->nz, ctx->opts.precision == SULI_PS_MIXED))) return rc;
    Workspace &ws = ctx->ws;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t n = (size_t)nx * ny * nz;
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz, nfz = (size_t)nx * ny * (nz + 1);
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(ws.u, u, sizeof(double) * (nx + 3) * (ny + 2) * (nz + 2), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ws.v, v, sizeof(double) * (nx + 2) * (ny + 3) * (nz + 2), hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ws.w, w, sizeof(double) * (nx + 2) * (ny + 2) * (nz + 3), hipMemcpyHostToDevice, s));
    if (rho) HIPCHK(ctx, hipMemcpyAsync(ws.rho, rho, sizeof(double) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ws.prd1, prd1, sizeof(double) * nbox, hipMemcpyHostToDevice, s));
    if (mask) {
        HIPCHK(ctx, hipMemcpyAsync(ws.obsx, obs_lsx, sizeof(double) * nfx, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ws.obsy, obs_lsy, sizeof(double) * nfy, hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ws.obsz, obs_lsz, sizeof(double) * nfz, hipMemcpyHostToDevice, s));
    }
    if (dz) {
        HIPCHK(ctx, hipMemcpyAsync(ws.dz, dz, sizeof(double) * (nz + 2), hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ws.dzz, dzz, sizeof(double) * (nz + 2), hipMemcpyHostToDevice, s));
    }
    rc = run_graddin_dev(ctx, g, ws.u, ws.v, ws.w, rho ? ws.rho : nullptr, ws.obsx, ws.obsy, ws.obsz,
                         dz ? ws.dz : nullptr, dz ? ws.dzz : nullptr, ws.prd1, stats);
    if (rc < 0) return rc;
    HIPCHK(ctx, hipMemcpyAsync(prd1, ws.prd1, sizeof(double) * nbox, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    return rc;
}

int suli_ps_posdin(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v, double *w, const double *rho,
                   const double *prd1, const double *dz, const double *dzz, double *prd)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    if (!u || !v || !w || !rho || !prd1) return fail(ctx, SULI_PS_EINVAL, "u, v, w, rho and prd1 must be non-NULL");
    if ((dz == nullptr) != (dzz == nullptr)) return fail(ctx, SULI_PS_EINVAL, "dz and dzz go together");
    if ((rc = ensure_ws(ctx, g->nx, g->ny, g->nz, false))) return rc;
    Workspace &ws = ctx->ws;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t n = (size_t)nx * ny * nz;
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t nu = (size_t)(nx + 3) * (ny + 2) * (nz + 2), nv = (size_t)(nx + 2) * (ny + 3) * (nz + 2),
                 nw = (size_t)(nx + 2) * (ny + 2) * (nz + 3);
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipMemcpyAsync(ws.u, u, sizeof(double) * nu, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ws.v, v, sizeof(double) * nv, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ws.w, w, sizeof(double) * nw, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ws.rho, rho, sizeof(double) * n, hipMemcpyHostToDevice, s));
    HIPCHK(ctx, hipMemcpyAsync(ws.prd1, prd1, sizeof(double) * nbox, hipMemcpyHostToDevice, s));
    if (dz) {
        HIPCHK(ctx, hipMemcpyAsync(ws.dz, dz, sizeof(double) * (nz + 2), hipMemcpyHostToDevice, s));
        HIPCHK(ctx, hipMemcpyAsync(ws.dzz, dzz, sizeof(double) * (nz + 2), hipMemcpyHostToDevice, s));
    }
    rc = run_posdin_dev(ctx, g, ws.u, ws.v, ws.w, ws.rho, ws.prd1, dz ? ws.dz : nullptr, dz ? ws.dzz : nullptr,
                        nullptr);
    if (rc) return rc;
    HIPCHK(ctx, hipMemcpyAsync(u, ws.u, sizeof(double) * nu, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(v, ws.v, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(w, ws.w, sizeof(double) * nw, hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    if (prd && prd != prd1) std::memcpy(prd, prd1, sizeof(double) * nbox);  // prd = prd1 (10_pos.f90:95)
    return 0;
}

void *suli_ps_dev_alloc(suli_ps_ctx *ctx, size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 1)) != hipSuccess) {
        if (ctx) ctx->err = "hipMalloc failed";
        return nullptr;
    }
    return p;
}

int suli_ps_dev_free(suli_ps_ctx *ctx, void *ptr)
{
    HIPCHK(ctx, hipFree(ptr));
    return 0;
}

int suli_ps_memcpy_h2d(suli_ps_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int suli_ps_memcpy_d2h(suli_ps_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int suli_ps_dev_zero(suli_ps_ctx *ctx, void *ptr, size_t bytes)
{
    if (!ctx || !ptr) return SULI_PS_EINVAL;
    HIPCHK(ctx, hipMemsetAsync(ptr, 0, bytes, ctx->stream));
    return 0;
}

int suli_ps_synchronize(suli_ps_ctx *ctx)
{
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int suli_ps_time_spmv(suli_ps_ctx *ctx, int32_t precision, int32_t reps, double *ms)
{
    if (!ctx || !ms || reps <= 0) return SULI_PS_EINVAL;
    Workspace &ws = ctx->ws;
    if (!ws.have_operator) return fail(ctx, SULI_PS_EINVAL, "no operator: call graddin first");
    const bool mixed = precision == SULI_PS_MIXED;
    if (mixed && !ws.have_mixed) return fail(ctx, SULI_PS_EINVAL, "no FP32 operator: solve in mixed mode first");
    hipStream_t s = ctx->stream;
    const Geo &g = ws.geo;
    // the plain SpMV q = A p (5 words per cell) on the last solve's operator:
    // one untimed launch, then `reps` launches between two events
    auto launch = [&]() {
        if (mixed)
            launch_spmv_only<float>(g, ws.p32, Op<float>{ws.sx32, ws.sy32, ws.sz32}, ws.q32, ws.P, s);
        else
            launch_spmv_only<double>(g, ws.p, Op<double>{ws.sx, ws.sy, ws.sz}, ws.q, ws.P, s);
    };
    launch();
    HIPCHK(ctx, hipEventRecord(ctx->ev_solve[0], s));
    for (int i = 0; i < reps; ++i) launch();
    HIPCHK(ctx, hipEventRecord(ctx->ev_solve[1], s));
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventSynchronize(ctx->ev_solve[1]));
    float t = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&t, ctx->ev_solve[0], ctx->ev_solve[1]));
    *ms = (double)t / reps;
    return 0;
}

}  // extern "C"
