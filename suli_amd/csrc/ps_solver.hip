// The C ABI declared in include/suli_ps.h: single solves, batches and the
// velocity correction, on top of the solve engine (ps_engine.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/suli_ps.h"
#include "ps_common.h"
#include "ps_engine.h"
#include "ps_internal.h"
#include "ps_transfer.h"

using namespace sps;

// Solver state of suli_ps_graddin_noalloc_dev: one problem's block (the
// arrays Engine::admit carves), SULI_PS_NOALLOC_MAX_BYTES on every GPU the
// library runs on, held until the process ends.  In host code `scratch` is
// only a stand-in for the device array: its device address is looked up
// once per device (scratch_address) and that address is what the engine
// and the kernels are given.
static __device__ __align__(256) unsigned char scratch[SULI_PS_NOALLOC_MAX_BYTES];

namespace {

constexpr int kMaxSlots = 256;   // problems resident at once per context
// Launch-table entries reserved by every context for the no-allocation entry
// point (3-D and 1-D tiles of its one problem).
constexpr int kNoAllocEntries = 32768;
constexpr int kMaxDevices = 64;

struct ScratchTable {
    std::mutex mu;
    bool looked_up[kMaxDevices] = {};
    hipError_t err[kMaxDevices] = {};
    void *ptr[kMaxDevices] = {};
};
ScratchTable &scratch_table()
{
    static ScratchTable t;
    return t;
}

// Device address of `scratch` on `device`, which must be current: looked up
// with hipGetSymbolAddress the first time, then kept (each GPU holds its own
// copy of the array at its own address).
hipError_t scratch_address(int device, void **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    ScratchTable &t = scratch_table();
    std::lock_guard<std::mutex> lk(t.mu);
    if (!t.looked_up[device]) {
        void *p = nullptr;
        t.err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        t.ptr[device] = t.err[device] == hipSuccess ? p : nullptr;
        t.looked_up[device] = true;
    }
    *out = t.ptr[device];
    return t.err[device];
}

// Device copies of the reference-layout arrays for the host-pointer API.
struct Staging {
    int nx = 0, ny = 0, nz = 0;
    void *block = nullptr;
    double *u = nullptr, *v = nullptr, *w = nullptr, *rho = nullptr, *prd1 = nullptr;
    double *obsx = nullptr, *obsy = nullptr, *obsz = nullptr, *dz = nullptr, *dzz = nullptr;
    double *hs = nullptr;
    double *dzx = nullptr, *dzy = nullptr, *dz3 = nullptr, *eta1 = nullptr, *deltah = nullptr;   // SuLi-IT
    double *box2 = nullptr, *eta0 = nullptr;   // a second halo box (prd0, SuLi-IT's dz) and eta0
    // What the last suli_ps_graddin left on the device (opts.reuse_staged):
    // its host pointers and grid, with u, v, w, rho as uploaded and prd1 as
    // solved.  Cleared by every other host-pointer call.
    struct Left {
        bool valid = false;
        suli_ps_grid grid{};
        const double *u = nullptr, *v = nullptr, *w = nullptr, *rho = nullptr, *prd1 = nullptr;
        const double *dz = nullptr, *dzz = nullptr;
    } left;
};

}  // namespace

struct suli_ps_ctx {
    suli_ps_opts opts{};
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    Engine eng;
    Staging st;
    HostPipe pipe;                 // pinned staging of the host-pointer copies
    std::vector<int> last_slots;   // slots of the last solve (kept for time_spmv)
    hipEvent_t ev[2] = {nullptr, nullptr};
    void *scratch = nullptr;       // device address of `scratch` on this device
    hipError_t scratch_err = hipSuccess;
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

// A host array of n doubles and its device copy, for HostPipe (n = 0 for
// an array the call does not have).
HostSeg seg(double *dev, const double *host, size_t n)
{
    return HostSeg{dev, const_cast<double *>(host), host ? n * sizeof(double) : 0};
}

// Host -> device copies of a host-pointer call, ahead of its work on the
// context's stream.
int upload(suli_ps_ctx *ctx, std::initializer_list<HostSeg> segs)
{
    HIPCHK(ctx, ctx->pipe.upload(ctx->device, ctx->stream, segs.begin(), (int)segs.size()));
    return 0;
}

// Device -> host copies after the work on the context's stream; the host
// arrays hold the results on return.
int download(suli_ps_ctx *ctx, std::initializer_list<HostSeg> segs)
{
    HIPCHK(ctx, ctx->pipe.download(ctx->device, ctx->stream, segs.begin(), (int)segs.size()));
    return 0;
}

// download(), and posdin's prd = prd1 (10_pos.f90:95) on the host arrays,
// done on a thread of its own while the copies run
int download_copy_prd(suli_ps_ctx *ctx, std::initializer_list<HostSeg> segs, double *prd, const double *prd1,
                      size_t nbox)
{
    const bool copy = prd && prd != prd1;
    std::thread t;
    if (copy) {
        try {
            t = std::thread([=] { std::memcpy(prd, prd1, sizeof(double) * nbox); });
        } catch (const std::exception &) {
        }
    }
    const int rc = download(ctx, segs);
    if (t.joinable())
        t.join();
    else if (copy)
        std::memcpy(prd, prd1, sizeof(double) * nbox);
    return rc;
}

suli_ps_opts default_opts()
{
    suli_ps_opts o{};
    o.precision = SULI_PS_FP64;
    o.stop_rule = SULI_PS_STOP_REFERENCE;
    o.tol = 1e-8;
    o.max_iter = 10000;
    o.verbose = 1;
    return o;
}

int check_grid(suli_ps_ctx *ctx, const suli_ps_grid *g)
{
    std::string e;
    const int rc = validate_grid(g, e);
    return rc ? fail(ctx, rc, "%s", e.c_str()) : 0;
}

int check_problem(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *u, const double *v, const double *w,
                  const double *rho, const double *ox, const double *oy, const double *oz, const double *dz,
                  const double *dzz, const double *prd1)
{
    std::string e;
    const int rc = validate_problem(g, u, v, w, rho, ox, oy, oz, dz, dzz, prd1, e);
    return rc ? fail(ctx, rc, "%s", e.c_str()) : 0;
}

void apply_engine_opts(suli_ps_ctx *ctx)
{
    ctx->eng.set_graphs(ctx->opts.no_graphs == 0);
    ctx->eng.set_autotune(ctx->opts.autotune != 0);
    ctx->eng.set_chunk(ctx->opts.chunk);
}

void release_last(suli_ps_ctx *ctx)
{
    for (int s : ctx->last_slots) ctx->eng.release(s);
    ctx->last_slots.clear();
}

void warn_capped(const suli_ps_ctx *ctx, const suli_ps_stats &st)
{
    if (st.status == SULI_PS_NOT_CONVERGED && ctx->opts.verbose) {
        // the reference's warning (9_grad.f90:590)
        std::printf(" pulou pressão; erro = %.15E\n", st.rr);
        std::fflush(stdout);
    }
}

// block: device memory for the problem's solver state (problem_bytes), or
// NULL for the engine's pool.
int run_graddin_dev(suli_ps_ctx *ctx, const suli_ps_grid *gr, const double *u, const double *v, const double *w,
                    const double *rho, const double *ox, const double *oy, const double *oz, const double *dz,
                    const double *dzz, double *prd1, suli_ps_stats *stats, const uint8_t *solid = nullptr,
                    void *block = nullptr)
{
    int rc = check_problem(ctx, gr, u, v, w, rho, ox, oy, oz, dz, dzz, prd1);
    if (rc) return rc;
    release_last(ctx);
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], s));
    const SetupIn in = make_setup(gr, u, v, w, rho, prd1, ox, oy, oz, dz, dzz, solid);
    int slot = -1;
    if ((rc = ctx->eng.admit(*gr, in, prd1, solve_opts(ctx->opts), 0, &slot, block)))
        return fail(ctx, rc, "%s", ctx->eng.error().c_str());
    ctx->last_slots.push_back(slot);
    if ((rc = ctx->eng.run_all(ctx->opts.timing != 0))) return fail(ctx, rc, "%s", ctx->eng.error().c_str());
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], s));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[1]));
    suli_ps_stats st;
    fill_stats(ctx->eng.slot(slot), &st);
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    st.solve_ms = ms;
    st.spmv_timed = ctx->eng.spmv_samples();
    st.spmv_ms = ctx->eng.spmv_ms_mean();
    st.graph_replays = ctx->eng.graph_replays();
    if (stats) *stats = st;
    warn_capped(ctx, st);
    return st.status;
}

int run_posdin_dev(suli_ps_ctx *ctx, const suli_ps_grid *gr, double *u, double *v, double *w, const double *rho,
                   const double *prd1, const double *dz, const double *dzz, double *prd,
                   const uint8_t *solid = nullptr)
{
    int rc = check_grid(ctx, gr);
    if (rc) return rc;
    if (!u || !v || !w || !rho || !prd1) return fail(ctx, SULI_PS_EINVAL, "u, v, w, rho and prd1 must be non-NULL");
    Geo g = make_geo(gr->nx, gr->ny, gr->nz, gr->ccx0 == 0, gr->ccy0 == 0);
    g.dtop = gr->top_bc == 1 ? 1 : 0;
    SetupIn in = make_setup(gr, nullptr, nullptr, nullptr, rho, prd1, nullptr, nullptr, nullptr, dz, dzz, solid);
    launch_posdin(g, in, u, v, w, ctx->stream);
    HIPCHK(ctx, hipGetLastError());
    if (prd && prd != prd1) {
        const size_t nbox = (size_t)(gr->nx + 2) * (gr->ny + 2) * (gr->nz + 2);
        HIPCHK(ctx, hipMemcpyAsync(prd, prd1, nbox * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int check_columns(suli_ps_ctx *ctx, const suli_ps_grid *gr, const double *dz)
{
    std::string e;
    const int rc = validate_shape(gr, e);
    if (rc) return fail(ctx, rc, "%s", e.c_str());
    if (!dz && !(gr->dz > 0)) return fail(ctx, SULI_PS_EINVAL, "dz must be positive");
    return 0;
}

int run_pressh_dev(suli_ps_ctx *ctx, const suli_ps_grid *gr, const double *rho, const double *dz, double gz,
                   double *prd1)
{
    int rc = check_columns(ctx, gr, dz);
    if (rc) return rc;
    if (!rho || !prd1) return fail(ctx, SULI_PS_EINVAL, "rho and prd1 must be non-NULL");
    const Geo g = make_geo(gr->nx, gr->ny, gr->nz, gr->ccx0 == 0, gr->ccy0 == 0);
    launch_pressh(g, rho, dz, gr->dz, gz, prd1, ctx->stream);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int run_posdin_hydro_dev(suli_ps_ctx *ctx, const suli_ps_grid *gr, double *u, double *v, double *w,
                         const double *rho, const double *hs, const double *prd1, const double *dz, double *prd)
{
    int rc = check_columns(ctx, gr, dz);
    if (rc) return rc;
    if (!u || !v || !w || !rho || !hs || !prd1)
        return fail(ctx, SULI_PS_EINVAL, "u, v, w, rho, hs and prd1 must be non-NULL");
    const Geo g = make_geo(gr->nx, gr->ny, gr->nz, gr->ccx0 == 0, gr->ccy0 == 0);
    const SetupIn in =
        make_setup(gr, nullptr, nullptr, nullptr, rho, prd1, nullptr, nullptr, nullptr, dz, nullptr, nullptr);
    launch_posdin_hydro(g, in, hs, u, v, w, ctx->stream);
    HIPCHK(ctx, hipGetLastError());
    if (prd && prd != prd1) {
        const size_t nbox = (size_t)(gr->nx + 2) * (gr->ny + 2) * (gr->nz + 2);
        HIPCHK(ctx, hipMemcpyAsync(prd, prd1, nbox * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

// SuLi-IT (SuLi-IT/9_grad.f90:10-263): the grid's sizes, spacings, dt and
// periodic flags; t_press, the IBM mask, IPA and the Dirichlet lid do not
// apply.
int check_it(suli_ps_ctx *ctx, const suli_ps_grid *gr, double tetah, double gz)
{
    std::string e;
    const int rc = validate_shape(gr, e);
    if (rc) return fail(ctx, rc, "%s", e.c_str());
    if (!(tetah > 0) || !(gz > 0)) return fail(ctx, SULI_PS_EINVAL, "tetah and gz must be positive");
    if (gr->top_bc != 0) return fail(ctx, SULI_PS_EINVAL, "SuLi-IT has its own free-surface top row: top_bc must be 0");
    return 0;
}

SetupIn make_setup_it(const suli_ps_grid *gr, double tetah, double gz, const double *u, const double *v,
                      const double *w, const double *dzx, const double *dzy, const double *dz3, const double *eta1,
                      const double *deltah, const double *prd1)
{
    SetupIn in{};
    in.u = u;
    in.v = v;
    in.w = w;
    in.prd1 = prd1;
    in.dx = gr->dx;
    in.dy = gr->dy;
    in.dt = gr->dt;
    // tetah dt/dx^2, tetah dt/dy^2 (SuLi-IT/9_grad.f90:44-45)
    in.aux1 = tetah * gr->dt / (gr->dx * gr->dx);
    in.aux2 = tetah * gr->dt / (gr->dy * gr->dy);
    in.it = 1;
    in.dzx = dzx;
    in.dzy = dzy;
    in.dz3 = dz3;
    in.eta1 = eta1;
    in.deltah = deltah;
    in.tetah = tetah;
    in.gz = gz;
    return in;
}

int run_it_graddin_dev(suli_ps_ctx *ctx, const suli_ps_grid *gr0, double tetah, double gz, const double *u,
                       const double *v, const double *w, const double *dzx, const double *dzy, const double *dz3,
                       const double *eta1, const double *deltah, double *prd1, suli_ps_stats *stats)
{
    int rc = check_it(ctx, gr0, tetah, gz);
    if (rc) return rc;
    if (!u || !v || !w || !dzx || !dzy || !dz3 || !eta1 || !deltah || !prd1)
        return fail(ctx, SULI_PS_EINVAL, "every SuLi-IT array must be non-NULL");
    suli_ps_grid gr = *gr0;
    gr.t_press = 1;   // no IPA tables, no IBM mask: the operator is SuLi-IT's
    gr.obst_t = 0;
    release_last(ctx);
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], s));
    const SetupIn in = make_setup_it(&gr, tetah, gz, u, v, w, dzx, dzy, dz3, eta1, deltah, prd1);
    int slot = -1;
    if ((rc = ctx->eng.admit(gr, in, prd1, solve_opts(ctx->opts), 0, &slot)))
        return fail(ctx, rc, "%s", ctx->eng.error().c_str());
    ctx->last_slots.push_back(slot);
    if ((rc = ctx->eng.run_all(ctx->opts.timing != 0))) return fail(ctx, rc, "%s", ctx->eng.error().c_str());
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], s));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[1]));
    suli_ps_stats st;
    fill_stats(ctx->eng.slot(slot), &st);
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    st.solve_ms = ms;
    st.spmv_timed = ctx->eng.spmv_samples();
    st.spmv_ms = ctx->eng.spmv_ms_mean();
    st.graph_replays = ctx->eng.graph_replays();
    if (stats) *stats = st;
    warn_capped(ctx, st);
    return st.status;
}

int run_it_posdin_dev(suli_ps_ctx *ctx, const suli_ps_grid *gr, double tetah, double gz, double *u, double *v,
                      double *w, const double *dz3, double *eta1, const double *prd1, double *prd)
{
    int rc = check_it(ctx, gr, tetah, gz);
    if (rc) return rc;
    if (!u || !v || !w || !dz3 || !prd1) return fail(ctx, SULI_PS_EINVAL, "u, v, w, dzz and prd1 must be non-NULL");
    const Geo g = make_geo(gr->nx, gr->ny, gr->nz, gr->ccx0 == 0, gr->ccy0 == 0);
    launch_it_posdin(g, gr->dx, gr->dy, gr->dt, tetah, gz, prd1, dz3, eta1, u, v, w, ctx->stream);
    HIPCHK(ctx, hipGetLastError());
    if (prd && prd != prd1) {
        const size_t nbox = (size_t)(gr->nx + 2) * (gr->ny + 2) * (gr->nz + 2);
        HIPCHK(ctx, hipMemcpyAsync(prd, prd1, nbox * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int run_it_posdin_layers_dev(suli_ps_ctx *ctx, const suli_ps_grid *gr, double tetah, double gz, double *u,
                             double *v, double *w, double *dz, double *dzx, double *dzy, double *dzz,
                             const double *eta0, double *eta1, const double *prd1, double *prd)
{
    int rc = check_it(ctx, gr, tetah, gz);
    if (rc) return rc;
    if (!u || !v || !w || !dz || !dzx || !dzy || !dzz || !eta0 || !eta1 || !prd1)
        return fail(ctx, SULI_PS_EINVAL, "u, v, w, dz, dzx, dzy, dzz, eta0, eta1 and prd1 must be non-NULL");
    const Geo g = make_geo(gr->nx, gr->ny, gr->nz, gr->ccx0 == 0, gr->ccy0 == 0);
    // the surface and its layer, dz_it, then the correction with the new
    // dzz: the reference's order (SuLi-IT/10_pos.f90:84-121)
    launch_it_layers(g, gz, prd1, eta0, eta1, dz, dzx, dzy, dzz, ctx->stream);
    launch_it_posdin(g, gr->dx, gr->dy, gr->dt, tetah, gz, prd1, dzz, nullptr, u, v, w, ctx->stream);
    HIPCHK(ctx, hipGetLastError());
    if (prd && prd != prd1) {
        const size_t nbox = (size_t)(gr->nx + 2) * (gr->ny + 2) * (gr->nz + 2);
        HIPCHK(ctx, hipMemcpyAsync(prd, prd1, nbox * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int check_exp_press(suli_ps_ctx *ctx, const suli_ps_grid *gr, const double *rho, const double *prd1,
                    const double *prd0, const double *dz, const double *dzz, const double *px, const double *py,
                    const double *pz)
{
    std::string e;
    const int rc = validate_shape(gr, e);
    if (rc) return fail(ctx, rc, "%s", e.c_str());
    if (!rho || !prd1 || !prd0 || !px || !py || !pz)
        return fail(ctx, SULI_PS_EINVAL, "rho, prd1, prd0, px, py and pz must be non-NULL");
    if ((dz == nullptr) != (dzz == nullptr)) return fail(ctx, SULI_PS_EINVAL, "dz and dzz go together");
    if (!dz && !(gr->dz > 0)) return fail(ctx, SULI_PS_EINVAL, "dz must be positive");
    return 0;
}

// Staging for the host-pointer API, kept while the grid shape is unchanged.
int ensure_staging(suli_ps_ctx *ctx, int nx, int ny, int nz)
{
    Staging &st = ctx->st;
    if (st.block && st.nx == nx && st.ny == ny && st.nz == nz) return 0;
    if (st.block) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        HIPCHK(ctx, hipFree(st.block));
        st = Staging{};
    }
    const size_t n = (size_t)nx * ny * nz;
    const size_t nu = (size_t)(nx + 3) * (ny + 2) * (nz + 2), nv = (size_t)(nx + 2) * (ny + 3) * (nz + 2),
                 nw = (size_t)(nx + 2) * (ny + 2) * (nz + 3);
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2), ne = (size_t)(nx + 2) * (ny + 2);
    const size_t sizes[18] = {nu, nv, nw, n, nbox,
                              (size_t)(nx + 1) * ny * nz, (size_t)nx * (ny + 1) * nz, (size_t)nx * ny * (nz + 1),
                              (size_t)nz + 2, (size_t)nz + 2, n, nu, nv, nw, ne,
                              (size_t)nx * ny, nbox, ne};
    size_t off[18], total = 0;
    for (int i = 0; i < 18; ++i) {
        off[i] = total;
        total += (sizes[i] * sizeof(double) + 255) / 256 * 256;
    }
    hipError_t e = hipMalloc(&st.block, total);
    if (e != hipSuccess) return fail(ctx, SULI_PS_ENOMEM, "hipMalloc(%zu) failed: %s", total, hipGetErrorString(e));
    double **ptrs[18] = {&st.u,   &st.v,   &st.w,   &st.rho, &st.prd1, &st.obsx,   &st.obsy, &st.obsz, &st.dz,
                         &st.dzz, &st.hs,  &st.dzx, &st.dzy, &st.dz3,  &st.eta1, &st.deltah, &st.box2, &st.eta0};
    for (int i = 0; i < 18; ++i) *ptrs[i] = reinterpret_cast<double *>(static_cast<char *>(st.block) + off[i]);
    st.nx = nx;
    st.ny = ny;
    st.nz = nz;
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
              hipEventCreate(&ctx->ev[0]) == hipSuccess && hipEventCreate(&ctx->ev[1]) == hipSuccess;
    ok = ok && ctx->eng.init(ctx->stream, kMaxSlots) == 0;
    // what suli_ps_graddin_noalloc_dev needs besides the scratch array
    ok = ok && ctx->eng.reserve(kNoAllocEntries) == 0;
    if (!ok) {
        suli_ps_destroy(ctx);
        return nullptr;
    }
    ctx->last_slots.reserve(kMaxSlots);
    // a failed lookup only disables the no-allocation entry point
    ctx->scratch_err = scratch_address(ctx->device, &ctx->scratch);
    apply_engine_opts(ctx);
    return ctx;
}

void suli_ps_destroy(suli_ps_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->pipe.release();   // pinned staging buffers, before their stream goes
    release_last(ctx);
    if (ctx->st.block) (void)hipFree(ctx->st.block);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    ctx->eng.~Engine();
    new (&ctx->eng) Engine();
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
    apply_engine_opts(ctx);
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

int suli_ps_graddin_noalloc_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *u, const double *v,
                                const double *w, const double *rho, const double *obs_lsx, const double *obs_lsy,
                                const double *obs_lsz, const double *dz, const double *dzz, double *prd1,
                                suli_ps_stats *stats)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_problem(ctx, g, u, v, w, rho, obs_lsx, obs_lsy, obs_lsz, dz, dzz, prd1);
    if (rc) return rc;
    if (ctx->scratch_err != hipSuccess || !ctx->scratch)
        return fail(ctx, SULI_PS_EHIP, "no-allocation solve: the scratch array of device %d is unavailable (%s)",
                    ctx->device, hipGetErrorString(ctx->scratch_err));
    int cur = -1;
    HIPCHK(ctx, hipGetDevice(&cur));
    if (cur != ctx->device)
        return fail(ctx, SULI_PS_EINVAL, "no-allocation solve: device %d is current, the context's is %d", cur,
                    ctx->device);
    const bool mixed = ctx->opts.precision == SULI_PS_MIXED;
    const size_t need = problem_bytes(g->nx, g->ny, g->nz, mixed);
    if (need > (size_t)SULI_PS_NOALLOC_MAX_BYTES)
        return fail(ctx, SULI_PS_EINVAL,
                    "no-allocation solve: a %dx%dx%d %s problem needs %zu bytes of solver state; the fixed array "
                    "holds %zu (SULI_PS_NOALLOC_MAX_BYTES)",
                    g->nx, g->ny, g->nz, mixed ? "mixed" : "fp64", need, (size_t)SULI_PS_NOALLOC_MAX_BYTES);
    // the launch tables this solve will fill (the shape Engine::admit picks)
    int kz = 0, vx = 0;
    if (ctx->opts.autotune) ctx->eng.cached_shape(g->nx, g->ny, g->nz, mixed, &kz, &vx);
    const Geo geo = make_geo(g->nx, g->ny, g->nz, g->ccx0 == 0, g->ccy0 == 0, kz > 0 ? std::max(kz, kMinKz) : 0,
                             mixed ? vx : 1);
    const long long entries = (long long)spmv_tiles(geo) + vec_tiles(geo.n);
    if (entries > ctx->eng.reserved_entries())
        return fail(ctx, SULI_PS_EINVAL, "no-allocation solve: %lld launch-table entries needed, %d reserved",
                    entries, ctx->eng.reserved_entries());
    NoAllocScope guard;
    return run_graddin_dev(ctx, g, u, v, w, rho, obs_lsx, obs_lsy, obs_lsz, dz, dzz, prd1, stats, nullptr,
                           ctx->scratch);
}

int suli_ps_posdin_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v, double *w, const double *rho,
                       const double *prd1, const double *dz, const double *dzz, double *prd)
{
    if (!ctx) return SULI_PS_EINVAL;
    return run_posdin_dev(ctx, g, u, v, w, rho, prd1, dz, dzz, prd);
}

int suli_ps_graddin_masked_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *u, const double *v,
                               const double *w, const double *rho, const double *obs_lsx, const double *obs_lsy,
                               const double *obs_lsz, const double *dz, const double *dzz, const uint8_t *solid,
                               double *prd1, suli_ps_stats *stats)
{
    if (!ctx) return SULI_PS_EINVAL;
    return run_graddin_dev(ctx, g, u, v, w, rho, obs_lsx, obs_lsy, obs_lsz, dz, dzz, prd1, stats, solid);
}

int suli_ps_posdin_masked_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v, double *w,
                              const double *rho, const double *prd1, const double *dz, const double *dzz,
                              const uint8_t *solid, double *prd)
{
    if (!ctx) return SULI_PS_EINVAL;
    return run_posdin_dev(ctx, g, u, v, w, rho, prd1, dz, dzz, prd, solid);
}

int suli_ps_graddin(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *u, const double *v, const double *w,
                    const double *rho, const double *obs_lsx, const double *obs_lsy, const double *obs_lsz,
                    const double *dz, const double *dzz, double *prd1, suli_ps_stats *stats)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_problem(ctx, g, u, v, w, rho, obs_lsx, obs_lsy, obs_lsz, dz, dzz, prd1);
    if (rc) return rc;
    ctx->st.left.valid = false;
    if ((rc = ensure_staging(ctx, g->nx, g->ny, g->nz))) return rc;
    Staging &st = ctx->st;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t n = (size_t)nx * ny * nz;
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz, nfz = (size_t)nx * ny * (nz + 1);
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const bool mask = g->obst_t != 0 && g->ibm_t == 2;
    const size_t nu = (size_t)(nx + 3) * (ny + 2) * (nz + 2), nv = (size_t)(nx + 2) * (ny + 3) * (nz + 2),
                 nw = (size_t)(nx + 2) * (ny + 2) * (nz + 3);
    int urc = upload(ctx, {seg(st.u, u, nu), seg(st.v, v, nv), seg(st.w, w, nw), seg(st.rho, rho, n),
                           seg(st.prd1, prd1, nbox), seg(st.obsx, obs_lsx, mask ? nfx : 0),
                           seg(st.obsy, obs_lsy, mask ? nfy : 0), seg(st.obsz, obs_lsz, mask ? nfz : 0),
                           seg(st.dz, dz, nz + 2), seg(st.dzz, dzz, nz + 2)});
    if (urc) return urc;
    rc = run_graddin_dev(ctx, g, st.u, st.v, st.w, rho ? st.rho : nullptr, st.obsx, st.obsy, st.obsz,
                         dz ? st.dz : nullptr, dz ? st.dzz : nullptr, st.prd1, stats);
    if (rc < 0) return rc;
    if ((urc = download(ctx, {seg(st.prd1, prd1, nbox)}))) return urc;
    if (rho) {
        Staging::Left &l = st.left;
        l.grid = *g;
        l.u = u;
        l.v = v;
        l.w = w;
        l.rho = rho;
        l.prd1 = prd1;
        l.dz = dz;
        l.dzz = dzz;
        l.valid = true;
    }
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
    // the arrays graddin just staged, as SuLi calls posdin right after it
    // (1_principal.f90:102-103), when the caller allows it
    const Staging::Left &l = ctx->st.left;
    const bool reuse = ctx->opts.reuse_staged && l.valid && std::memcmp(&l.grid, g, sizeof *g) == 0 && l.u == u &&
                       l.v == v && l.w == w && l.rho == rho && l.prd1 == prd1 && l.dz == dz && l.dzz == dzz;
    ctx->st.left.valid = false;
    if ((rc = ensure_staging(ctx, g->nx, g->ny, g->nz))) return rc;
    Staging &st = ctx->st;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t n = (size_t)nx * ny * nz;
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t nu = (size_t)(nx + 3) * (ny + 2) * (nz + 2), nv = (size_t)(nx + 2) * (ny + 3) * (nz + 2),
                 nw = (size_t)(nx + 2) * (ny + 2) * (nz + 3);
    if (!reuse && (rc = upload(ctx, {seg(st.u, u, nu), seg(st.v, v, nv), seg(st.w, w, nw), seg(st.rho, rho, n),
                                     seg(st.prd1, prd1, nbox), seg(st.dz, dz, nz + 2), seg(st.dzz, dzz, nz + 2)})))
        return rc;
    rc = run_posdin_dev(ctx, g, st.u, st.v, st.w, st.rho, st.prd1, dz ? st.dz : nullptr, dz ? st.dzz : nullptr,
                        nullptr);
    if (rc) return rc;
    return download_copy_prd(ctx, {seg(st.u, u, nu), seg(st.v, v, nv), seg(st.w, w, nw)}, prd, prd1, nbox);
}

int suli_ps_pressh_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *rho, const double *dz, double gz,
                       double *prd1)
{
    if (!ctx) return SULI_PS_EINVAL;
    return run_pressh_dev(ctx, g, rho, dz, gz, prd1);
}

int suli_ps_pressh(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *rho, const double *dz, double gz,
                   double *prd1)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_columns(ctx, g, dz);
    if (rc) return rc;
    if (!rho || !prd1) return fail(ctx, SULI_PS_EINVAL, "rho and prd1 must be non-NULL");
    ctx->st.left.valid = false;
    if ((rc = ensure_staging(ctx, g->nx, g->ny, g->nz))) return rc;
    Staging &st = ctx->st;
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    const size_t nbox = (size_t)(g->nx + 2) * (g->ny + 2) * (g->nz + 2);
    if ((rc = upload(ctx, {seg(st.rho, rho, n), seg(st.prd1, prd1, nbox), seg(st.dz, dz, g->nz + 2)}))) return rc;
    if ((rc = run_pressh_dev(ctx, g, st.rho, dz ? st.dz : nullptr, gz, st.prd1))) return rc;
    return download(ctx, {seg(st.prd1, prd1, nbox)});
}

int suli_ps_it_graddin_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz, const double *u,
                           const double *v, const double *w, const double *dzx, const double *dzy, const double *dzz,
                           const double *eta1, const double *deltah, double *prd1, suli_ps_stats *stats)
{
    if (!ctx) return SULI_PS_EINVAL;
    return run_it_graddin_dev(ctx, g, tetah, gz, u, v, w, dzx, dzy, dzz, eta1, deltah, prd1, stats);
}

int suli_ps_it_posdin_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz, double *u, double *v,
                          double *w, const double *dzz, double *eta1, const double *prd1, double *prd)
{
    if (!ctx) return SULI_PS_EINVAL;
    return run_it_posdin_dev(ctx, g, tetah, gz, u, v, w, dzz, eta1, prd1, prd);
}

int suli_ps_it_graddin(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz, const double *u,
                       const double *v, const double *w, const double *dzx, const double *dzy, const double *dzz,
                       const double *eta1, const double *deltah, double *prd1, suli_ps_stats *stats)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_it(ctx, g, tetah, gz);
    if (rc) return rc;
    if (!u || !v || !w || !dzx || !dzy || !dzz || !eta1 || !deltah || !prd1)
        return fail(ctx, SULI_PS_EINVAL, "every SuLi-IT array must be non-NULL");
    ctx->st.left.valid = false;
    if ((rc = ensure_staging(ctx, g->nx, g->ny, g->nz))) return rc;
    Staging &st = ctx->st;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t nu = (size_t)(nx + 3) * (ny + 2) * (nz + 2), nv = (size_t)(nx + 2) * (ny + 3) * (nz + 2),
                 nw = (size_t)(nx + 2) * (ny + 2) * (nz + 3);
    int urc = upload(ctx, {seg(st.u, u, nu), seg(st.v, v, nv), seg(st.w, w, nw), seg(st.dzx, dzx, nu),
                           seg(st.dzy, dzy, nv), seg(st.dz3, dzz, nw), seg(st.eta1, eta1, (size_t)(nx + 2) * (ny + 2)),
                           seg(st.deltah, deltah, (size_t)nx * ny), seg(st.prd1, prd1, nbox)});
    if (urc) return urc;
    rc = run_it_graddin_dev(ctx, g, tetah, gz, st.u, st.v, st.w, st.dzx, st.dzy, st.dz3, st.eta1, st.deltah,
                            st.prd1, stats);
    if (rc < 0) return rc;
    if ((urc = download(ctx, {seg(st.prd1, prd1, nbox)}))) return urc;
    return rc;
}

int suli_ps_it_posdin(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz, double *u, double *v,
                      double *w, const double *dzz, double *eta1, const double *prd1, double *prd)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_it(ctx, g, tetah, gz);
    if (rc) return rc;
    if (!u || !v || !w || !dzz || !prd1) return fail(ctx, SULI_PS_EINVAL, "u, v, w, dzz and prd1 must be non-NULL");
    ctx->st.left.valid = false;
    if ((rc = ensure_staging(ctx, g->nx, g->ny, g->nz))) return rc;
    Staging &st = ctx->st;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t nu = (size_t)(nx + 3) * (ny + 2) * (nz + 2), nv = (size_t)(nx + 2) * (ny + 3) * (nz + 2),
                 nw = (size_t)(nx + 2) * (ny + 2) * (nz + 3), ne = (size_t)(nx + 2) * (ny + 2);
    if ((rc = upload(ctx, {seg(st.u, u, nu), seg(st.v, v, nv), seg(st.w, w, nw), seg(st.dz3, dzz, nw),
                           seg(st.eta1, eta1, ne), seg(st.prd1, prd1, nbox)})))
        return rc;
    if ((rc = run_it_posdin_dev(ctx, g, tetah, gz, st.u, st.v, st.w, st.dz3, eta1 ? st.eta1 : nullptr, st.prd1,
                                nullptr)))
        return rc;
    return download_copy_prd(ctx, {seg(st.u, u, nu), seg(st.v, v, nv), seg(st.w, w, nw), seg(st.eta1, eta1, ne)},
                             prd, prd1, nbox);
}

int suli_ps_heaviside_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *ls, double eps, double rho_f2,
                          double mu_f1, double mu_f2, double *hs, double *rho, double *mu)
{
    if (!ctx) return SULI_PS_EINVAL;
    std::string e;
    const int rc = validate_shape(g, e);
    if (rc) return fail(ctx, rc, "%s", e.c_str());
    if (!ls || !rho) return fail(ctx, SULI_PS_EINVAL, "ls and rho must be non-NULL");
    if (!(eps > 0)) return fail(ctx, SULI_PS_EINVAL, "eps = alpha1*dx1 must be positive");
    launch_heaviside((long long)g->nx * g->ny * g->nz, ls, eps, g->rho_f1, rho_f2, mu_f1, mu_f2, hs, rho, mu,
                     ctx->stream);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int suli_ps_heaviside_tens_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *ls, double eps,
                               double rho_f2, double mu_f1, double mu_f2, double *hs, double *rho, double *mu,
                               double *hsx, double *hsy, double *hsz)
{
    if (!ctx) return SULI_PS_EINVAL;
    std::string e;
    const int rc = validate_shape(g, e);
    if (rc) return fail(ctx, rc, "%s", e.c_str());
    if (g->nx < 4 || g->ny < 4 || g->nz < 4)
        return fail(ctx, SULI_PS_EINVAL, "grid %dx%dx%d: the WENO5 stencils need 4 cells in every direction",
                    g->nx, g->ny, g->nz);
    if (!(g->dz > 0)) return fail(ctx, SULI_PS_EINVAL, "dz must be positive (a uniform grid)");
    if (!ls || !rho || !hsx || !hsy || !hsz) return fail(ctx, SULI_PS_EINVAL, "ls, rho, hsx, hsy and hsz must be non-NULL");
    if (!(eps > 0)) return fail(ctx, SULI_PS_EINVAL, "eps = alpha1*dx1 must be positive");
    const Geo geo = make_geo(g->nx, g->ny, g->nz, g->ccx0 == 0, g->ccy0 == 0);
    launch_heaviside_tens(geo, ls, eps, g->rho_f1, rho_f2, mu_f1, mu_f2, g->dx, g->dy, g->dz, hs, rho, mu, hsx, hsy,
                          hsz, ctx->stream);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int suli_ps_exp_press_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *rho, const double *prd1,
                          double *prd0, const double *dz, const double *dzz, double *px, double *py, double *pz)
{
    if (!ctx) return SULI_PS_EINVAL;
    const int rc = check_exp_press(ctx, g, rho, prd1, prd0, dz, dzz, px, py, pz);
    if (rc) return rc;
    const Geo geo = make_geo(g->nx, g->ny, g->nz, g->ccx0 == 0, g->ccy0 == 0);
    launch_exp_press(geo, rho, dz, dzz, g->dx, g->dy, g->dz, prd1, prd0, px, py, pz, ctx->stream);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int suli_ps_exp_press(suli_ps_ctx *ctx, const suli_ps_grid *g, const double *rho, const double *prd1, double *prd0,
                      const double *dz, const double *dzz, double *px, double *py, double *pz)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_exp_press(ctx, g, rho, prd1, prd0, dz, dzz, px, py, pz);
    if (rc) return rc;
    ctx->st.left.valid = false;
    if ((rc = ensure_staging(ctx, g->nx, g->ny, g->nz))) return rc;
    Staging &st = ctx->st;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t n = (size_t)nx * ny * nz, nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t nfx = (size_t)(nx + 1) * ny * nz, nfy = (size_t)nx * (ny + 1) * nz, nfz = (size_t)nx * ny * (nz + 1);
    if ((rc = upload(ctx, {seg(st.rho, rho, n), seg(st.prd1, prd1, nbox), seg(st.box2, prd0, nbox),
                           seg(st.dz, dz, nz + 2), seg(st.dzz, dzz, nz + 2)})))
        return rc;
    if ((rc = suli_ps_exp_press_dev(ctx, g, st.rho, st.prd1, st.box2, dz ? st.dz : nullptr, dz ? st.dzz : nullptr,
                                    st.obsx, st.obsy, st.obsz)))
        return rc;
    return download(ctx, {seg(st.box2, prd0, nbox), seg(st.obsx, px, nfx), seg(st.obsy, py, nfy),
                          seg(st.obsz, pz, nfz)});
}

int suli_ps_it_posdin_layers_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz, double *u,
                                 double *v, double *w, double *dz, double *dzx, double *dzy, double *dzz,
                                 const double *eta0, double *eta1, const double *prd1, double *prd)
{
    if (!ctx) return SULI_PS_EINVAL;
    return run_it_posdin_layers_dev(ctx, g, tetah, gz, u, v, w, dz, dzx, dzy, dzz, eta0, eta1, prd1, prd);
}

int suli_ps_it_posdin_layers(suli_ps_ctx *ctx, const suli_ps_grid *g, double tetah, double gz, double *u, double *v,
                             double *w, double *dz, double *dzx, double *dzy, double *dzz, const double *eta0,
                             double *eta1, const double *prd1, double *prd)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_it(ctx, g, tetah, gz);
    if (rc) return rc;
    if (!u || !v || !w || !dz || !dzx || !dzy || !dzz || !eta0 || !eta1 || !prd1)
        return fail(ctx, SULI_PS_EINVAL, "u, v, w, dz, dzx, dzy, dzz, eta0, eta1 and prd1 must be non-NULL");
    ctx->st.left.valid = false;
    if ((rc = ensure_staging(ctx, g->nx, g->ny, g->nz))) return rc;
    Staging &st = ctx->st;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t nu = (size_t)(nx + 3) * (ny + 2) * (nz + 2), nv = (size_t)(nx + 2) * (ny + 3) * (nz + 2),
                 nw = (size_t)(nx + 2) * (ny + 2) * (nz + 3), ne = (size_t)(nx + 2) * (ny + 2);
    // what the call updates, both ways, and what it only reads
    const std::initializer_list<HostSeg> io = {seg(st.u, u, nu),     seg(st.v, v, nv),     seg(st.w, w, nw),
                                               seg(st.box2, dz, nbox), seg(st.dzx, dzx, nu), seg(st.dzy, dzy, nv),
                                               seg(st.dz3, dzz, nw),  seg(st.eta1, eta1, ne)};
    if ((rc = upload(ctx, io)) || (rc = upload(ctx, {seg(st.eta0, eta0, ne), seg(st.prd1, prd1, nbox)})))
        return rc;
    if ((rc = run_it_posdin_layers_dev(ctx, g, tetah, gz, st.u, st.v, st.w, st.box2, st.dzx, st.dzy, st.dz3,
                                       st.eta0, st.eta1, st.prd1, nullptr)))
        return rc;
    return download_copy_prd(ctx, io, prd, prd1, nbox);
}

int suli_ps_posdin_hydro_dev(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v, double *w,
                             const double *rho, const double *hs, const double *prd1, const double *dz, double *prd)
{
    if (!ctx) return SULI_PS_EINVAL;
    return run_posdin_hydro_dev(ctx, g, u, v, w, rho, hs, prd1, dz, prd);
}

int suli_ps_posdin_hydro(suli_ps_ctx *ctx, const suli_ps_grid *g, double *u, double *v, double *w,
                         const double *rho, const double *hs, const double *prd1, const double *dz, double *prd)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_columns(ctx, g, dz);
    if (rc) return rc;
    if (!u || !v || !w || !rho || !hs || !prd1)
        return fail(ctx, SULI_PS_EINVAL, "u, v, w, rho, hs and prd1 must be non-NULL");
    ctx->st.left.valid = false;
    if ((rc = ensure_staging(ctx, g->nx, g->ny, g->nz))) return rc;
    Staging &st = ctx->st;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t n = (size_t)nx * ny * nz;
    const size_t nbox = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    const size_t nu = (size_t)(nx + 3) * (ny + 2) * (nz + 2), nv = (size_t)(nx + 2) * (ny + 3) * (nz + 2),
                 nw = (size_t)(nx + 2) * (ny + 2) * (nz + 3);
    if ((rc = upload(ctx, {seg(st.u, u, nu), seg(st.v, v, nv), seg(st.w, w, nw), seg(st.rho, rho, n),
                           seg(st.hs, hs, n), seg(st.prd1, prd1, nbox), seg(st.dz, dz, nz + 2)})))
        return rc;
    rc = run_posdin_hydro_dev(ctx, g, st.u, st.v, st.w, st.rho, st.hs, st.prd1, dz ? st.dz : nullptr, nullptr);
    if (rc) return rc;
    return download_copy_prd(ctx, {seg(st.u, u, nu), seg(st.v, v, nv), seg(st.w, w, nw)}, prd, prd1, nbox);
}

int suli_ps_graddin_batch_dev(suli_ps_ctx *ctx, suli_ps_problem *probs, int32_t count)
{
    if (!ctx || (!probs && count > 0) || count < 0) return SULI_PS_EINVAL;
    if (count > kMaxSlots) return fail(ctx, SULI_PS_EINVAL, "batch of %d exceeds %d problems", count, kMaxSlots);
    int rc;
    for (int i = 0; i < count; ++i) {
        const suli_ps_problem &p = probs[i];
        if ((rc = check_problem(ctx, &p.grid, p.u, p.v, p.w, p.rho, p.obs_lsx, p.obs_lsy, p.obs_lsz, p.dz, p.dzz,
                                p.prd1)))
            return fail(ctx, rc, "problem %d: %s", i, ctx->err.c_str());
    }
    release_last(ctx);
    hipStream_t s = ctx->stream;
    HIPCHK(ctx, hipEventRecord(ctx->ev[0], s));
    const SolveOpts so = solve_opts(ctx->opts);
    std::vector<int> slots(count, -1);
    for (int i = 0; i < count; ++i) {
        const suli_ps_problem &p = probs[i];
        const SetupIn in = make_setup(&p.grid, p.u, p.v, p.w, p.rho, p.prd1, p.obs_lsx, p.obs_lsy, p.obs_lsz, p.dz,
                                      p.dzz, p.solid);
        if ((rc = ctx->eng.admit(p.grid, in, p.prd1, so, i, &slots[i]))) {
            release_last(ctx);
            for (int s2 : slots)
                if (s2 >= 0) ctx->eng.release(s2);
            return fail(ctx, rc, "problem %d: %s", i, ctx->eng.error().c_str());
        }
        ctx->last_slots.push_back(slots[i]);
    }
    if ((rc = ctx->eng.run_all(false))) return fail(ctx, rc, "%s", ctx->eng.error().c_str());
    HIPCHK(ctx, hipEventRecord(ctx->ev[1], s));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev[1]));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    int worst = SULI_PS_OK;
    for (int i = 0; i < count; ++i) {
        fill_stats(ctx->eng.slot(slots[i]), &probs[i].stats);
        probs[i].stats.solve_ms = ms;
        probs[i].stats.graph_replays = ctx->eng.graph_replays();
        worst = std::max(worst, probs[i].stats.status);
    }
    return worst;
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

// Staged through the context's pinned chunks above a few MiB, as the
// host-pointer entry points are; both return once the copy is complete.
int suli_ps_memcpy_h2d(suli_ps_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return SULI_PS_EINVAL;
    const HostSeg sg{dst, const_cast<void *>(src), bytes};
    HIPCHK(ctx, ctx->pipe.upload(ctx->device, ctx->stream, &sg, 1));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return 0;
}

int suli_ps_memcpy_d2h(suli_ps_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return SULI_PS_EINVAL;
    const HostSeg sg{const_cast<void *>(src), dst, bytes};
    HIPCHK(ctx, ctx->pipe.download(ctx->device, ctx->stream, &sg, 1));
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
    int rc = ctx->eng.time_spmv(precision == SULI_PS_MIXED, reps, ms);
    if (rc) return fail(ctx, rc, "%s", ctx->eng.error().c_str());
    return 0;
}

int suli_ps_autotune(suli_ps_ctx *ctx, const suli_ps_grid *g, int32_t precision, int32_t *kz, int32_t *vx,
                     double *step_us)
{
    if (!ctx || !kz || !vx || !step_us) return SULI_PS_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    int k = 0, v = 1;
    double us = 0.0;
    if ((rc = ctx->eng.tune(g->nx, g->ny, g->nz, precision == SULI_PS_MIXED, &k, &v, &us)))
        return fail(ctx, rc, "%s", ctx->eng.error().c_str());
    *kz = k;
    *vx = v;
    *step_us = us;
    return 0;
}

int suli_ps_set_tiles(suli_ps_ctx *ctx, const suli_ps_grid *g, int32_t precision, int32_t kz, int32_t vx)
{
    if (!ctx) return SULI_PS_EINVAL;
    int rc = check_grid(ctx, g);
    if (rc) return rc;
    const int w = vx & ~SULI_PS_TILES_PIPELINED;
    if (kz < 1 || (w != 1 && w != 2 && w != 4) || (w == 1 && w != vx))
        return fail(ctx, SULI_PS_EINVAL, "tile shape kz=%d vx=%d", kz, vx);
    ctx->eng.set_tiles(g->nx, g->ny, g->nz, precision == SULI_PS_MIXED, kz, vx);
    return 0;
}

int suli_ps_copy_bench(suli_ps_ctx *ctx, size_t bytes, int32_t reps, double *ms)
{
    if (!ctx || !ms || reps <= 0 || bytes < 16) return SULI_PS_EINVAL;
    const size_t n16 = bytes / 16;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n16 * 16) != hipSuccess) return fail(ctx, SULI_PS_ENOMEM, "hipMalloc(%zu) failed", bytes);
    if (hipMalloc(&b, n16 * 16) != hipSuccess) {
        (void)hipFree(a);
        return fail(ctx, SULI_PS_ENOMEM, "hipMalloc(%zu) failed", bytes);
    }
    hipStream_t s = ctx->stream;
    int rc = 0;
    hipError_t e = hipMemsetAsync(a, 0, n16 * 16, s);
    launch_copy(a, b, (long long)n16, s);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[0], s);
    for (int i = 0; i < reps && e == hipSuccess; ++i) launch_copy(a, b, (long long)n16, s);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], s);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev[1]);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]);
    if (e != hipSuccess) rc = fail(ctx, SULI_PS_EHIP, "copy bench: %s", hipGetErrorString(e));
    *ms = (double)t / reps;
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

int suli_ps_read_bench(suli_ps_ctx *ctx, size_t bytes, int32_t reps, double *ms)
{
    if (!ctx || !ms || reps <= 0 || bytes < 16) return SULI_PS_EINVAL;
    const size_t n16 = bytes / 16;
    void *a = nullptr;
    float *out = nullptr;
    if (hipMalloc(&a, n16 * 16) != hipSuccess) return fail(ctx, SULI_PS_ENOMEM, "hipMalloc(%zu) failed", bytes);
    if (hipMalloc(&out, sizeof(float) * kReadGrid * kThreads) != hipSuccess) {
        (void)hipFree(a);
        return fail(ctx, SULI_PS_ENOMEM, "hipMalloc of the read bench's sums failed");
    }
    hipStream_t s = ctx->stream;
    int rc = 0;
    hipError_t e = hipMemsetAsync(a, 0, n16 * 16, s);
    launch_read(a, out, (long long)n16, s);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[0], s);
    for (int i = 0; i < reps && e == hipSuccess; ++i) launch_read(a, out, (long long)n16, s);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev[1], s);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev[1]);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]);
    if (e != hipSuccess) rc = fail(ctx, SULI_PS_EHIP, "read bench: %s", hipGetErrorString(e));
    *ms = (double)t / reps;
    (void)hipFree(a);
    (void)hipFree(out);
    return rc;
}

}  // extern "C"
