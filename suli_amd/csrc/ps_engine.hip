// Host-side solve engine (see ps_engine.h).
#include "ps_engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace sps {

namespace {

constexpr int kMaxGrid = 2048;     // persistent workgroups (8 per CU)
constexpr int kMaxTimed = 256;     // timing windows per run_all
constexpr int kTimeStride = 16;    // one window per 16 steps ...
constexpr int kTimeSteps = 8;      // ... over steps 16k+1 .. 16k+8 (no replacement inside;
                                    // an even count: the mixed step alternates two kinds)

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Carves one device block into a problem's arrays (256-byte aligned).
//
// Each array starts kSkew bytes past the end of the previous one.  The
// iteration kernels read six to eighteen arrays at the same cell offset;
// when the arrays are back to back and their size is a multiple of a large
// power of two (an FP32 cell array of 256x128x96 is 12 MiB), those reads
// fall on the same DRAM channel and bank at once.  Measured on this GPU: a
// kernel that reads nine 12 MiB FP32 arrays and writes one took 24.2 us with
// the arrays back to back and 18.0 us with them 4352 bytes apart.
constexpr size_t kSkew = 17 * 256;
struct Carver {
    char *base;
    size_t off = 0;
    template <typename T>
    T *take(size_t count)
    {
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off = align_up(off + std::max<size_t>(count, 1) * sizeof(T), 256) + kSkew;
        return p;
    }
};

void carve(Carver &cv, const Geo &g, bool mixed, int nt3, int nt1, ProbDev &pd)
{
    const size_t n = (size_t)g.n;
    const size_t nfx = (size_t)(g.nx + 1) * g.ny * g.nz, nfy = (size_t)g.nx * (g.ny + 1) * g.nz,
                 nfz = (size_t)g.nx * g.ny * (g.nz + 1);
    pd.sx = cv.take<double>(nfx);
    pd.sy = cv.take<double>(nfy);
    pd.sz = cv.take<double>(nfz);
    pd.d = cv.take<double>(n);
    pd.b = cv.take<double>(n);
    pd.e = cv.take<double>(n);
    pd.r = cv.take<double>(n);   // FP64 residual; the mixed solver's setup scratch
    if (!mixed) {
        pd.q = cv.take<double>(n);
        pd.p0 = cv.take<double>(n);
        pd.p1 = cv.take<double>(n);
    } else {
        pd.sx32 = cv.take<float>(nfx);
        pd.sy32 = cv.take<float>(nfy);
        pd.sz32 = cv.take<float>(nfz);
        pd.c32 = cv.take<float>(n);
        pd.r32 = cv.take<float>(n);
        pd.q32 = cv.take<float>(n);
        // the p ring: kFold slots kSkew bytes apart
        pd.pslot = (long long)((align_up(n * sizeof(float), 256) + kSkew) / sizeof(float));
        pd.p32 = cv.take<float>((size_t)pd.pslot * kFold);
        pd.q32b = cv.take<float>(n);
        pd.r32b = cv.take<float>(n);
        pd.P.m4 = cv.take<double>(8 * (size_t)nt3);
    }
    pd.tx = cv.take<double>((size_t)g.nz + 1);
    pd.ty = cv.take<double>((size_t)g.nz + 1);
    pd.tz = cv.take<double>((size_t)g.nz + 1);
    if (mixed) {
        pd.tx32 = cv.take<float>((size_t)g.nz + 1);
        pd.ty32 = cv.take<float>((size_t)g.nz + 1);
        pd.tz32 = cv.take<float>((size_t)g.nz + 1);
        pd.rx32 = cv.take<float>(8 * (size_t)g.nz);
        pd.ry32 = cv.take<float>(4 * (size_t)g.nz);
        pd.rz32 = cv.take<float>(4 * ((size_t)g.nz + 1));
        pd.rowreg = cv.take<unsigned>((size_t)(g.ny + 31) / 32 * g.nz);
    }
    pd.P.spmv = cv.take<double>(nt3);
    pd.P.upd = cv.take<double>(nt1);
    pd.P.tres = cv.take<double>(3 * (size_t)nt3);
    pd.P.init = cv.take<double>(5 * (size_t)init_blocks(g));
    pd.P.proj = cv.take<double>(2 * (size_t)init_blocks(g));
}

}  // namespace

// ------------------------------------------------------ allocation guard ----
namespace {
thread_local bool t_no_alloc = false;
}

bool no_alloc_active() { return t_no_alloc; }
NoAllocScope::NoAllocScope() : prev_(t_no_alloc) { t_no_alloc = true; }
NoAllocScope::~NoAllocScope() { t_no_alloc = prev_; }

// ---------------------------------------------------------------- Pool ----
Pool::~Pool() { release_all(); }

void *Pool::get(size_t bytes, std::string &err)
{
    for (size_t i = 0; i < free_.size(); ++i)
        if (free_[i].first == bytes) {
            void *p = free_[i].second;
            free_[i] = free_.back();
            free_.pop_back();
            return p;
        }
    if (no_alloc_active()) {
        err = "no-allocation call: the pool holds no free block of this size";
        return nullptr;
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        // give back cached blocks of other sizes and retry once
        for (auto &kv : free_) (void)hipFree(kv.second), held_ -= kv.first;
        free_.clear();
        e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
            return nullptr;
        }
    }
    held_ += bytes;
    return p;
}

void Pool::put(void *p, size_t bytes)
{
    if (p) free_.emplace_back(bytes, p);
}

void Pool::release_all()
{
    for (auto &kv : free_) (void)hipFree(kv.second);
    free_.clear();
    held_ = 0;
}

size_t problem_bytes(int nx, int ny, int nz, bool mixed)
{
    // the partial-sum arrays sized for the most tiles (the thinnest slabs);
    // everything else does not depend on the tiling
    const Geo g = make_geo(nx, ny, nz, 0, 0, kMinKz);
    Carver cv{nullptr};
    ProbDev pd{};
    carve(cv, g, mixed, spmv_tiles(g), vec_tiles(g.n), pd);
    return align_up(cv.off, 2 << 20);
}

// -------------------------------------------------------------- Engine ----
#define ECHK(call)                                                                                     \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return fail(SULI_PS_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int Engine::fail(int code, const std::string &msg)
{
    err_ = msg;
    return code;
}

// Fails an allocating path while a NoAllocScope is active.
#define NO_ALLOC_HERE(what)                                                                            \
    do {                                                                                               \
        if (no_alloc_active()) return fail(SULI_PS_ENOMEM, std::string("no-allocation call would ") + (what)); \
    } while (0)

Engine::~Engine()
{
    if (s_) (void)hipStreamSynchronize(s_);
    for (auto &st : side_)
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    if (fork_ev_) (void)hipEventDestroy(fork_ev_);
    for (auto &e : join_ev_)
        if (e) (void)hipEventDestroy(e);
    for (auto &sl : slots_)
        if (sl.block && sl.own_block) (void)hipFree(sl.block);
    pool_.release_all();
    if (d_probs_) (void)hipFree(d_probs_);
    if (d_status_) (void)hipFree(d_status_);
    if (h_status_) (void)hipHostFree(h_status_);
    if (h_final_) (void)hipHostFree(h_final_);
    if (d_fin_) (void)hipFree(d_fin_);
    for (auto &e : ev_status_)
        if (e) (void)hipEventDestroy(e);
    if (d_tab_) (void)hipFree(d_tab_);
    if (d_probs_list_) (void)hipFree(d_probs_list_);
    for (int i = 0; i < 4; ++i) {
        if (h_stage_[i]) (void)hipHostFree(h_stage_[i]);
        if (ev_stage_[i]) (void)hipEventDestroy(ev_stage_[i]);
    }
    for (auto &e : ev_spmv_)
        if (e) (void)hipEventDestroy(e);
    if (d_ran_) (void)hipFree(d_ran_);
    retire_graphs();
    reap_graphs(true);
    if (d_step_) (void)hipFree(d_step_);
}

int Engine::init(hipStream_t stream, int max_slots)
{
    s_ = stream;
    max_slots_ = max_slots;
    slots_.assign(max_slots, Slot{});
    ECHK(hipMalloc(&d_probs_, sizeof(ProbDev) * max_slots));
    ECHK(hipMalloc(&d_status_, sizeof(CgScalars) * max_slots));
    ECHK(hipMemset(d_status_, 0, sizeof(CgScalars) * max_slots));
    ECHK(hipHostMalloc((void **)&h_status_, sizeof(CgScalars) * 2 * max_slots, hipHostMallocDefault));
    ECHK(hipHostMalloc((void **)&h_final_, sizeof(CgScalars) * max_slots, hipHostMallocDefault));
    for (auto &e : ev_status_) ECHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (auto &e : ev_stage_) ECHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ev_spmv_.resize(2 * kMaxTimed);
    for (auto &e : ev_spmv_) ECHK(hipEventCreate(&e));
    ECHK(hipMalloc(&d_ran_, sizeof(int) * 2 * kMaxTimed));
    ECHK(hipMalloc(&d_step_, 2 * sizeof(int)));
    ECHK(hipMemset(d_step_, 0, 2 * sizeof(int)));
    dstep_valid_ = true;
    // default priority, as the engine's stream
    for (auto &st : side_) ECHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    ECHK(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
    for (auto &e : join_ev_) ECHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return 0;
}

int Engine::reserve(int entries)
{
    if (entries <= reserved_) return 0;
    NO_ALLOC_HERE("reserve tables");
    ECHK(hipStreamSynchronize(s_));
    const size_t ntab = (size_t)entries;
    if (ntab > tab_cap_ || (size_t)max_slots_ > probs_cap_ || !d_tab_) {
        if (d_tab_) ECHK(hipFree(d_tab_));
        if (d_probs_list_) ECHK(hipFree(d_probs_list_));
        d_tab_ = nullptr;
        d_probs_list_ = nullptr;
        tab_cap_ = std::max(tab_cap_, ntab);
        probs_cap_ = std::max(probs_cap_, (size_t)max_slots_);
        ECHK(hipMalloc(&d_tab_, sizeof(TileRef) * tab_cap_));
        ECHK(hipMalloc(&d_probs_list_, sizeof(int) * probs_cap_));
    }
    // finish tables: every 3-D tile plus the problem list
    const size_t fin = align_up(sizeof(TileRef) * ntab, 256) + sizeof(int) * max_slots_;
    if (fin > fin_cap_) {
        if (d_fin_) ECHK(hipFree(d_fin_));
        d_fin_ = nullptr;
        fin_cap_ = fin;
        ECHK(hipMalloc(&d_fin_, fin_cap_));
    }
    const size_t stage = std::max({sizeof(TileRef) * ntab + sizeof(int) * max_slots_ + 64, fin,
                                   sizeof(CgScalars) + sizeof(ProbDev), (size_t)65536});
    if (stage > stage_cap_) {
        for (int i = 0; i < 4; ++i) {
            if (h_stage_[i]) ECHK(hipHostFree(h_stage_[i]));
            h_stage_[i] = nullptr;
            ECHK(hipHostMalloc(&h_stage_[i], stage, hipHostMallocDefault));
        }
        stage_cap_ = stage;
    }
    tab_.t3_64.reserve(ntab);
    tab_.t1_64.reserve(ntab);
    tab_.t3_32.reserve(ntab);
    tab_.t1_32.reserve(ntab);
    tab_.mixed_probs.reserve(max_slots_);
    pool_.reserve(max_slots_);
    fin_t3_.reserve(ntab);
    run_slots_.reserve(max_slots_);
    run_finished_.reserve(max_slots_);
    done_.reserve(max_slots_);
    ran_.reserve(2 * kMaxTimed);
    reserved_ = entries;
    tables_dirty_ = true;
    return 0;
}

void Engine::cached_shape(int nx, int ny, int nz, bool mixed, int *kz, int *vx) const
{
    auto hit = tuned_.find(std::make_tuple(nx, ny, nz, mixed));
    *kz = hit != tuned_.end() ? hit->second.kz : 0;
    *vx = hit != tuned_.end() ? hit->second.vx : 0;
}

int Engine::active_count() const
{
    int n = 0;
    for (const auto &s : slots_) n += s.active && !s.finished;
    return n;
}

int Engine::admit(const suli_ps_grid &gr, const SetupIn &in, double *prd1_out, const SolveOpts &o, int user,
                  int *slot_out, void *block, hipEvent_t wait_ev)
{
    int si = -1;
    for (int i = 0; i < max_slots_; ++i)
        if (!slots_[i].used) {
            si = i;
            break;
        }
    if (si < 0) return fail(SULI_PS_ENOMEM, "no free problem slot");
    if (step_ % kRefreshGrain != 0) return fail(SULI_PS_EINVAL, "engine: admission off the refresh grain");
    Slot &sl = slots_[si];
    sl = Slot{};
    int kz = o.kz > 0 ? std::max(o.kz, kMinKz) : 0;
    int vx = o.vx > 0 ? o.vx : 0;
    if ((kz == 0 || vx == 0) && autotune_ && no_alloc_active()) {
        // tuning allocates a scratch problem: take what is recorded instead
        int kz_c = 0, vx_c = 0;
        cached_shape(gr.nx, gr.ny, gr.nz, o.mixed, &kz_c, &vx_c);
        if (kz == 0) kz = kz_c;
        if (vx == 0) vx = vx_c;
    } else if ((kz == 0 || vx == 0) && autotune_) {
        int kz_t = 0, vx_t = 1;
        double us = 0.0;
        int rc_t = tune(gr.nx, gr.ny, gr.nz, o.mixed, &kz_t, &vx_t, &us);
        if (rc_t) return rc_t;
        if (kz == 0) kz = kz_t;
        if (vx == 0) vx = vx_t;
    }
    sl.g = make_geo(gr.nx, gr.ny, gr.nz, gr.ccx0 == 0, gr.ccy0 == 0, kz, o.mixed ? vx : 1);
    sl.g.dtop = gr.top_bc == 1 ? 1 : 0;
    sl.o = o;
    sl.o.kz = sl.g.kz;
    sl.o.vx = lane_code(sl.g);
    const int nt3 = spmv_tiles(sl.g), nt1 = vec_tiles(sl.g.n);
    sl.bytes = problem_bytes(gr.nx, gr.ny, gr.nz, o.mixed);
    sl.own_block = block == nullptr;
    sl.block = block ? block : pool_.get(sl.bytes, err_);
    if (!sl.block) return SULI_PS_ENOMEM;
    Carver cv{static_cast<char *>(sl.block)};
    ProbDev &pd = sl.pd;
    pd = ProbDev{};
    pd.g = sl.g;
    carve(cv, sl.g, o.mixed, nt3, nt1, pd);
    // constant density (IPA, 9_grad.f90:109-116) and no solid mask: the
    // iteration kernels read per-plane tables instead of the face arrays
    if (gr.t_press != 2 || in.solid) {
        pd.tx = pd.ty = pd.tz = nullptr;
        pd.tx32 = pd.ty32 = pd.tz32 = nullptr;
    } else {
        pd.rx32 = pd.ry32 = pd.rz32 = nullptr;
        pd.rowreg = nullptr;
    }
    pd.nt3 = nt3;
    pd.nt1 = nt1;
    pd.mixed = o.mixed ? 1 : 0;
    pd.store_r = o.mixed && o.store_r ? 1 : 0;
    pd.S = d_status_ + si;
    sl.n0 = step_ + 1;   // iteration 1 runs at the next global step
    pd.n0 = sl.n0;
    sl.prd1_out = prd1_out;
    sl.solid = in.solid;
    sl.user = user;
    sl.used = true;
    sl.active = true;

    // initial scalars and the descriptor, staged through pinned memory
    CgScalars s0{};
    s0.max_iter = o.max_iter > 0 ? o.max_iter : 10000;
    s0.mixed = pd.mixed;
    // a Dirichlet lid, or SuLi-IT's free-surface row, leaves the operator
    // without a null space
    s0.project = (o.keep_null || sl.g.dtop || in.it) ? 0 : 1;
    s0.rel_rule = o.stop_rule == SULI_PS_STOP_RELATIVE;
    s0.tol = o.tol;
    // threshold of the reference stop rule: 0.0001/(nx*ny*nz) (9_grad.f90:393)
    s0.thresh_ref = 0.0001 / (double)(gr.nx * gr.ny * gr.nz);
    if (in.it) {
        // SuLi-IT: |r.r| > 1e-11, tested on the previous iteration's r.r
        // (SuLi-IT/9_grad.f90:165)
        s0.thresh_ref = 0.00000000001;
        s0.lag_stop = 1;
    }
    s0.done_at = kNotDone;
    s0.skip_e_at = -1;
    s0.next_refresh = kFirstRefresh;
    sl.status = s0;
    const size_t need = sizeof(CgScalars) + sizeof(ProbDev);
    int stg;
    int rc = 0;
    {
        // a staging buffer whose previous copy has completed
        if (stage_cap_ < need) {
            NO_ALLOC_HERE("grow the staging ring");
            ECHK(hipStreamSynchronize(s_));
            for (int i = 0; i < 4; ++i) {
                if (h_stage_[i]) ECHK(hipHostFree(h_stage_[i]));
                ECHK(hipHostMalloc(&h_stage_[i], std::max(need, (size_t)65536), hipHostMallocDefault));
            }
            stage_cap_ = std::max(need, (size_t)65536);
        }
        stg = stage_next_;
        stage_next_ = (stage_next_ + 1) & 3;
        ECHK(hipEventSynchronize(ev_stage_[stg]));
        char *h = static_cast<char *>(h_stage_[stg]);
        std::memcpy(h, &s0, sizeof s0);
        std::memcpy(h + sizeof s0, &pd, sizeof pd);
        ECHK(hipMemcpyAsync(pd.S, h, sizeof s0, hipMemcpyHostToDevice, s_));
        ECHK(hipMemcpyAsync(d_probs_ + si, h + sizeof s0, sizeof pd, hipMemcpyHostToDevice, s_));
        ECHK(hipEventRecord(ev_stage_[stg], s_));
    }

    // setup kernels (ps_setup.hip), after the inputs have arrived
    if (wait_ev) ECHK(hipStreamWaitEvent(s_, wait_ev, 0));
    double *sx = const_cast<double *>(pd.sx), *sy = const_cast<double *>(pd.sy), *sz = const_cast<double *>(pd.sz);
    launch_face_coefs(sl.g, in, sx, sy, sz, s_);
    launch_diag_rhs(sl.g, in, sx, sy, sz, const_cast<double *>(pd.d), pd.b, pd.e, s_);
    launch_scale_faces(sl.g, pd.d, sx, sy, sz, const_cast<float *>(pd.sx32), const_cast<float *>(pd.sy32),
                       const_cast<float *>(pd.sz32), s_);
    if (pd.tx)
        launch_face_tables(sl.g, pd.sx, pd.sy, pd.sz, pd.sx32, pd.sy32, pd.sz32, const_cast<double *>(pd.tx),
                           const_cast<double *>(pd.ty), const_cast<double *>(pd.tz), const_cast<float *>(pd.tx32),
                           const_cast<float *>(pd.ty32), const_cast<float *>(pd.tz32), s_);
    else if (pd.rowreg)
        launch_row_tables(sl.g, pd.sx32, pd.sy32, pd.sz32, const_cast<float *>(pd.rx32), const_cast<float *>(pd.ry32),
                          const_cast<float *>(pd.rz32), const_cast<unsigned *>(pd.rowreg), s_);
    launch_init(sl.g, in, Op<double>{pd.sx, pd.sy, pd.sz}, pd.d, pd.b, pd.e, pd.r, o.mixed ? nullptr : pd.r,
                pd.r32, pd.c32, pd.P, init_blocks(sl.g), pd.S, s_);
    ECHK(hipGetLastError());
    tables_dirty_ = true;
    if (slot_out) *slot_out = si;
    return rc;
}

void Engine::release(int si)
{
    Slot &sl = slots_[si];
    if (!sl.used) return;
    if (sl.own_block) pool_.put(sl.block, sl.bytes);
    sl = Slot{};
    tables_dirty_ = true;
}

int Engine::rebuild_tables()
{
    tab_.t3_64.clear();
    tab_.t1_64.clear();
    tab_.t3_32.clear();
    tab_.t1_32.clear();
    tab_.mixed_probs.clear();
    table_cells_ = 0;
    for (int i = 0; i < max_slots_; ++i) {
        const Slot &sl = slots_[i];
        if (!sl.active || sl.finished) continue;
        table_cells_ += sl.g.n;
        auto &t1 = sl.o.mixed ? tab_.t1_32 : tab_.t1_64;
        if (!sl.o.mixed)
            for (int t = 0; t < sl.pd.nt3; ++t) tab_.t3_64.push_back({i, t});
        for (int t = 0; t < sl.pd.nt1; ++t) t1.push_back({i, t});
        if (sl.o.mixed) tab_.mixed_probs.push_back(i);
    }
    // FP32 3-D entries grouped by lane kind (one step kernel per kind)
    for (int w = 0; w < kLaneKinds; ++w) {
        sub_t3_32[w] = (int)tab_.t3_32.size();
        for (int i = 0; i < max_slots_; ++i) {
            const Slot &sl = slots_[i];
            if (!sl.active || sl.finished || !sl.o.mixed || lane_kind(sl.g) != w) continue;
            for (int t = 0; t < sl.pd.nt3; ++t) tab_.t3_32.push_back({i, t});
        }
        cnt_t3_32[w] = (int)tab_.t3_32.size() - sub_t3_32[w];
    }
    off_t3_64 = 0;
    off_t1_64 = off_t3_64 + (int)tab_.t3_64.size();
    off_t3_32 = off_t1_64 + (int)tab_.t1_64.size();
    off_t1_32 = off_t3_32 + (int)tab_.t3_32.size();
    const size_t ntab = off_t1_32 + tab_.t1_32.size();
    const size_t nprobs = tab_.mixed_probs.size();
    if (ntab > tab_cap_ || nprobs > probs_cap_ || !d_tab_) {
        NO_ALLOC_HERE("grow the launch tables");
        // grow: buffers still read by queued kernels are freed after a sync
        ECHK(hipStreamSynchronize(s_));
        if (d_tab_) ECHK(hipFree(d_tab_));
        if (d_probs_list_) ECHK(hipFree(d_probs_list_));
        tab_cap_ = std::max<size_t>(ntab, 1) * 2;
        probs_cap_ = std::max<size_t>(nprobs, (size_t)max_slots_);
        ECHK(hipMalloc(&d_tab_, sizeof(TileRef) * tab_cap_));
        ECHK(hipMalloc(&d_probs_list_, sizeof(int) * probs_cap_));
    }
    const size_t need = sizeof(TileRef) * ntab + sizeof(int) * nprobs + 64;
    if (need > stage_cap_) {
        NO_ALLOC_HERE("grow the staging ring");
        ECHK(hipStreamSynchronize(s_));
        for (int i = 0; i < 4; ++i) {
            if (h_stage_[i]) ECHK(hipHostFree(h_stage_[i]));
            ECHK(hipHostMalloc(&h_stage_[i], need * 2, hipHostMallocDefault));
        }
        stage_cap_ = need * 2;
    }
    const int stg = stage_next_;
    stage_next_ = (stage_next_ + 1) & 3;
    ECHK(hipEventSynchronize(ev_stage_[stg]));
    char *h = static_cast<char *>(h_stage_[stg]);
    TileRef *ht = reinterpret_cast<TileRef *>(h);
    std::copy(tab_.t3_64.begin(), tab_.t3_64.end(), ht + off_t3_64);
    std::copy(tab_.t1_64.begin(), tab_.t1_64.end(), ht + off_t1_64);
    std::copy(tab_.t3_32.begin(), tab_.t3_32.end(), ht + off_t3_32);
    std::copy(tab_.t1_32.begin(), tab_.t1_32.end(), ht + off_t1_32);
    int *hp = reinterpret_cast<int *>(h + align_up(sizeof(TileRef) * ntab, 64));
    std::copy(tab_.mixed_probs.begin(), tab_.mixed_probs.end(), hp);
    if (ntab) ECHK(hipMemcpyAsync(d_tab_, ht, sizeof(TileRef) * ntab, hipMemcpyHostToDevice, s_));
    if (nprobs) ECHK(hipMemcpyAsync(d_probs_list_, hp, sizeof(int) * nprobs, hipMemcpyHostToDevice, s_));
    ECHK(hipEventRecord(ev_stage_[stg], s_));
    tables_dirty_ = false;
    graph_valid_ = false;   // captured launches name the old table sizes
    return 0;
}

// Steps of the next chunk (a multiple of kRefreshGrain, so that problems
// admitted between chunks start on the refresh grain).  Mixed problems:
// up to the earliest replacement known to be scheduled, so the host reads
// the state soon after it; FP64 problems decide their stop on the device
// every iteration and only need the host to notice it.
int Engine::next_chunk_len() const
{
    if (chunk_ > 0) return (chunk_ + kRefreshGrain - 1) / kRefreshGrain * kRefreshGrain;
    bool any_mixed = false;
    long long biggest = 0;
    int best = kMaxRefreshGap;
    for (int i = 0; i < max_slots_; ++i) {
        const Slot &sl = slots_[i];
        if (!sl.active || sl.finished) continue;
        biggest = std::max(biggest, sl.g.n);
        if (!sl.o.mixed) continue;
        any_mixed = true;
        const int queued = step_ - sl.n0 + 1;   // iterations already enqueued
        const int ahead = sl.status.next_refresh - queued;
        best = std::min(best, ahead > 0 ? ahead : 2 * kRefreshGrain);
    }
    if (!any_mixed) return biggest <= 200000 ? 64 : 32;
    best = std::max(best, kRefreshGrain);
    return (best + kRefreshGrain - 1) / kRefreshGrain * kRefreshGrain;
}

// Could a mixed problem's residual replacement be due after global step s?
// Its scheduled iteration as last read, any later iteration (the schedule
// read may be stale) or an iteration cap.  Launching when none is due only
// costs time: the kernels test refresh_due on the device.
bool Engine::refresh_may_be_due(int s) const
{
    for (int i = 0; i < max_slots_; ++i) {
        const Slot &sl = slots_[i];
        if (!sl.active || sl.finished || !sl.o.mixed) continue;
        const int n = s - sl.n0 + 1;
        if (n >= sl.status.next_refresh || n > sl.o.max_iter || sl.status.done_at != kNotDone) return true;
    }
    return false;
}

int Engine::launch_steps_m(StepRef sr, int *ran, bool concurrent)
{
    int kinds[kLaneKinds], nk = 0;
    for (int w = 0; w < kLaneKinds; ++w)
        if (cnt_t3_32[w]) kinds[nk++] = w;
    // kind i on stream i mod (nside + 1), 0 being the engine's own
    const int nside = concurrent ? std::min(nk - 1, kSideStreams) : 0;
    if (nside > 0) ECHK(hipEventRecord(fork_ev_, s_));
    for (int i = 0; i < nk; ++i) {
        const int w = kinds[i], q = i % (nside + 1);
        const hipStream_t st = q == 0 ? s_ : side_[q - 1];
        if (q > 0 && i == q) ECHK(hipStreamWaitEvent(st, fork_ev_, 0));
        const int n = cnt_t3_32[w];
        launch_cg_m(d_probs_, d_tab_ + off_t3_32 + sub_t3_32[w], n, std::min(n, kMaxGrid), sr, ran, kind_vx(w),
                    kind_pipe(w), st);
        sr.next = nullptr;   // one launch updates the counter
    }
    for (int q = 1; q <= nside; ++q) {
        ECHK(hipEventRecord(join_ev_[q - 1], side_[q - 1]));
        ECHK(hipStreamWaitEvent(s_, join_ev_[q - 1], 0));
    }
    return 0;
}

// Kernels of global steps step_+1 .. step_+nsteps launched from the host,
// each with its step number as an argument.
int Engine::enqueue_steps_direct(int nsteps, bool timing)
{
    const int n3_64 = (int)tab_.t3_64.size(), n1_64 = (int)tab_.t1_64.size();
    const int n3_32 = (int)tab_.t3_32.size(), n1_32 = (int)tab_.t1_32.size();
    const int nmix = (int)tab_.mixed_probs.size();
    for (int it = 0; it < nsteps; ++it) {
        ++step_;
        const StepRef sr{nullptr, step_, nullptr};
        // timing windows of kTimeSteps consecutive steps (events only at
        // their ends, so the per-step figure is the launches' own time)
        const int ph = step_ % kTimeStride;
        const bool t_open = timing && ph == 1 && ntimed_ < kMaxTimed;
        const bool t_close = timing && ph == kTimeSteps && ntimed_ < kMaxTimed;
        int *ran = t_open ? d_ran_ + 2 * ntimed_ : t_close ? d_ran_ + 2 * ntimed_ + 1 : nullptr;
        if (t_open) ECHK(hipEventRecord(ev_spmv_[2 * ntimed_], s_));
        if (n3_64) launch_cg_a(false, d_probs_, d_tab_ + off_t3_64, n3_64, std::min(n3_64, kMaxGrid), sr, ran, s_);
        if (n3_32) {
            const int rc = launch_steps_m(sr, ran, true);
            if (rc) return rc;
        }
        if (t_close) {
            ECHK(hipEventRecord(ev_spmv_[2 * ntimed_ + 1], s_));
            ++ntimed_;
        }
        if (n1_64) launch_cg_b(false, d_probs_, d_tab_ + off_t1_64, n1_64, std::min(n1_64, kMaxGrid), sr, s_);
        if (nmix && step_ % kRefreshGrain == 0 && refresh_may_be_due(step_))
            launch_refresh(d_probs_, d_tab_ + off_t1_32, n1_32, d_tab_ + off_t3_32, n3_32, d_probs_list_, nmix,
                           kMaxGrid, sr, s_);
    }
    dstep_valid_ = false;   // the device counters no longer follow step_
    ECHK(hipGetLastError());
    return 0;
}

// Launches of the current graphs may still be queued: keep them until the
// stream has passed its present tail.
void Engine::retire_graphs()
{
    for (auto &g : gx_)
        if (g) {
            hipEvent_t ev = nullptr;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess &&
                hipEventRecord(ev, s_) == hipSuccess) {
                retired_.emplace_back(g, ev);
            } else {
                if (ev) (void)hipEventDestroy(ev);
                (void)hipStreamSynchronize(s_);
                (void)hipGraphExecDestroy(g);
            }
            g = nullptr;
        }
    graph_valid_ = false;
}

void Engine::reap_graphs(bool wait)
{
    size_t keep = 0;
    for (size_t i = 0; i < retired_.size(); ++i) {
        auto &r = retired_[i];
        if (wait) (void)hipEventSynchronize(r.second);
        if (wait || hipEventQuery(r.second) == hipSuccess) {
            (void)hipGraphExecDestroy(r.first);
            (void)hipEventDestroy(r.second);
        } else {
            retired_[keep++] = r;
        }
    }
    retired_.resize(keep);
}

// Captures one group of kRefreshGrain steps of the current tables as a
// graph, with and without the residual replacement at its end, for each
// of the two device step counters: a group of parity p reads counter p and
// its first launch stores the next group's step in counter 1 - p.
int Engine::capture_graphs()
{
    NO_ALLOC_HERE("capture a graph");
    std::unique_lock<std::mutex> lk;
    if (capture_mutex_) lk = std::unique_lock<std::mutex>(*capture_mutex_);
    retire_graphs();
    reap_graphs(false);
    const int n3_64 = (int)tab_.t3_64.size(), n1_64 = (int)tab_.t1_64.size();
    const int n3_32 = (int)tab_.t3_32.size(), n1_32 = (int)tab_.t1_32.size();
    const int nmix = (int)tab_.mixed_probs.size();
    for (int gi = 0; gi < 4; ++gi) {
        const int par = gi >> 1, v = gi & 1;
        if (v == 1 && !nmix) continue;
        const int *base = d_step_ + par;
        int *next = d_step_ + (1 - par);
        ECHK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        for (int off = 1; off <= kRefreshGrain; ++off) {
            // the group's first launch (cg_a, else cg_m) carries the counter update
            const StepRef sr{base, off, off == 1 ? next : nullptr};
            const StepRef sr_m{base, off, (off == 1 && !n3_64) ? next : nullptr};
            if (n3_64) launch_cg_a(false, d_probs_, d_tab_ + off_t3_64, n3_64, std::min(n3_64, kMaxGrid), sr, nullptr, s_);
            if (n3_32) (void)launch_steps_m(sr_m, nullptr);   // one stream: only launches
            if (n1_64)
                launch_cg_b(false, d_probs_, d_tab_ + off_t1_64, n1_64, std::min(n1_64, kMaxGrid),
                            StepRef{base, off, nullptr}, s_);
        }
        if (v == 1)
            launch_refresh(d_probs_, d_tab_ + off_t1_32, n1_32, d_tab_ + off_t3_32, n3_32, d_probs_list_, nmix,
                           kMaxGrid, StepRef{base, kRefreshGrain, nullptr}, s_);
        hipGraph_t g = nullptr;
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(s_, &g);
        if (le != hipSuccess || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(SULI_PS_EHIP, std::string("graph capture failed: ") +
                                          hipGetErrorString(le != hipSuccess ? le : ce));
        }
        const hipError_t ie = hipGraphInstantiate(&gx_[gi], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) return fail(SULI_PS_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    }
    graph_valid_ = true;
    ++captures_;
    return 0;
}

// Steps step_+1 .. step_+kRefreshGrain by replaying a captured group.
int Engine::enqueue_group_graph()
{
    int rc;
    if (!graph_valid_ && (rc = capture_graphs())) return rc;
    const int par = (step_ / kRefreshGrain) & 1;
    // counter par holds step_ after a replayed group; after direct launches set it
    if (!dstep_valid_) launch_set_step(d_step_ + par, step_, s_);
    const bool refresh = !tab_.mixed_probs.empty() && refresh_may_be_due(step_ + kRefreshGrain);
    ECHK(hipGraphLaunch(gx_[2 * par + (refresh ? 1 : 0)], s_));
    ++replays_;
    step_ += kRefreshGrain;
    dstep_valid_ = true;
    return 0;
}

// Replayed graphs save the host its launches, which matters when the
// steps are short.  From kDirectCells cells in the tables on, the host keeps
// ahead of the GPU with direct launches, and they were faster: at 256x128x96
// a mixed solve took 24.6 ms with direct launches against 25.5 ms replayed
// (FP64: 47.9 against 48.8 ms).
constexpr long long kDirectCells = 1LL << 20;

int Engine::enqueue_chunk(int K, bool timing)
{
    int rc;
    // capture and instantiation allocate: a no-allocation call launches directly
    if (graphs_ && !timing && table_cells_ < kDirectCells && !no_alloc_active()) {
        for (int it = 0; it < K; it += kRefreshGrain)
            if ((rc = enqueue_group_graph())) return rc;
    } else if ((rc = enqueue_steps_direct(K, timing))) {
        return rc;
    }
    const int slot = chunk_idx_ & 1;
    int hi = 0;
    for (int i = 0; i < max_slots_; ++i)
        if (slots_[i].used) hi = i + 1;
    ECHK(hipMemcpyAsync(h_status_ + (size_t)slot * max_slots_, d_status_, sizeof(CgScalars) * std::max(hi, 1),
                        hipMemcpyDeviceToHost, s_));
    ECHK(hipEventRecord(ev_status_[slot], s_));
    status_step_[slot] = step_;
    ++chunk_idx_;
    return 0;
}

int Engine::finish(const std::vector<int> &done)
{
    if (done.empty()) return 0;
    // tables for the finished problems, staged like the launch tables
    std::vector<TileRef> &t3 = fin_t3_;
    t3.clear();
    for (int si : done)
        for (int t = 0; t < slots_[si].pd.nt3; ++t) t3.push_back({si, t});
    const size_t poff = align_up(sizeof(TileRef) * t3.size(), 256);
    const size_t need = poff + sizeof(int) * done.size();
    if (need > fin_cap_) {
        NO_ALLOC_HERE("grow the finish tables");
        // grow: the old buffer may still be read by queued kernels
        ECHK(hipStreamSynchronize(s_));
        if (d_fin_) ECHK(hipFree(d_fin_));
        fin_cap_ = need * 2;
        ECHK(hipMalloc(&d_fin_, fin_cap_));
    }
    if (need > stage_cap_) {
        NO_ALLOC_HERE("grow the staging ring");
        ECHK(hipStreamSynchronize(s_));
        for (int i = 0; i < 4; ++i) {
            if (h_stage_[i]) ECHK(hipHostFree(h_stage_[i]));
            ECHK(hipHostMalloc(&h_stage_[i], need * 2, hipHostMallocDefault));
        }
        stage_cap_ = need * 2;
    }
    const int stg = stage_next_;
    stage_next_ = (stage_next_ + 1) & 3;
    ECHK(hipEventSynchronize(ev_stage_[stg]));
    char *h = static_cast<char *>(h_stage_[stg]);
    std::memcpy(h, t3.data(), sizeof(TileRef) * t3.size());
    std::memcpy(h + poff, done.data(), sizeof(int) * done.size());
    ECHK(hipMemcpyAsync(d_fin_, h, need, hipMemcpyHostToDevice, s_));
    ECHK(hipEventRecord(ev_stage_[stg], s_));
    const TileRef *dt = static_cast<const TileRef *>(d_fin_);
    const int *dp = reinterpret_cast<const int *>(static_cast<char *>(d_fin_) + poff);
    launch_finish(d_probs_, dt, (int)t3.size(), dp, (int)done.size(), kMaxGrid, s_);
    for (int si : done) {
        Slot &sl = slots_[si];
        launch_unscale(sl.g, sl.pd.e, sl.pd.d, sl.prd1_out, sl.pd.S, sl.solid, s_);
        sl.finished = true;
    }
    ECHK(hipGetLastError());
    tables_dirty_ = true;
    return 0;
}

int Engine::pump(bool timing, std::vector<int> &finished)
{
    int rc;
    if (tables_dirty_ && (rc = rebuild_tables())) return rc;
    // keep two chunks in flight: the host reads one chunk's state while the
    // next runs
    while (active_count() > 0 && npend_ < 2) {
        if ((rc = enqueue_chunk(next_chunk_len(), timing))) return rc;
        pend_[(pend_head_ + npend_) & 1] = (chunk_idx_ - 1) & 1;
        ++npend_;
    }
    if (npend_ == 0) return 0;
    const int which = pend_[pend_head_];
    pend_head_ = (pend_head_ + 1) & 1;
    --npend_;
    ECHK(hipEventSynchronize(ev_status_[which]));
    const CgScalars *h = h_status_ + (size_t)which * max_slots_;
    const int last_step = status_step_[which];
    std::vector<int> &done = done_;
    done.clear();
    for (int i = 0; i < max_slots_; ++i) {
        Slot &sl = slots_[i];
        // problems admitted after this chunk was queued took no part in it
        if (!sl.used || !sl.active || sl.finished || sl.n0 > last_step) continue;
        sl.status = h[i];
        if (h[i].final_) done.push_back(i);
    }
    if ((rc = finish(done))) return rc;
    for (int i : done) finished.push_back(i);
    return 0;
}

int Engine::read_final(const std::vector<int> &slots)
{
    int hi = 0;
    for (int i : slots) hi = std::max(hi, i + 1);
    if (hi == 0) return 0;
    ECHK(hipMemcpyAsync(h_final_, d_status_, sizeof(CgScalars) * hi, hipMemcpyDeviceToHost, s_));
    ECHK(hipStreamSynchronize(s_));
    for (int i : slots) slots_[i].status = h_final_[i];
    return 0;
}

int Engine::run_all(bool timing)
{
    int rc;
    ntimed_ = 0;
    replays_ = 0;
    if (timing) ECHK(hipMemsetAsync(d_ran_, 0, sizeof(int) * 2 * kMaxTimed, s_));
    std::vector<int> &all = run_slots_;
    all.clear();
    for (int i = 0; i < max_slots_; ++i)
        if (slots_[i].used && slots_[i].active && !slots_[i].finished) all.push_back(i);
    std::vector<int> &finished = run_finished_;
    finished.clear();
    const int max_rounds = 1 << 20;
    int r = 0;
    for (; r < max_rounds && (active_count() > 0 || npend_ > 0); ++r)
        if ((rc = pump(timing, finished))) return rc;
    if (active_count() > 0) return fail(SULI_PS_EHIP, "engine: problems did not finish");
    // final scalars (rr_true, gshift) after the finish kernels
    if ((rc = read_final(all))) return rc;
    // SpMV timing samples that did the work
    spmv_ms_ = 0.0;
    spmv_n_ = 0;
    if (timing && ntimed_ > 0) {
        // windows in which problem 0 iterated at both ends
        std::vector<int> &ran = ran_;
        ran.assign(2 * ntimed_, 0);
        ECHK(hipMemcpyAsync(ran.data(), d_ran_, sizeof(int) * 2 * ntimed_, hipMemcpyDeviceToHost, s_));
        ECHK(hipStreamSynchronize(s_));
        double sum = 0.0;
        for (int t = 0; t < ntimed_; ++t) {
            if (!ran[2 * t] || !ran[2 * t + 1]) continue;
            float ms = 0.f;
            ECHK(hipEventElapsedTime(&ms, ev_spmv_[2 * t], ev_spmv_[2 * t + 1]));
            sum += ms / kTimeSteps;
            ++spmv_n_;
        }
        spmv_ms_ = spmv_n_ ? sum / spmv_n_ : 0.0;
    }
    return 0;
}

void Engine::set_tiles(int nx, int ny, int nz, bool mixed, int kz, int vx)
{
    const Geo g = make_geo(nx, ny, nz, 0, 0, std::max(kz, kMinKz), mixed ? vx : 1);
    tuned_[std::make_tuple(nx, ny, nz, mixed)] = Tuned{g.kz, lane_code(g), 0.0};
}

int Engine::tune(int nx, int ny, int nz, bool mixed, int *kz_out, int *vx_out, double *us_out)
{
    const auto key = std::make_tuple(nx, ny, nz, mixed);
    auto hit = tuned_.find(key);
    if (hit != tuned_.end()) {
        *kz_out = hit->second.kz;
        *vx_out = hit->second.vx;
        *us_out = hit->second.us;
        return 0;
    }
    NO_ALLOC_HERE("tune a tile shape");
    const Geo g_def = make_geo(nx, ny, nz, 0, 0, 0);
    std::vector<int> cands = {2, 3, 4, 6, 8, 12, 16, 24, 32, 48};
    if (std::find(cands.begin(), cands.end(), g_def.kz) == cands.end()) cands.push_back(g_def.kz);
    const size_t bytes = problem_bytes(nx, ny, nz, mixed);
    void *blk = pool_.get(bytes, err_);
    if (!blk) return SULI_PS_ENOMEM;
    ProbDev *dP = nullptr;
    CgScalars *dS = nullptr;
    TileRef *dT = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = 0;
    const Geo g_min = make_geo(nx, ny, nz, 0, 0, kMinKz);
    const int max_nt3 = spmv_tiles(g_min);
    auto cleanup = [&] {
        (void)hipStreamSynchronize(s_);
        if (dP) (void)hipFree(dP);
        if (dS) (void)hipFree(dS);
        if (dT) (void)hipFree(dT);
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        pool_.put(blk, bytes);
    };
#define TCHK(call)                                                                                       \
    do {                                                                                                 \
        hipError_t e_ = (call);                                                                          \
        if (e_ != hipSuccess) {                                                                          \
            cleanup();                                                                                   \
            return fail(SULI_PS_EHIP, std::string("autotune: ") + #call + ": " + hipGetErrorString(e_)); \
        }                                                                                                \
    } while (0)
    TCHK(hipMalloc(&dP, sizeof(ProbDev)));
    TCHK(hipMalloc(&dS, sizeof(CgScalars)));
    TCHK(hipMalloc(&dT, sizeof(TileRef) * max_nt3));
    TCHK(hipEventCreate(&a));
    TCHK(hipEventCreate(&b));
    // state that keeps the per-step kernel on its working path: zero
    // vectors; FP64: entering r.r 0 over a previous 1 (beta 0), threshold
    // -1; mixed: zero partials (alpha and beta NaN, which no branch reads)
    CgScalars s0{};
    s0.rr[0] = s0.rr[1] = 1.0;
    s0.thresh = -1.0;
    s0.done_at = kNotDone;
    s0.skip_e_at = -1;
    s0.max_iter = 1 << 30;
    s0.mixed = mixed ? 1 : 0;
    s0.next_refresh = kNotDone;
    std::vector<TileRef> tab(max_nt3);
    // FP64: n = 3 every launch, the same inputs each time.  Mixed: n runs
    // through 2 .. 9 again and again, so that the timed launches alternate
    // their buffers and include the fold steps (n = 5, 9) in their real
    // proportion; repeating one plain step on the same buffers timed
    // several-cells-per-lane shapes up to 25 % faster than they run in a
    // solve.  The zero vectors keep every value finite or NaN, and no
    // branch of the step depends on them.
    const int reps = 24;
    double best_us = 1e300;
    int best_kz = g_def.kz, best_vx = 1;
    std::vector<std::pair<int, int>> shapes;
    for (int vx : {1, 2, 2 | kPipelined, 4, 4 | kPipelined}) {
        const int w = vx & ~kPipelined;
        if (w > 1 && (!mixed || nx % w)) continue;
        for (int kz : cands)
            if (kz <= nz || kz == cands.front()) shapes.emplace_back(kz, vx);
    }
    for (const auto &shape : shapes) {
        const int kz = shape.first;
        const Geo g = make_geo(nx, ny, nz, 0, 0, kz, shape.second);
        ProbDev pd{};
        Carver cv{static_cast<char *>(blk)};
        const int nt3 = spmv_tiles(g), nt1 = vec_tiles(g.n);
        carve(cv, g, mixed, nt3, nt1, pd);
        // tuned on the face-array operator whose rows all hold their
        // plane's values (a layered or constant-density field: the FP32
        // step reads the row tables instead of the arrays)
        pd.tx = pd.ty = pd.tz = nullptr;
        pd.tx32 = pd.ty32 = pd.tz32 = nullptr;
        const size_t row_bytes = mixed ? sizeof(unsigned) * (size_t)(ny + 31) / 32 * nz : 0;
        pd.g = g;
        pd.nt3 = nt3;
        pd.nt1 = nt1;
        pd.mixed = mixed ? 1 : 0;
        pd.n0 = 1;
        pd.S = dS;
        for (int t = 0; t < nt3; ++t) tab[t] = {0, t};
        TCHK(hipMemsetAsync(blk, 0, cv.off, s_));
        if (row_bytes) TCHK(hipMemsetAsync(const_cast<unsigned *>(pd.rowreg), 0xff, row_bytes, s_));
        TCHK(hipMemcpyAsync(dP, &pd, sizeof pd, hipMemcpyHostToDevice, s_));
        TCHK(hipMemcpyAsync(dS, &s0, sizeof s0, hipMemcpyHostToDevice, s_));
        TCHK(hipMemcpyAsync(dT, tab.data(), sizeof(TileRef) * nt3, hipMemcpyHostToDevice, s_));
        TCHK(hipStreamSynchronize(s_));
        const int grid = std::min(nt3, kMaxGrid);
        int seq = 0;
        auto launch = [&] {
            const int step = mixed ? 2 + (seq++ % 8) : 3;
            if (mixed)
                launch_cg_m(dP, dT, nt3, grid, StepRef{nullptr, step, nullptr}, nullptr, g.vx, g.pipe, s_);
            else
                launch_cg_a(false, dP, dT, nt3, grid, StepRef{nullptr, step, nullptr}, nullptr, s_);
        };
        for (int r = 0; r < 8; ++r) launch();
        TCHK(hipEventRecord(a, s_));
        for (int r = 0; r < reps; ++r) launch();
        TCHK(hipEventRecord(b, s_));
        TCHK(hipEventSynchronize(b));
        TCHK(hipGetLastError());
        float ms = 0.f;
        TCHK(hipEventElapsedTime(&ms, a, b));
        const double us = 1e3 * ms / reps;
        if (us < best_us) {
            best_us = us;
            best_kz = g.kz;
            best_vx = lane_code(g);
        }
    }
#undef TCHK
    cleanup();
    tuned_[key] = Tuned{best_kz, best_vx, best_us};
    *kz_out = best_kz;
    *vx_out = best_vx;
    *us_out = best_us;
    (void)rc;
    return 0;
}

int Engine::time_spmv(bool fp32, int reps, double *ms)
{
    const Slot &sl = slots_[0];
    if (!sl.used) return fail(SULI_PS_EINVAL, "no operator: call graddin first");
    if (fp32 != (bool)sl.o.mixed)
        return fail(SULI_PS_EINVAL, "the last solve's operator is in the other precision");
    // the SpMV's own tile shapes: depths of 2 .. 24 planes, and 1, 2 or 4
    // cells per lane in FP32 (where nx is a multiple); the fastest counts
    std::vector<std::pair<int, int>> shapes;
    for (int vx : {4, 2, 1}) {
        if (vx > 1 && (!fp32 || sl.g.nx % vx)) continue;
        for (int kz : {2, 3, 4, 6, 8, 12, 16, 24})
            if (kz <= sl.g.nz) shapes.emplace_back(kz, vx);
    }
    // every shape's tiles fit the partial-sum array (sized for the most tiles)
    const int max_nt3 = spmv_tiles(make_geo(sl.g.nx, sl.g.ny, sl.g.nz, 0, 0, kMinKz));
    ProbDev *dP = nullptr;
    TileRef *dt = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    auto cleanup = [&] {
        (void)hipStreamSynchronize(s_);
        if (dP) (void)hipFree(dP);
        if (dt) (void)hipFree(dt);
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    };
#define SCHK(call)                                                                                      \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess) {                                                                         \
            cleanup();                                                                                  \
            return fail(SULI_PS_EHIP, std::string("time_spmv: ") + #call + ": " + hipGetErrorString(e_)); \
        }                                                                                               \
    } while (0)
    SCHK(hipMalloc(&dP, sizeof(ProbDev)));
    SCHK(hipMalloc(&dt, sizeof(TileRef) * max_nt3));
    SCHK(hipEventCreate(&a));
    SCHK(hipEventCreate(&b));
    std::vector<TileRef> t3(max_nt3);
    for (int t = 0; t < max_nt3; ++t) t3[t] = {0, t};
    SCHK(hipMemcpy(dt, t3.data(), sizeof(TileRef) * max_nt3, hipMemcpyHostToDevice));
    double best = 1e300;
    for (const auto &sh : shapes) {
        ProbDev pd = sl.pd;
        pd.g = make_geo(sl.g.nx, sl.g.ny, sl.g.nz, sl.g.px, sl.g.py, sh.first, sh.second);
        pd.g.dtop = sl.g.dtop;
        pd.nt3 = spmv_tiles(pd.g);
        if (pd.nt3 > max_nt3) continue;
        SCHK(hipMemcpy(dP, &pd, sizeof pd, hipMemcpyHostToDevice));
        const int n3 = pd.nt3, grid = std::min(n3, kMaxGrid);
        launch_spmv_only(fp32, dP, dt, n3, grid, s_);
        SCHK(hipEventRecord(a, s_));
        for (int i = 0; i < reps; ++i) launch_spmv_only(fp32, dP, dt, n3, grid, s_);
        SCHK(hipEventRecord(b, s_));
        SCHK(hipEventSynchronize(b));
        SCHK(hipGetLastError());
        float t = 0.f;
        SCHK(hipEventElapsedTime(&t, a, b));
        best = std::min(best, (double)t / reps);
    }
#undef SCHK
    cleanup();
    *ms = best;
    return 0;
}

}  // namespace sps
