// Serving loop (include/suli_ps.h, suli_ps_server_*): pressure solves
// submitted as they arrive, admitted by an HBM budget and solved together
// by one engine (continuous batching).
//
// Two threads per server:
//   copy thread    admits queued requests in arrival order while their device
//                  memory fits the budget, copies their inputs on the copy
//                  stream (the setup kernels wait for that copy's event), and
//                  copies finished results back to the caller's prd1;
//   device thread  moves uploaded requests into the engine between chunks
//                  and pumps the engine: every resident request runs in the
//                  same launches, joining and leaving at chunk boundaries.
// Each request's arithmetic depends only on its own grid (ps_cg.hip), so its
// result is the same bits as a solo solve with the same options.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/suli_ps.h"
#include "ps_common.h"
#include "ps_transfer.h"
#include "ps_engine.h"
#include "ps_internal.h"

using namespace sps;

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kMaxResident = 256;   // engine slots

struct Request {
    enum State { QUEUED, UPLOADING, READY, RUNNING, FINISHING, DOWNLOADING, DONE, FAILED };
    int64_t ticket = 0;
    State state = QUEUED;
    suli_ps_grid g{};
    // caller's host arrays
    const double *u, *v, *w, *rho, *ox, *oy, *oz, *dz, *dzz;
    double *prd1;
    // device block: fields (reference layouts) then the solver state
    size_t field_bytes = 0, bytes = 0;
    void *block = nullptr;
    double *du = nullptr, *dv = nullptr, *dw = nullptr, *drho = nullptr, *dprd1 = nullptr;
    double *dox = nullptr, *doy = nullptr, *doz = nullptr, *ddz = nullptr, *ddzz = nullptr;
    void *state_block = nullptr;
    hipEvent_t ev_up = nullptr, ev_done = nullptr;
    int slot = -1;
    Slot final_slot{};
    int rc = 0;
    std::string err;
    Clock::time_point t_submit;
    double latency_ms = 0.0;
};

size_t up256(size_t x) { return (x + 255) / 256 * 256; }

// Carves a request's block: sizes of the reference-layout fields it copies.
size_t carve_fields(Request &r, char *base)
{
    const suli_ps_grid &g = r.g;
    const size_t nx = g.nx, ny = g.ny, nz = g.nz;
    const bool mask = g.obst_t != 0 && g.ibm_t == 2;
    size_t off = 0;
    auto take = [&](double *&dst, size_t count, bool present) {
        dst = present && base ? reinterpret_cast<double *>(base + off) : nullptr;
        if (present) off += up256(count * sizeof(double));
    };
    take(r.du, (nx + 3) * (ny + 2) * (nz + 2), true);
    take(r.dv, (nx + 2) * (ny + 3) * (nz + 2), true);
    take(r.dw, (nx + 2) * (ny + 2) * (nz + 3), true);
    take(r.drho, nx * ny * nz, r.rho != nullptr);
    take(r.dprd1, (nx + 2) * (ny + 2) * (nz + 2), true);
    take(r.dox, (nx + 1) * ny * nz, mask);
    take(r.doy, nx * (ny + 1) * nz, mask);
    take(r.doz, nx * ny * (nz + 1), mask);
    take(r.ddz, nz + 2, r.dz != nullptr);
    take(r.ddzz, nz + 2, r.dz != nullptr);
    return up256(off);
}

}  // namespace

struct suli_ps_server {
    suli_ps_server_opts opts{};
    int device = 0;
    hipStream_t s_dev = nullptr, s_copy = nullptr;
    HostPipe pipe;   // pinned staging of the copy thread's copies (that thread only)
    Engine eng;
    std::mutex m;
    std::condition_variable cv_copy, cv_dev, cv_done;
    std::map<int64_t, Request *> reqs;     // every request not yet waited for
    std::deque<Request *> queue;           // QUEUED, arrival order
    std::vector<Request *> ready;          // READY for the engine
    std::deque<Request *> downloads;       // DOWNLOADING
    std::vector<int> to_release;           // engine slots whose results are copied
    int64_t next_ticket = 0, submitted = 0, completed = 0, chunks = 0;
    int resident = 0;
    uint64_t resident_bytes = 0, peak_bytes = 0, budget = 0;
    // device blocks kept for reuse by size (held = in use + cached)
    std::multimap<size_t, void *> cache;
    uint64_t held = 0;
    bool stop = false;
    bool broken = false;     // the engine failed: no further solves
    std::mutex mcap;         // device allocation and free vs graph capture
    // events of waited-for requests, reused by later ones (destroying an
    // event recorded on the device stream is not allowed while it captures)
    // (kept apart: copy-stream events and device-stream events)
    std::vector<hipEvent_t> spare_up, spare_done;
    std::string err;
    std::thread t_copy, t_dev;

    void copy_loop();
    void dev_loop();
    // With the lock held: a resident request ends with an error and gives
    // its device memory back.
    void fail_resident(Request *r, int rc, const std::string &msg)
    {
        r->rc = rc;
        r->err = msg;
        if (r->block) cache.emplace(r->bytes, r->block);
        r->block = nullptr;
        r->state = Request::FAILED;
        --resident;
        resident_bytes -= r->bytes;
        ++completed;
    }
    void *alloc_block(size_t bytes, std::string &e);
};

void *suli_ps_server::alloc_block(size_t bytes, std::string &e)
{
    {
        std::lock_guard<std::mutex> lk(m);
        auto it = cache.find(bytes);
        if (it != cache.end()) {
            void *p = it->second;
            cache.erase(it);
            return p;
        }
    }
    // a new block: keep device memory held within the budget by giving back
    // cached blocks of other sizes first
    std::vector<std::pair<size_t, void *>> drop;
    {
        std::lock_guard<std::mutex> lk(m);
        while (held + bytes > budget && !cache.empty()) {
            auto it = std::prev(cache.end());
            drop.emplace_back(it->first, it->second);
            held -= it->first;
            cache.erase(it);
        }
        held += bytes;
    }
    void *p = nullptr;
    hipError_t he;
    {
        // a free synchronises the device, which would break a graph the
        // device thread is capturing
        std::lock_guard<std::mutex> cl(mcap);
        for (auto &d : drop) (void)hipFree(d.second);
        he = hipMalloc(&p, bytes);
    }
    if (he != hipSuccess) {
        std::lock_guard<std::mutex> lk(m);
        held -= bytes;
        e = std::string("hipMalloc failed: ") + hipGetErrorString(he);
        return nullptr;
    }
    return p;
}

void suli_ps_server::copy_loop()
{
    (void)hipSetDevice(device);
    for (;;) {
        Request *down = nullptr, *up = nullptr;
        {
            std::unique_lock<std::mutex> lk(m);
            cv_copy.wait(lk, [&] {
                if (!downloads.empty()) return true;
                if (!queue.empty()) {
                    Request *r = queue.front();
                    if (resident < opts.max_active && resident_bytes + r->bytes <= budget) return true;
                }
                return stop && queue.empty() && downloads.empty() && resident == 0;
            });
            if (!downloads.empty()) {
                down = downloads.front();
                downloads.pop_front();
            } else if (!queue.empty()) {
                up = queue.front();
                queue.pop_front();
                up->state = Request::UPLOADING;
                ++resident;
                resident_bytes += up->bytes;
                peak_bytes = std::max(peak_bytes, resident_bytes);
            } else {
                return;   // stopping and drained
            }
        }
        if (down) {
            // result and final scalars; the device thread hands a request
            // over once its finish kernels have completed (it alone touches
            // the device stream's events: that stream may be capturing a
            // graph, and no other thread may wait on its events then)
            Request &r = *down;
            const size_t nbox = (size_t)(r.g.nx + 2) * (r.g.ny + 2) * (r.g.nz + 2);
            CgScalars fin{};
            const HostSeg res{r.dprd1, r.prd1, nbox * sizeof(double)};
            hipError_t e = pipe.download(device, s_copy, &res, 1);
            if (e == hipSuccess)
                e = hipMemcpyAsync(&fin, eng.device_status(r.slot), sizeof fin, hipMemcpyDeviceToHost, s_copy);
            if (e == hipSuccess) e = hipStreamSynchronize(s_copy);
            std::lock_guard<std::mutex> lk(m);
            if (e != hipSuccess) {
                r.rc = SULI_PS_EHIP;
                r.err = std::string("result copy failed: ") + hipGetErrorString(e);
                r.state = Request::FAILED;
            } else {
                r.final_slot.status = fin;
                r.state = Request::DONE;
            }
            r.latency_ms = std::chrono::duration<double, std::milli>(Clock::now() - r.t_submit).count();
            to_release.push_back(r.slot);
            cache.emplace(r.bytes, r.block);
            r.block = nullptr;
            --resident;
            resident_bytes -= r.bytes;
            ++completed;
            cv_done.notify_all();
            cv_dev.notify_all();
            continue;
        }
        // upload
        Request &r = *up;
        std::string e;
        r.block = alloc_block(r.bytes, e);
        hipError_t he = hipSuccess;
        if (r.block) {
            carve_fields(r, static_cast<char *>(r.block));
            r.state_block = static_cast<char *>(r.block) + r.field_bytes;
            const size_t nx = r.g.nx, ny = r.g.ny, nz = r.g.nz;
            const bool mask = r.g.obst_t != 0 && r.g.ibm_t == 2;
            struct Cp {
                double *dst;
                const double *src;
                size_t count;
            } cps[] = {
                {r.du, r.u, (nx + 3) * (ny + 2) * (nz + 2)},
                {r.dv, r.v, (nx + 2) * (ny + 3) * (nz + 2)},
                {r.dw, r.w, (nx + 2) * (ny + 2) * (nz + 3)},
                {r.drho, r.rho, nx * ny * nz},
                {r.dprd1, r.prd1, (nx + 2) * (ny + 2) * (nz + 2)},
                {mask ? r.dox : nullptr, r.ox, (nx + 1) * ny * nz},
                {mask ? r.doy : nullptr, r.oy, nx * (ny + 1) * nz},
                {mask ? r.doz : nullptr, r.oz, nx * ny * (nz + 1)},
                {r.ddz, r.dz, nz + 2},
                {r.ddzz, r.dzz, nz + 2},
            };
            HostSeg segs[sizeof cps / sizeof cps[0]];
            int nseg = 0;
            for (const Cp &c : cps)
                if (c.dst && c.src) segs[nseg++] = HostSeg{c.dst, const_cast<double *>(c.src), c.count * sizeof(double)};
            he = pipe.upload(device, s_copy, segs, nseg);
            if (he == hipSuccess) he = hipEventRecord(r.ev_up, s_copy);
        }
        std::lock_guard<std::mutex> lk(m);
        if (!r.block || he != hipSuccess) {
            r.rc = r.block ? SULI_PS_EHIP : SULI_PS_ENOMEM;
            r.err = r.block ? std::string("input copy failed: ") + hipGetErrorString(he) : e;
            if (r.block) cache.emplace(r.bytes, r.block);
            r.block = nullptr;
            r.state = Request::FAILED;
            --resident;
            resident_bytes -= r.bytes;
            ++completed;
            cv_done.notify_all();
            continue;
        }
        r.state = Request::READY;
        ready.push_back(&r);
        cv_dev.notify_all();
    }
}

void suli_ps_server::dev_loop()
{
    (void)hipSetDevice(device);
    const SolveOpts so = solve_opts(opts.solve);
    std::vector<int> finished;
    std::deque<Request *> finishing;   // finish kernels queued, not yet known complete
    // hands requests whose finish kernels have completed to the copy thread
    auto promote = [&](bool block) {
        while (!finishing.empty()) {
            Request *r = finishing.front();
            const hipError_t q = block ? hipEventSynchronize(r->ev_done) : hipEventQuery(r->ev_done);
            if (q == hipErrorNotReady) break;
            finishing.pop_front();
            std::lock_guard<std::mutex> lk(m);
            if (q != hipSuccess) {
                to_release.push_back(r->slot);
                fail_resident(r, SULI_PS_EHIP, std::string("finish: ") + hipGetErrorString(q));
                cv_done.notify_all();
                continue;
            }
            r->state = Request::DOWNLOADING;
            downloads.push_back(r);
            cv_copy.notify_all();
            block = false;
        }
    };
    for (;;) {
        promote(false);
        std::vector<Request *> incoming;
        std::vector<int> rel;
        {
            std::unique_lock<std::mutex> lk(m);
            cv_dev.wait(lk, [&] {
                return !ready.empty() || !to_release.empty() || !finishing.empty() ||
                       (!broken && (eng.active_count() > 0 || eng.pending_chunks() > 0)) ||
                       (stop && resident == 0 && queue.empty());
            });
            if (stop && resident == 0 && queue.empty() && ready.empty() && finishing.empty() &&
                (broken || (eng.active_count() == 0 && eng.pending_chunks() == 0))) {
                for (int s : to_release) eng.release(s);
                to_release.clear();
                return;
            }
            incoming.swap(ready);
            rel.swap(to_release);
        }
        for (int s : rel) eng.release(s);
        for (Request *r : incoming) {
            const SetupIn in = make_setup(&r->g, r->du, r->dv, r->dw, r->drho, r->dprd1, r->dox, r->doy, r->doz,
                                          r->ddz, r->ddzz);
            int slot = -1;
            const int rc = broken ? SULI_PS_EHIP
                                  : eng.admit(r->g, in, r->dprd1, so, (int)(r->ticket & 0x7fffffff), &slot,
                                              r->state_block, r->ev_up);
            std::lock_guard<std::mutex> lk(m);
            if (rc) {
                fail_resident(r, rc, broken ? err : eng.error());
                cv_done.notify_all();
                cv_copy.notify_all();
                continue;
            }
            r->slot = slot;
            r->state = Request::RUNNING;
        }
        if (broken || (eng.active_count() == 0 && eng.pending_chunks() == 0)) {
            // nothing to launch: wait for the oldest finishing request
            if (incoming.empty() && !finishing.empty()) promote(true);
            continue;
        }
        finished.clear();
        const int before = eng.chunks_enqueued();
        int rc = eng.pump(false, finished);
        std::vector<Request *> done_reqs;
        for (int s : finished) {
            hipEvent_t ev = nullptr;
            {
                std::lock_guard<std::mutex> lk(m);
                for (auto &kv : reqs)
                    if (kv.second->state == Request::RUNNING && kv.second->slot == s) ev = kv.second->ev_done;
            }
            if (ev) (void)hipEventRecord(ev, s_dev);
        }
        std::lock_guard<std::mutex> lk(m);
        chunks += eng.chunks_enqueued() - before;
        for (int s : finished)
            for (auto &kv : reqs) {
                Request *r = kv.second;
                if (r->state != Request::RUNNING || r->slot != s) continue;
                r->final_slot = eng.slot(s);
                r->state = Request::FINISHING;
                finishing.push_back(r);
            }
        if (rc) {
            // an engine failure ends every request on the device; later
            // ones fail at admission
            broken = true;
            err = eng.error();
            for (auto &kv : reqs) {
                Request *r = kv.second;
                if (r->state == Request::RUNNING) fail_resident(r, rc, err);
            }
            cv_done.notify_all();
            cv_copy.notify_all();
        }
    }
}

extern "C" {

suli_ps_server *suli_ps_server_create(const suli_ps_server_opts *opts)
{
    suli_ps_server *srv = new suli_ps_server();
    if (opts) srv->opts = *opts;
    if (!opts) {
        srv->opts.solve.precision = SULI_PS_FP64;
        srv->opts.solve.stop_rule = SULI_PS_STOP_RELATIVE;
        srv->opts.solve.tol = 1e-8;
    }
    if (srv->opts.max_active <= 0 || srv->opts.max_active > kMaxResident) srv->opts.max_active = kMaxResident;
    bool ok = hipGetDevice(&srv->device) == hipSuccess &&
              hipStreamCreateWithFlags(&srv->s_dev, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&srv->s_copy, hipStreamNonBlocking) == hipSuccess;
    size_t free_b = 0, total_b = 0;
    ok = ok && hipMemGetInfo(&free_b, &total_b) == hipSuccess;
    ok = ok && srv->eng.init(srv->s_dev, kMaxResident) == 0;
    if (!ok) {
        if (srv->s_dev) (void)hipStreamDestroy(srv->s_dev);
        if (srv->s_copy) (void)hipStreamDestroy(srv->s_copy);
        delete srv;
        return nullptr;
    }
    srv->budget = srv->opts.hbm_budget ? srv->opts.hbm_budget : (uint64_t)(0.8 * (double)free_b);
    srv->eng.set_graphs(srv->opts.solve.no_graphs == 0);
    srv->eng.set_capture_mutex(&srv->mcap);
    srv->eng.set_autotune(srv->opts.solve.autotune != 0);
    srv->eng.set_chunk(srv->opts.solve.chunk);
    srv->t_copy = std::thread([srv] { srv->copy_loop(); });
    srv->t_dev = std::thread([srv] { srv->dev_loop(); });
    return srv;
}

void suli_ps_server_destroy(suli_ps_server *srv)
{
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->m);
        srv->stop = true;
    }
    srv->cv_copy.notify_all();
    srv->cv_dev.notify_all();
    // the copy thread ends once everything is drained; wake it as the device
    // thread finishes requests (it is notified on every completion)
    srv->t_dev.join();
    {
        std::lock_guard<std::mutex> lk(srv->m);
    }
    srv->cv_copy.notify_all();
    srv->t_copy.join();
    (void)hipStreamSynchronize(srv->s_dev);
    (void)hipStreamSynchronize(srv->s_copy);
    srv->pipe.release();
    for (auto &kv : srv->reqs) {
        Request *r = kv.second;
        if (r->ev_up) (void)hipEventDestroy(r->ev_up);
        if (r->ev_done) (void)hipEventDestroy(r->ev_done);
        if (r->block) (void)hipFree(r->block);
        delete r;
    }
    for (auto &kv : srv->cache) (void)hipFree(kv.second);
    for (hipEvent_t e : srv->spare_up) (void)hipEventDestroy(e);
    for (hipEvent_t e : srv->spare_done) (void)hipEventDestroy(e);
    srv->eng.~Engine();
    new (&srv->eng) Engine();
    (void)hipStreamDestroy(srv->s_dev);
    (void)hipStreamDestroy(srv->s_copy);
    delete srv;
}

const char *suli_ps_server_last_error(const suli_ps_server *srv) { return srv ? srv->err.c_str() : "no server"; }

int64_t suli_ps_server_submit(suli_ps_server *srv, const suli_ps_grid *g, const double *u, const double *v,
                              const double *w, const double *rho, const double *obs_lsx, const double *obs_lsy,
                              const double *obs_lsz, const double *dz, const double *dzz, double *prd1)
{
    if (!srv) return SULI_PS_EINVAL;
    std::string e;
    int rc = validate_problem(g, u, v, w, rho, obs_lsx, obs_lsy, obs_lsz, dz, dzz, prd1, e);
    if (rc) {
        std::lock_guard<std::mutex> lk(srv->m);
        srv->err = e;
        return rc;
    }
    Request *r = new Request();
    r->g = *g;
    r->u = u;
    r->v = v;
    r->w = w;
    r->rho = g->t_press == 1 ? rho : nullptr;
    const bool mask = g->obst_t != 0 && g->ibm_t == 2;
    r->ox = mask ? obs_lsx : nullptr;
    r->oy = mask ? obs_lsy : nullptr;
    r->oz = mask ? obs_lsz : nullptr;
    r->dz = dz;
    r->dzz = dzz;
    r->prd1 = prd1;
    r->field_bytes = carve_fields(*r, nullptr);
    r->bytes = r->field_bytes + problem_bytes(g->nx, g->ny, g->nz, srv->opts.solve.precision == SULI_PS_MIXED);
    std::lock_guard<std::mutex> lk(srv->m);
    for (auto pr : {std::make_pair(&r->ev_up, &srv->spare_up), std::make_pair(&r->ev_done, &srv->spare_done)}) {
        if (!pr.second->empty()) {
            *pr.first = pr.second->back();
            pr.second->pop_back();
        } else if (hipEventCreateWithFlags(pr.first, hipEventDisableTiming) != hipSuccess) {
            *pr.first = nullptr;
        }
    }
    if (!r->ev_up || !r->ev_done) {
        if (r->ev_up) srv->spare_up.push_back(r->ev_up);
        if (r->ev_done) srv->spare_done.push_back(r->ev_done);
        delete r;
        return SULI_PS_EHIP;
    }
    if (r->bytes > srv->budget) {
        char buf[160];
        std::snprintf(buf, sizeof buf, "request needs %zu bytes of device memory, the budget is %llu", r->bytes,
                      (unsigned long long)srv->budget);
        srv->err = buf;
        srv->spare_up.push_back(r->ev_up);
        srv->spare_done.push_back(r->ev_done);
        delete r;
        return SULI_PS_ENOMEM;
    }
    if (srv->stop) {
        srv->err = "server is stopping";
        srv->spare_up.push_back(r->ev_up);
        srv->spare_done.push_back(r->ev_done);
        delete r;
        return SULI_PS_EINVAL;
    }
    r->ticket = srv->next_ticket++;
    r->t_submit = Clock::now();
    srv->reqs[r->ticket] = r;
    srv->queue.push_back(r);
    ++srv->submitted;
    srv->cv_copy.notify_all();
    return r->ticket;
}

int suli_ps_server_wait(suli_ps_server *srv, int64_t ticket, double timeout_s, suli_ps_stats *stats)
{
    if (!srv) return SULI_PS_EINVAL;
    std::unique_lock<std::mutex> lk(srv->m);
    auto it = srv->reqs.find(ticket);
    if (it == srv->reqs.end()) {
        srv->err = "unknown ticket (already waited for?)";
        return SULI_PS_EINVAL;
    }
    Request *r = it->second;
    auto finished = [&] { return r->state == Request::DONE || r->state == Request::FAILED; };
    if (timeout_s < 0) {
        srv->cv_done.wait(lk, finished);
    } else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished)) {
        return SULI_PS_EAGAIN;
    }
    int rc = r->rc;
    if (r->state == Request::DONE) {
        suli_ps_stats st;
        fill_stats(r->final_slot, &st);
        st.solve_ms = r->latency_ms;
        if (stats) *stats = st;
        rc = st.status;
    } else {
        srv->err = r->err;
    }
    srv->reqs.erase(it);
    srv->spare_up.push_back(r->ev_up);
    srv->spare_done.push_back(r->ev_done);
    lk.unlock();
    delete r;
    return rc;
}

int suli_ps_server_get_info(suli_ps_server *srv, suli_ps_server_info *info)
{
    if (!srv || !info) return SULI_PS_EINVAL;
    std::lock_guard<std::mutex> lk(srv->m);
    std::memset(info, 0, sizeof *info);
    info->submitted = srv->submitted;
    info->completed = srv->completed;
    info->queued = (int32_t)srv->queue.size();
    info->resident = srv->resident;
    info->resident_bytes = srv->resident_bytes;
    info->peak_resident_bytes = srv->peak_bytes;
    info->budget = srv->budget;
    info->chunks = srv->chunks;
    return 0;
}

}  // extern "C"
