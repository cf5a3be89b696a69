// Pinned, multi-threaded staging of the host-pointer entry points' copies.
#include "ps_transfer.h"

#include <cstring>
#include <exception>
#include <thread>
#include <vector>

namespace sps {

namespace {

struct Chunk {
    char *dev;
    char *host;
    size_t bytes;
};

// One worker's chunks, host -> device on stream st: each chunk is copied
// into the pinned buffer whose previous DMA has finished, then sent.
hipError_t upload_chunks(hipStream_t st, char *const pin[2], hipEvent_t const ev[2], const std::vector<Chunk> &cs)
{
    hipError_t e = hipSuccess;
    for (size_t k = 0; e == hipSuccess && k < cs.size(); ++k) {
        const int b = (int)(k & 1);
        if ((e = hipEventSynchronize(ev[b])) != hipSuccess) break;
        std::memcpy(pin[b], cs[k].host, cs[k].bytes);
        if ((e = hipMemcpyAsync(cs[k].dev, pin[b], cs[k].bytes, hipMemcpyHostToDevice, st)) != hipSuccess) break;
        e = hipEventRecord(ev[b], st);
    }
    return e;
}

// One worker's chunks, device -> host on stream st (after the work already
// on it): the DMA of chunk k + 1 runs while chunk k is copied out of its
// pinned buffer.
hipError_t download_chunks(hipStream_t st, char *const pin[2], hipEvent_t const ev[2], const std::vector<Chunk> &cs)
{
    // the buffers may still be the source of an earlier upload's DMA
    hipError_t e = hipEventSynchronize(ev[0]);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    auto issue = [&](size_t k) {
        const int b = (int)(k & 1);
        hipError_t r = hipMemcpyAsync(pin[b], cs[k].dev, cs[k].bytes, hipMemcpyDeviceToHost, st);
        return r == hipSuccess ? hipEventRecord(ev[b], st) : r;
    };
    for (size_t k = 0; e == hipSuccess && k < cs.size() && k < 2; ++k) e = issue(k);
    for (size_t k = 0; e == hipSuccess && k < cs.size(); ++k) {
        const int b = (int)(k & 1);
        if ((e = hipEventSynchronize(ev[b])) != hipSuccess) break;
        std::memcpy(cs[k].host, pin[b], cs[k].bytes);
        if (k + 2 < cs.size()) e = issue(k + 2);
    }
    return e;
}

}  // namespace

HostPipe::~HostPipe() { release(); }

void HostPipe::release()
{
    for (Worker &w : w_) {
        // no DMA may still use a buffer when it is freed
        for (hipEvent_t &e : w.ev)
            if (e) {
                (void)hipEventSynchronize(e);
                (void)hipEventDestroy(e);
                e = nullptr;
            }
        for (char *&p : w.pin)
            if (p) {
                (void)hipHostFree(p);
                p = nullptr;
            }
    }
    ready_ = false;
    device_ = -1;
}

hipError_t HostPipe::ensure(int device)
{
    if (ready_ && device_ == device) return hipSuccess;
    release();
    hipError_t e = hipSuccess;
    for (Worker &w : w_)
        for (int b = 0; b < 2 && e == hipSuccess; ++b) {
            e = hipHostMalloc(reinterpret_cast<void **>(&w.pin[b]), kChunk, hipHostMallocDefault);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&w.ev[b], hipEventDisableTiming);
        }
    if (e != hipSuccess) {
        release();
        return e;
    }
    ready_ = true;
    device_ = device;
    return hipSuccess;
}

hipError_t HostPipe::upload(int device, hipStream_t s, const HostSeg *segs, int n)
{
    return run(device, s, segs, n, true);
}

hipError_t HostPipe::download(int device, hipStream_t s, const HostSeg *segs, int n)
{
    return run(device, s, segs, n, false);
}

hipError_t HostPipe::run(int device, hipStream_t s, const HostSeg *segs, int n, bool up)
{
    const hipMemcpyKind kind = up ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    size_t total = 0;
    for (int i = 0; i < n; ++i) total += segs[i].bytes;
    if (total < kDirect || ensure(device) != hipSuccess) {
        // small, or no pinned buffers to be had: the runtime's own path
        for (int i = 0; i < n; ++i) {
            if (!segs[i].bytes) continue;
            void *dst = up ? segs[i].dev : segs[i].host;
            const void *src = up ? segs[i].host : segs[i].dev;
            hipError_t e = hipMemcpyAsync(dst, src, segs[i].bytes, kind, s);
            if (e != hipSuccess) return e;
        }
        // (the host memory has been read, or written, on return either way)
        return hipStreamSynchronize(s);
    }

    // chunks dealt round robin over the workers
    std::vector<Chunk> lists[kWorkers];
    std::vector<std::thread> th;
    int c = 0;
    try {
        th.reserve(kWorkers);
        for (int i = 0; i < n; ++i)
            for (size_t off = 0; off < segs[i].bytes; off += kChunk, ++c)
                lists[c % kWorkers].push_back({static_cast<char *>(segs[i].dev) + off,
                                               static_cast<char *>(segs[i].host) + off,
                                               segs[i].bytes - off < kChunk ? segs[i].bytes - off : kChunk});
    } catch (const std::exception &) {
        return hipErrorOutOfMemory;
    }
    const int nw = c < kWorkers ? c : kWorkers;

    hipError_t res[kWorkers] = {};
    auto work = [&](int t) {
        res[t] = up ? upload_chunks(s, w_[t].pin, w_[t].ev, lists[t]) : download_chunks(s, w_[t].pin, w_[t].ev, lists[t]);
    };
    // worker 0 on the calling thread (whose device is the context's), the
    // others on threads of their own, made current on the same device
    int inline_from = nw;   // workers the calling thread runs itself
    for (int t = 1; t < nw; ++t) {
        try {
            th.emplace_back([&work, device, t, &res] {
                res[t] = hipSetDevice(device);
                if (res[t] == hipSuccess) work(t);
            });
        } catch (const std::exception &) {
            inline_from = t;
            break;
        }
    }
    work(0);
    for (int t = inline_from; t < nw; ++t) work(t);
    for (std::thread &x : th) x.join();
    for (int t = 0; t < nw; ++t)
        if (res[t] != hipSuccess) return res[t];
    return hipSuccess;
}

}  // namespace sps
