// Host <-> device copies of the host-pointer entry points (host code).
//
// The caller's arrays (SuLi's Fortran module arrays) are pageable memory.
// A pageable hipMemcpyAsync goes through the runtime's own staging at about
// 28 GB/s to the device and 17 GB/s back on MI355X; pinned memory moves at
// about 57 GB/s both ways.  HostPipe stages through pinned chunks of its
// own: kWorkers host threads, each with two kChunk pinned buffers, copy
// their share of the chunks between the caller's memory and the pinned
// buffers with memcpy while the DMA of their previous chunk runs.  The DMAs
// go on the caller's stream (one stream carries the full PCIe rate; the
// host copies are what needs the threads), so they are ordered with the
// work already on it and with what follows.  Transfers below kDirect bytes
// go straight through hipMemcpyAsync.  Every call returns with the host
// memory read (upload) or written (download); the calling thread's current
// device must be the stream's.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace sps {

struct HostSeg {
    void *dev;
    void *host;
    size_t bytes;
};

class HostPipe {
public:
    static constexpr int kWorkers = 4;
    static constexpr size_t kChunk = size_t(4) << 20;
    static constexpr size_t kDirect = size_t(2) << 20;

    HostPipe() = default;
    HostPipe(const HostPipe &) = delete;
    HostPipe &operator=(const HostPipe &) = delete;
    ~HostPipe();

    // Host -> device, after the work already on `s` and before whatever is
    // enqueued on `s` after the call; the host memory has been read when it
    // returns.
    hipError_t upload(int device, hipStream_t s, const HostSeg *segs, int n);
    // Device -> host, after the work already on `s`; the host memory holds
    // the data when it returns.
    hipError_t download(int device, hipStream_t s, const HostSeg *segs, int n);

    // Frees the pinned buffers, streams and events (also done on destruction).
    void release();

private:
    struct Worker {
        char *pin[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};   // the last DMA from or into each buffer
    };
    hipError_t ensure(int device);
    hipError_t run(int device, hipStream_t s, const HostSeg *segs, int n, bool up);

    Worker w_[kWorkers];
    int device_ = -1;
    bool ready_ = false;
};

}  // namespace sps
