// Device helpers shared by the setup and CG kernels: wave64/workgroup sums,
// the last-block ticket, XCD-aware tile order and ghost index rules.
#pragma once

#include "ps_internal.h"

namespace sps {

// Sum over a wave64 with xor butterflies.  Lane 0 holds the sum in a fixed
// order (the value in other lanes is not used).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Workgroup sum, fixed order; the result is valid in thread 0 only.
__device__ __forceinline__ double block_sum(double v)
{
    __shared__ double red[kThreads / 64];
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        s = red[0];
#pragma unroll
        for (int w = 1; w < kThreads / 64; ++w) s += red[w];
    }
    return s;
}

// Thread 0 publishes this block's partial sum(s); returns true in every
// thread of the block that finished last.  The ticket is reset by that
// block (finish_ticket).  __threadfence() is an agent-scope release/acquire,
// which on gfx950 writes back / invalidates the per-XCD L2 so that the last
// block sees partials written on other XCDs.
__device__ __forceinline__ bool publish_and_ticket(double *partial, double v, double *partial2, double v2,
                                                   unsigned *ticket)
{
    __shared__ int last;
    if (threadIdx.x == 0) {
        if (partial) __hip_atomic_store(partial + blockIdx.x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (partial2) __hip_atomic_store(partial2 + blockIdx.x, v2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const unsigned t = atomicAdd(ticket, 1u);
        last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

__device__ __forceinline__ void finish_ticket(unsigned *ticket)
{
    if (threadIdx.x == 0) __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of n partials in a fixed order (thread-strided, then block_sum).
// The order depends only on n, so a problem's result does not depend on
// what else runs on the device.
__device__ __forceinline__ double sum_partials(const double *partial, int n)
{
    double s = 0.0;
    for (int t = threadIdx.x; t < n; t += kThreads)
        s += __hip_atomic_load(partial + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return block_sum(s);
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8).  Map them so that each XCD walks one contiguous range of
// tiles: neighbouring tiles (which share halo planes) then share an L2.
__device__ __forceinline__ int xcd_tile(int b, int n)
{
    const int per = n / kNumXcd, rem = n % kNumXcd;
    const int xcd = b % kNumXcd, local = b / kNumXcd;
    return xcd * per + min(xcd, rem) + local;
}

// Neighbour index along one direction with the reference's ghost rule:
// mirror (ghost = first interior cell, 9_grad.f90:531-579) or periodic.
__device__ __forceinline__ int nb_lo(int i, int n, int periodic) { return i > 0 ? i - 1 : (periodic ? n - 1 : 0); }
__device__ __forceinline__ int nb_hi(int i, int n, int periodic) { return i < n - 1 ? i + 1 : (periodic ? 0 : n - 1); }

// Halo-box index (Fortran 0..n+1) -> interior cell (0-based) under the same
// rule.  Applied per direction this reproduces the reference's ghost fills
// done in x, then y, then z order (9_grad.f90:594-643).
__device__ __forceinline__ int box_to_cell(int I, int n, int periodic)
{
    if (I == 0) return periodic ? n - 1 : 0;
    if (I == n + 1) return periodic ? 0 : n - 1;
    return I - 1;
}

}  // namespace sps
