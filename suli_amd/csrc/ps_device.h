// Device helpers shared by the setup and CG kernels: wave64/workgroup sums,
// partial-sum totals, XCD-aware tile order and ghost index rules.
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

// Total of n per-workgroup partial sums written by an earlier kernel,
// returned to every thread.  The summation order depends only on n, so every
// workgroup (and every kernel) that sums the same partials gets the same
// bits, and a problem's result does not depend on what else runs.
__device__ __forceinline__ double total_of(const double *partial, int n)
{
    __shared__ double tot;
    double s = 0.0;
    for (int t = threadIdx.x; t < n; t += kThreads) s += partial[t];
    s = block_sum(s);
    if (threadIdx.x == 0) tot = s;
    __syncthreads();
    const double r = tot;
    __syncthreads();
    return r;
}

// Four workgroup sums at once (fixed order), valid in thread 0 only.
__device__ __forceinline__ void block_sum4(double v[4])
{
    __shared__ double red4[4][kThreads / 64];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = wave_sum(v[c]);
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < 4; ++c) red4[c][threadIdx.x >> 6] = v[c];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double s = red4[c][0];
#pragma unroll
            for (int w = 1; w < kThreads / 64; ++w) s += red4[c][w];
            v[c] = s;
        }
}

// Totals of n partial 4-vectors (m[4t + c]), returned to every thread; the
// order depends only on n.
__device__ __forceinline__ void total4_of(const double *m, int n, double out[4])
{
    __shared__ double tot4[4];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < n; t += kThreads)
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] += m[4 * t + c];
    block_sum4(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < 4; ++c) tot4[c] = v[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) out[c] = tot4[c];
    __syncthreads();
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
