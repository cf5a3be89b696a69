// Device helpers shared by the setup and CG kernels: wave64/workgroup sums,
// partial-sum totals, XCD-aware tile order and ghost index rules.
#pragma once

#include "ps_internal.h"

namespace sps {

// Pointers in global memory.  The arrays the kernels work on are reached
// through device-side problem descriptors, so to the compiler they are
// generic pointers, which it accesses with flat instructions: those also
// count on the LDS/scalar-load counter (every wait for them waits for
// pending scalar loads too) and take no scalar base register.  Every such
// array is device memory; accessing it through these pointers gives global
// instructions.
template <class T>
using gptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr<T> glob(T *p)
{
    return (gptr<T>)p;
}

// Read-only device memory: arrays that no kernel writes while the reading
// kernel runs (tables built by set-up kernels).  A wave-uniform load through
// such a pointer is a scalar load, which waits on its own counter instead of
// behind the wave's pending vector loads.
template <class T>
using cptr = __attribute__((address_space(4))) const T *;
template <class T>
__device__ __forceinline__ cptr<T> konst(const T *p)
{
    return (cptr<T>)p;
}

// One double moved between lanes of each row of 16 (DPP row rotate right
// by N: lane l of a row gets lane (l - N) mod 16 of the same row), both
// 32-bit halves.
template <int N>
__device__ __forceinline__ double row_ror(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0x120 + N, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0x120 + N, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double read_lane(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over a wave64, returned to every lane (a wave-uniform value).  Each
// row of 16 lanes is summed with DPP rotations (lane 0 of the row adds in a
// fixed order), then the four rows' sums are added in order.  The bits
// depend only on the 64 inputs' lane positions, so every wave that sums
// the same values in the same lanes gets the same total.
__device__ __forceinline__ double wave_sum(double v)
{
    v += row_ror<8>(v);
    v += row_ror<4>(v);
    v += row_ror<2>(v);
    v += row_ror<1>(v);
    return ((read_lane(v, 0) + read_lane(v, 16)) + read_lane(v, 32)) + read_lane(v, 48);
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
// returned to every thread.  Each wave sums all n itself (lane l the
// partials l, l + 64, ..., then wave_sum), so no wave waits for another and
// every wave of every workgroup and kernel that sums the same partials gets
// the same bits: a problem's result does not depend on what else runs.
__device__ __forceinline__ double total_of(const double *partial, int n)
{
    const gptr<const double> p = glob(partial);
    double s = 0.0;
    for (int t = threadIdx.x & 63; t < n; t += 64) s += p[t];
    return wave_sum(s);
}

// The same total summed once per workgroup (every thread reads some of the
// partials) and handed to its waves through LDS: for kernels whose waves
// would otherwise each read many partials.  Every workgroup gets the same bits; a kernel
// uses one of the two forms for a given set of partials.
__device__ __forceinline__ double block_total_of(const double *partial, int n)
{
    __shared__ double tot;
    const gptr<const double> p = glob(partial);
    double s = 0.0;
    for (int t = threadIdx.x; t < n; t += kThreads) s += p[t];
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

// Totals of n partial 4-vectors (m[4t + c]), returned to every thread.  The
// workgroup sums them once: each thread the partials t = tid, tid + 256,
// ..., each wave its lanes' sums, then every thread adds the four waves'
// sums in wave order (through LDS).  Every workgroup and kernel that sums
// the same partials gets the same bits.  Called by all threads of the
// workgroup together.
__device__ __forceinline__ void total4_of(const double *m, int n, double out[4])
{
    __shared__ double part[kThreads / 64][4];
    const gptr<const double> p = glob(m);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int t = threadIdx.x; t < n; t += kThreads)
#pragma unroll
        for (int c = 0; c < 4; ++c) v[c] += p[4 * t + c];
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] = wave_sum(v[c]);
    __syncthreads();   // the previous call's readers are done with part
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < 4; ++c) part[threadIdx.x >> 6][c] = v[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        double s = part[0][c];
#pragma unroll
        for (int w = 1; w < kThreads / 64; ++w) s += part[w][c];
        out[c] = s;
    }
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
