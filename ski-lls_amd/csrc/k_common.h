// k_common.h -- device-side helpers shared by the HIP kernels (wave64).
#pragma once

#include "skl_internal.h"

namespace skl {

template <class T> void fill2d(long m, long n, T *A, long lda, T v);
template <class Ts, class Td> void copy2d(long m, long n, const Ts *A, long lda, Td *B, long ldb);

// MI355X dispatches workgroup i of a grid to XCD i mod 8, and each XCD has
// its own L2.  This maps the hardware index to a logical one under which
// consecutive logical workgroups run on the same XCD (XCD x takes a
// contiguous range of the logical indices), so tiles that share cache lines
// (the blocks of one column) meet in one L2 instead of fetching the lines
// once per XCD.  A bijection on [0, nblocks) for any nblocks.
__device__ inline long xcd_local_block(long bid, long nblocks) {
    const long q = nblocks / 8, r = nblocks % 8, x = bid % 8, slot = bid / 8;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + slot;
}

__device__ inline double shfl_xor_d(double v, int mask) {
    return __shfl_xor(v, mask, 64);
}

template <class T> __device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// K independent wave sums, interleaved level by level (same result per value
// as wave_sum)
template <int K, class T> __device__ inline void wave_sum_k(T (&v)[K]) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] += __shfl_xor(v[k], o, 64);
}

// value of v in lane l (l wave-uniform)
__device__ inline double readlane(double v, int l) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)b, l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ inline float readlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Barrier over every workgroup of a launch.  The kernels that use it keep
// their grids far below what the device holds at once (at most a quarter to
// a half of the CUs, one workgroup per CU), so all workgroups are resident
// without a cooperative launch.  ctr is zeroed by the host before the launch;
// `target` starts at 0 in every workgroup.  Agent-scope fences around the
// arrival publish this workgroup's global writes and expose the others'.
//
// The wait is bounded: after `spins` polls it gives up, sets the error word
// *err (kept per host context, read back by the host after the
// factorization) and returns false.  A workgroup that finds *err set while
// waiting gives up at once, so after one failure every workgroup of the grid
// leaves its barriers within a few polls.  The result is uniform over the
// workgroup; a kernel returns as soon as it is false (all its threads
// together), and the kernels that consume indices the failed launch should
// have written check *err before using them.
__device__ inline bool grid_barrier(const GridSync &g, unsigned &target) {
    __shared__ int s_ok;
    __syncthreads();
    target += gridDim.x;
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned old =
            __hip_atomic_fetch_add(g.ctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        if (old + 1u < target) {
            unsigned polls = 0;
            for (;;) {
                if (__hip_atomic_load(g.ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target)
                    break;
                if (polls >= g.spins ||
                    ((polls & 63u) == 63u &&
                     __hip_atomic_load(g.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)) {
                    ok = 0;
                    break;
                }
                ++polls;
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (!ok) __hip_atomic_fetch_or(g.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// true when a grid barrier of an earlier launch gave up (see GridSync)
__device__ inline bool grid_failed(const unsigned *err) {
    return __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}

template <class T> __device__ inline T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Sum over a block of BS threads (multiple of 64); result valid in every thread.
// Fixed summation order: lanes by butterfly, then waves in index order.
template <int BS, class T> __device__ inline T block_sum(T v) {
    __shared__ T red[BS / 64];
    v = wave_sum(v);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    T s = 0;
#pragma unroll
    for (int i = 0; i < BS / 64; ++i) s += red[i];
    return s;
}

template <int BS, class T> __device__ inline T block_max(T v) {
    __shared__ T red[BS / 64];
    v = wave_max(v);
    int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    T s = red[0];
#pragma unroll
    for (int i = 1; i < BS / 64; ++i) s = fmax(s, red[i]);
    return s;
}

template <class T> struct cplx {
    T x, y;
};
template <class T> __device__ __host__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T> __device__ __host__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) {
    return {a.x + b.x, a.y + b.y};
}
template <class T> __device__ __host__ inline cplx<T> csub(cplx<T> a, cplx<T> b) {
    return {a.x - b.x, a.y - b.y};
}

} // namespace skl
