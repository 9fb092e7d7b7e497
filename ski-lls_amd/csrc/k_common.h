// k_common.h -- device-side helpers shared by the HIP kernels (wave64).
#pragma once

#include "skl_internal.h"

namespace skl {

template <class T> void fill2d(long m, long n, T *A, long lda, T v);
template <class Ts, class Td> void copy2d(long m, long n, const Ts *A, long lda, Td *B, long ldb);

__device__ inline double shfl_xor_d(double v, int mask) {
    return __shfl_xor(v, mask, 64);
}

template <class T> __device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
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
