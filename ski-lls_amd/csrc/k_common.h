// k_common.h -- device-side helpers shared by the HIP kernels (wave64).
#pragma once

#include "skl_internal.h"

#include <type_traits>

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

// 1/sqrt(d) and 1/x from the hardware estimates with Newton steps (two in
// FP64, one in FP32; quadratic convergence from the estimate's 2^-22 or
// better): a chain of a few fused multiply-adds instead of the IEEE
// square-root and division sequences on a step's critical path.  Within an
// ulp or two of the correctly rounded values (d > 0, x != 0 finite).
__device__ inline double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
    y = y * fma(-0.5 * d * y, y, 1.5);
    y = y * fma(-0.5 * d * y, y, 1.5);
    return y;
}
__device__ inline double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = y * fma(-x, y, 2.0);
    y = y * fma(-x, y, 2.0);
    return y;
}
__device__ inline float rsqrt_nr(float d) {
    float y = __builtin_amdgcn_rsqf(d);
    return y * fmaf(-0.5f * d * y, y, 1.5f);
}
__device__ inline float rcp_nr(float x) {
    float y = __builtin_amdgcn_rcpf(x);
    return y * fmaf(-x, y, 2.0f);
}
// sqrt(x) for x >= 0 as x / sqrt(x) (0 for 0)
template <class T> __device__ inline T sqrt_nr(T x) { return x > T(0) ? x * rsqrt_nr(x) : T(0); }

// Division by a run-time divisor d (1 <= d < 2^16) of dividends x < 2^16 as
// one multiply-high: q = (x m) >> 32 with m = ceil(2^32 / d).  m d = 2^32 + e
// with 0 <= e < d, so x m / 2^32 = x / d + x e / (d 2^32), and the second term
// stays below 1 / d (x d < 2^32): the floor is exact.  m comes from a
// correctly rounded double quotient, which stays below the next integer when
// d is not a power of two (2^32 / d is then at least 2^-48 of itself away from
// it).  Build once per divisor; each div/mod then costs a 32 x 64-bit product
// instead of the ~20-instruction integer division sequence.
struct FastDiv {
    unsigned d;
    unsigned long long m;
    __device__ explicit FastDiv(unsigned d_)
        : d(d_), m((unsigned long long)ceil(4294967296.0 / (double)d_)) {}
    __device__ unsigned div(unsigned x) const {
        return (unsigned)(((unsigned long long)x * m) >> 32);
    }
    __device__ unsigned mod(unsigned x) const { return x - div(x) * d; }
};

// Compile-time loop: f(std::integral_constant<int, i>{}) for i in [I0, I1).
template <int I0, int I1, class F> __device__ inline void static_for(F &&f) {
    if constexpr (I0 < I1) {
        f(std::integral_constant<int, I0>{});
        static_for<I0 + 1, I1>(f);
    }
}

// acc += v * x where v is lane R % 16 of the caller's row of sixteen lanes of
// src (DPP row_newbcast on the FMA's first operand: lane l of src holding
// entry 16 c + l % 16 of a vector, every lane gets entry R from its own row).
// A DPP source must not be written by the two VALU instructions before the
// read: callers fence src (dpp_src_fence) after writing it.
template <int R> __device__ inline void bcast16_fma(double &acc, double src, double x) {
    asm volatile("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
                 : "+v"(acc)
                 : "v"(src), "v"(x), "i"(R & 15));
}
template <int N> __device__ inline void dpp_src_fence(double (&v)[N]) {
    static_assert(N <= 5, "fence up to five sources");
    if constexpr (N == 4)
        asm volatile("s_nop 1" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]));
    else
        asm volatile("s_nop 1" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]));
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

// Lane moves within rows of 16 (DPP, VALU ops without an LDS round trip):
// quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror, row_mirror.  After the
// four the value combined over its row of 16 sits in every lane of the row.
template <int CTRL> __device__ inline int dpp_move(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL> __device__ inline float dpp_move(float v) {
    return __int_as_float(dpp_move<CTRL>(__float_as_int(v)));
}
template <int CTRL> __device__ inline double dpp_move(double v) {
    const int lo = dpp_move<CTRL>(__double2loint(v)), hi = dpp_move<CTRL>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;

// max over the wave, the same value in every lane (rows by DPP, then the four
// row results by readlane)
template <class T> __device__ inline T wave_max_bcast(T v) {
    v = fmax(v, dpp_move<DPP_XOR1>(v));
    v = fmax(v, dpp_move<DPP_XOR2>(v));
    v = fmax(v, dpp_move<DPP_HALF_MIRROR>(v));
    v = fmax(v, dpp_move<DPP_MIRROR>(v));
    return fmax(fmax(readlane(v, 0), readlane(v, 16)), fmax(readlane(v, 32), readlane(v, 48)));
}
// sum over the wave, the same value in every lane (same lane moves as
// wave_max_bcast; a fixed order, so equal inputs give equal sums everywhere)
template <class T> __device__ inline T wave_sum_bcast(T v) {
    v += dpp_move<DPP_XOR1>(v);
    v += dpp_move<DPP_XOR2>(v);
    v += dpp_move<DPP_HALF_MIRROR>(v);
    v += dpp_move<DPP_MIRROR>(v);
    return (readlane(v, 0) + readlane(v, 16)) + (readlane(v, 32) + readlane(v, 48));
}
__device__ inline int wave_min_bcast(int k) {
    k = min(k, dpp_move<DPP_XOR1>(k));
    k = min(k, dpp_move<DPP_XOR2>(k));
    k = min(k, dpp_move<DPP_HALF_MIRROR>(k));
    k = min(k, dpp_move<DPP_MIRROR>(k));
    return min(min(__builtin_amdgcn_readlane(k, 0), __builtin_amdgcn_readlane(k, 16)),
               min(__builtin_amdgcn_readlane(k, 32), __builtin_amdgcn_readlane(k, 48)));
}

// Lane of the largest v over the wave's candidates (lanes with v >= 0; v < 0
// marks a lane that is not one, so NaN must be mapped below 0 by the caller),
// ties to the smallest key; -1 when no lane is a candidate.  Wave uniform.
// The same choice as a butterfly over (v, key) pairs, from one max, one
// ballot and, only on ties, one min over the keys.
template <class T> __device__ inline int wave_argmax(T v, int key) {
    const T m = wave_max_bcast(v);
    if (m < T(0)) return -1;
    const unsigned long long hit = __ballot(v == m);
    if (__popcll(hit) == 1) return __ffsll((long long)hit) - 1;
    const int kmin = wave_min_bcast(v == m ? key : 0x7fffffff);
    return __ffsll((long long)__ballot(v == m && key == kmin)) - 1;
}

__device__ inline int wave_max_int(int k) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) k = max(k, __shfl_xor(k, o, 64));
    return k;
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
