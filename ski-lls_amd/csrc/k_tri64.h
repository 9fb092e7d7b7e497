// k_tri64.h -- 64 x 64 FP64 products and the upper-triangular inverse in LDS
// (one 256-thread workgroup), shared by the panel and Cholesky kernels.
#pragma once

#include "k_common.h"

namespace skl {

constexpr int CQ_T = 256;
constexpr int CQ_LD = 65;

// Fills 64 x 64 LDS tiles (ld CQ_LD) from fa(row, col) and fb(row, col),
// then zeroes the entries where ka(row, col) / kb(row, col) is false.  Each
// thread's loads of both tiles are issued before the first is waited on, so
// the fill costs one memory latency rather than one per loop trip (a loop of
// load-then-store pays the latency sixteen times).  The loaders must read a
// valid address for every (row, col) -- clamp the indices -- and the masking
// is a separate LDS pass after a barrier: a select on the loaded value lets
// the compiler put each load behind a branch of its own and wait on it there.
// All CQ_T threads call it.
template <class FA, class KA, class FB, class KB>
__device__ inline void tile64_fill2(double *da, FA fa, KA ka, double *db, FB fb, KB kb) {
    constexpr int NQ = 64 * 64 / CQ_T;
    double va[NQ], vb[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = threadIdx.x + q * CQ_T;
        va[q] = fa(e & 63, e >> 6);
        if (db) vb[q] = fb(e & 63, e >> 6);
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = threadIdx.x + q * CQ_T;
        da[(e & 63) + (e >> 6) * CQ_LD] = va[q];
        if (db) db[(e & 63) + (e >> 6) * CQ_LD] = vb[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = threadIdx.x + q * CQ_T, r = e & 63, c = e >> 6;
        if (!ka(r, c)) da[r + c * CQ_LD] = 0.0;
        if (db && !kb(r, c)) db[r + c * CQ_LD] = 0.0;
    }
}
template <class FA, class KA> __device__ inline void tile64_fill(double *da, FA fa, KA ka) {
    tile64_fill2(da, fa, ka, (double *)nullptr, fa, ka);
}

// C = A B for 64 x 64 matrices in LDS (ld CQ_LD) on the matrix cores: wave w
// of the CQ_T threads forms the 32 x 32 quadrant (32 (w & 1), 32 (w >> 1)) as
// 2 x 2 blocks of v_mfma_f64_16x16x4f64 (A(i, k) from lane i + 16 k, B(k, j)
// from lane j + 16 k; accumulator element r of lane l holds row l / 16 + 4 r,
// column l % 16 of its block).  Each entry goes to store(row, col, value).
// No barrier inside: A and B must be ready, and the caller synchronises
// before anything store wrote is read.
template <class F> __device__ inline void mm64_lds(const double *A, const double *B, F store) {
    typedef double acc4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int qr = (w & 1) * 32, qc = (w >> 1) * 32;
    const int li = lane & 15, lk = lane >> 4;
    acc4 c00 = {0, 0, 0, 0}, c01 = c00, c10 = c00, c11 = c00;
#pragma unroll 4
    for (int k0 = 0; k0 < 64; k0 += 4) {
        const double a0 = A[(qr + li) + (k0 + lk) * CQ_LD];
        const double a1 = A[(qr + 16 + li) + (k0 + lk) * CQ_LD];
        const double b0 = B[(k0 + lk) + (qc + li) * CQ_LD];
        const double b1 = B[(k0 + lk) + (qc + 16 + li) * CQ_LD];
        c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, c00, 0, 0, 0);
        c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, c01, 0, 0, 0);
        c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, c10, 0, 0, 0);
        c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, c11, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = lk + 4 * r, col = li;
        store(qr + row, qc + col, c00[r]);
        store(qr + row, qc + 16 + col, c01[r]);
        store(qr + 16 + row, qc + col, c10[r]);
        store(qr + 16 + row, qc + 16 + col, c11[r]);
    }
}

// X = R^-1 for the upper triangular 64 x 64 R in LDS (ld CQ_LD; rows and
// columns past the panel width hold the identity), by recursive doubling:
// the 8 x 8 diagonal blocks by back substitution in registers, one thread per
// column (reciprocals of the pivots from rcp_nr), then three merge levels
// with X(A, C) = -X_A R(A, C) X_C (the 8 x 8 blocks' level on the vector
// units, the 16 and 32 levels on the matrix cores).  Every inner product runs
// over the whole block with a fixed length (the triangles' zeros included),
// so its loads are issued together rather than one dependent LDS round trip
// per term.
// Y is LDS scratch of the same shape.  All CQ_T threads call it; it ends
// with a barrier.
template <int SZ> __device__ inline void tri_inv_merge(const double *Rs, double *X, double *Y) {
    constexpr int PER = SZ * SZ, NE = (64 / (2 * SZ)) * PER;
    const int tid = threadIdx.x;
    // Y(A, C) = R(A, C) X_C
#pragma unroll
    for (int e0 = 0; e0 < NE; e0 += CQ_T) {
        const int e = e0 + tid;
        const int pr = e / PER, w = e % PER, i = w % SZ, j = w / SZ;
        const int a0 = pr * 2 * SZ, c0 = a0 + SZ;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < SZ; ++k)
            acc = fma(Rs[(a0 + i) + (c0 + k) * CQ_LD], X[(c0 + k) + (c0 + j) * CQ_LD], acc);
        Y[(a0 + i) + (c0 + j) * CQ_LD] = acc;
    }
    __syncthreads();
    // X(A, C) = -X_A Y(A, C)
#pragma unroll
    for (int e0 = 0; e0 < NE; e0 += CQ_T) {
        const int e = e0 + tid;
        const int pr = e / PER, w = e % PER, i = w % SZ, j = w / SZ;
        const int a0 = pr * 2 * SZ, c0 = a0 + SZ;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < SZ; ++k)
            acc = fma(X[(a0 + i) + (a0 + k) * CQ_LD], Y[(a0 + k) + (c0 + j) * CQ_LD], acc);
        X[(a0 + i) + (c0 + j) * CQ_LD] = -acc;
    }
    __syncthreads();
}

// The same merge for SZ = 16, 32 on the matrix cores: the (A, C) blocks in
// 16 x 16 tiles, one tile per wave (SZ / 8 tiles), SZ / 4 steps of
// v_mfma_f64_16x16x4f64 each (operand and accumulator layout as in mm64_lds).
template <int SZ> __device__ inline void tri_inv_merge_mfma(const double *Rs, double *X, double *Y) {
    typedef double acc4 __attribute__((ext_vector_type(4)));
    constexpr int TPB = SZ / 16, NT = (64 / (2 * SZ)) * TPB * TPB;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const bool mine = w < NT;
    const int pr = w / (TPB * TPB), ti = (w % (TPB * TPB)) % TPB, tj = (w % (TPB * TPB)) / TPB;
    const int a0 = pr * 2 * SZ, c0 = a0 + SZ;
    const int r0 = a0 + 16 * ti, q0 = c0 + 16 * tj;
    if (mine) {
        // Y(A, C) = R(A, C) X_C
        acc4 c = {0, 0, 0, 0};
#pragma unroll
        for (int k0 = 0; k0 < SZ; k0 += 4)
            c = __builtin_amdgcn_mfma_f64_16x16x4f64(Rs[(r0 + li) + (c0 + k0 + lk) * CQ_LD],
                                                     X[(c0 + k0 + lk) + (q0 + li) * CQ_LD], c, 0,
                                                     0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) Y[(r0 + lk + 4 * r) + (q0 + li) * CQ_LD] = c[r];
    }
    __syncthreads();
    if (mine) {
        // X(A, C) = -X_A Y(A, C)
        acc4 c = {0, 0, 0, 0};
#pragma unroll
        for (int k0 = 0; k0 < SZ; k0 += 4)
            c = __builtin_amdgcn_mfma_f64_16x16x4f64(X[(r0 + li) + (a0 + k0 + lk) * CQ_LD],
                                                     Y[(a0 + k0 + lk) + (q0 + li) * CQ_LD], c, 0,
                                                     0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) X[(r0 + lk + 4 * r) + (q0 + li) * CQ_LD] = -c[r];
    }
    __syncthreads();
}

__device__ inline void tri_inv_upper_lds(const double *Rs, double *X, double *Y) {
    static_assert(CQ_T == 256, "the merge levels assume 256 threads");
    const int tid = threadIdx.x;
    for (int e = tid; e < 64 * 64; e += CQ_T) X[(e & 63) + (e >> 6) * CQ_LD] = 0.0;
    __syncthreads();
    if (tid < 64) {
        const int c = tid, o = c & ~7, cl = c & 7;
        double rb[8][8], ri[8], x[8];
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int l = i; l < 8; ++l) rb[i][l] = Rs[(o + i) + (o + l) * CQ_LD];
#pragma unroll
        for (int i = 0; i < 8; ++i) ri[i] = rcp_nr(rb[i][i]);
#pragma unroll
        for (int i = 7; i >= 0; --i) {
            double acc = 0.0;
#pragma unroll
            for (int l = i + 1; l < 8; ++l) acc = fma(rb[i][l], x[l], acc);
            x[i] = i > cl ? 0.0 : (i == cl ? ri[i] : -acc * ri[i]);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) X[(o + i) + c * CQ_LD] = x[i];
    }
    __syncthreads();
    tri_inv_merge<8>(Rs, X, Y);
    tri_inv_merge_mfma<16>(Rs, X, Y);
    tri_inv_merge_mfma<32>(Rs, X, Y);
}


} // namespace skl
