// k_sparse.hip -- sparse-input kernels: hashing sketch S A of a CSC matrix,
// CSC -> CSR transpose, index narrowing and CSR products.
#include <algorithm>
#include <climits>

#include "k_common.h"

namespace skl {

// ---------------------------------------------------------------------------
// Device exclusive scan of int32 counts (n entries) into out (n + 1 entries).
// ---------------------------------------------------------------------------
constexpr int SCAN_B = 1024;

__global__ __launch_bounds__(SCAN_B) void k_scan_block(long n, const int32_t *in, int32_t *out,
                                                       int32_t *bsum) {
    __shared__ int32_t s[SCAN_B];
    long i = (long)blockIdx.x * SCAN_B + threadIdx.x;
    int32_t v = (i < n) ? in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_B; o <<= 1) {
        int32_t t = (threadIdx.x >= (unsigned)o) ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - v; // exclusive
    if (threadIdx.x == SCAN_B - 1) bsum[blockIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(SCAN_B) void k_scan_sums(long nb, int32_t *bsum, int32_t *total) {
    // sequential carry over chunks of SCAN_B block sums
    __shared__ int32_t s[SCAN_B];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (long base = 0; base < nb; base += SCAN_B) {
        long i = base + threadIdx.x;
        int32_t v = (i < nb) ? bsum[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < SCAN_B; o <<= 1) {
            int32_t t = (threadIdx.x >= (unsigned)o) ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = s[threadIdx.x] - v + carry;
        __syncthreads();
        if (threadIdx.x == SCAN_B - 1) carry += s[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_add(long n, int32_t *out, const int32_t *bsum, const int32_t *total) {
    long i = (long)blockIdx.x * SCAN_B + threadIdx.x;
    if (i < n) out[i] += bsum[blockIdx.x];
    if (i == 0) out[n] = *total;
}

void exclusive_scan(long n, const int32_t *in, int32_t *out, Arena &ar) {
    long nb = (n + SCAN_B - 1) / SCAN_B;
    int32_t *bsum = ar.alloc_n<int32_t>((size_t)std::max<long>(nb, 1) + 1);
    hipStream_t s = ctx().stream;
    if (nb > 0) k_scan_block<<<(unsigned)nb, SCAN_B, 0, s>>>(n, in, out, bsum);
    k_scan_sums<<<1, SCAN_B, 0, s>>>(nb, bsum, bsum + nb);
    k_scan_add<<<(unsigned)std::max<long>(nb, 1), SCAN_B, 0, s>>>(n, out, bsum, bsum + nb);
    SKL_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
__global__ void k_narrow(long n, const long *in, int32_t *out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (int32_t)in[i];
}
void narrow_indices(long n, const long *in, int32_t *out) {
    if (n <= 0) return;
    k_narrow<<<(unsigned)((n + 255) / 256), 256, 0, ctx().stream>>>(n, in, out);
    SKL_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// CSC (n x d) -> CSR with each row's entries in increasing column order.
// ---------------------------------------------------------------------------
__global__ void k_row_count(long nnz, const int32_t *rowind, int32_t *cnt) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) atomicAdd(&cnt[rowind[p]], 1);
}
// Entries land in their row's slots in arrival order (atomics); each keeps its
// CSC position so that the ranking below can order duplicates.
__global__ __launch_bounds__(256) void k_csc_scatter(long d, const int32_t *colptr,
                                                     const int32_t *rowind, const double *val,
                                                     int32_t *cur, int32_t *ccol, double *cval,
                                                     int32_t *csrc) {
    long j = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= d) return;
    for (int p = colptr[j] + (threadIdx.x & 63); p < colptr[j + 1]; p += 64) {
        int pos = atomicAdd(&cur[rowind[p]], 1);
        ccol[pos] = (int32_t)j;
        cval[pos] = val[p];
        csrc[pos] = p;
    }
}
// One wave per row: every entry goes to the position given by its rank in the
// row: the number of entries with a smaller column, or with the same column
// (a duplicate (i, j) of the CSC input) and a smaller CSC position.  The
// result depends on the input only, not on the order the scatter's atomics
// happened to take.  The keys are read 64 at a time into registers and
// broadcast with readlane; the positions are compared only in rows that hold
// a duplicate.
__global__ __launch_bounds__(256) void k_row_sort_rank(long n, const int32_t *__restrict__ ptr,
                                                       const int32_t *__restrict__ tcol,
                                                       const double *__restrict__ tval,
                                                       const int32_t *__restrict__ tsrc,
                                                       int32_t *__restrict__ ccol,
                                                       double *__restrict__ cval) {
    const long i = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const int lane = threadIdx.x & 63;
    const int p0 = ptr[i], L = ptr[i + 1] - p0;
    for (int e0 = 0; e0 < L; e0 += 64) {
        const int e = e0 + lane;
        const int key = e < L ? tcol[p0 + e] : INT_MAX;
        const double v = e < L ? tval[p0 + e] : 0.0;
        int rank = 0, eq = 0;
        for (int q0 = 0; q0 < L; q0 += 64) {
            const int kl = q0 + lane < L ? tcol[p0 + q0 + lane] : INT_MAX;
            const int qn = min(64, L - q0);
            for (int q = 0; q < qn; ++q) {
                const int kq = __builtin_amdgcn_readlane(kl, q);
                rank += kq < key;
                eq += (kq == key) && (q0 + q != e);
            }
        }
        if (__ballot(e < L && eq > 0)) {
            // duplicates in this row: order equal columns by CSC position
            const int src = e < L ? tsrc[p0 + e] : INT_MAX;
            for (int q0 = 0; q0 < L; q0 += 64) {
                const int kl = q0 + lane < L ? tcol[p0 + q0 + lane] : INT_MAX;
                const int sl = q0 + lane < L ? tsrc[p0 + q0 + lane] : INT_MAX;
                const int qn = min(64, L - q0);
                for (int q = 0; q < qn; ++q) {
                    const int kq = __builtin_amdgcn_readlane(kl, q);
                    const int sq = __builtin_amdgcn_readlane(sl, q);
                    rank += (kq == key) && (sq < src);
                }
            }
        }
        if (e < L) {
            ccol[p0 + rank] = key;
            cval[p0 + rank] = v;
        }
    }
}

void csc_to_csr(long n, long d, long nnz, const int32_t *colptr, const int32_t *rowind,
                const double *val, int32_t *rptr, int32_t *ccol, double *cval, Arena &ar) {
    hipStream_t s = ctx().stream;
    int32_t *cnt = ar.alloc_n<int32_t>((size_t)n + 1);
    fill2d<int32_t>(n, 1, cnt, n, 0);
    if (nnz > 0) k_row_count<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(nnz, rowind, cnt);
    exclusive_scan(n, cnt, rptr, ar);
    int32_t *cur = ar.alloc_n<int32_t>((size_t)n + 1);
    SKL_HIP(hipMemcpyAsync(cur, rptr, sizeof(int32_t) * (size_t)(n + 1), hipMemcpyDeviceToDevice, s));
    // scatter in arrival order into scratch, then place each entry by its rank
    const size_t mark = ar.used();
    int32_t *tcol = ar.alloc_n<int32_t>((size_t)std::max<long>(nnz, 1));
    double *tval = ar.alloc_n<double>((size_t)std::max<long>(nnz, 1));
    int32_t *tsrc = ar.alloc_n<int32_t>((size_t)std::max<long>(nnz, 1));
    if (d > 0)
        k_csc_scatter<<<(unsigned)((d + 3) / 4), 256, 0, s>>>(d, colptr, rowind, val, cur, tcol,
                                                              tval, tsrc);
    k_row_sort_rank<<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, rptr, tcol, tval, tsrc, ccol, cval);
    SKL_CHECK_LAUNCH();
    ar.rewind(mark);
}

// y = beta y + alpha A x  (CSR, one thread-group of 8 lanes per row)
__global__ __launch_bounds__(256) void k_csr_gemv(long n, const int32_t *ptr, const int32_t *col,
                                                  const double *val, const double *x,
                                                  double alpha, double beta, double *y) {
    const int sub = threadIdx.x & 7;
    long i = (long)blockIdx.x * 32 + (threadIdx.x >> 3);
    double s = 0.0;
    if (i < n)
        for (int p = ptr[i] + sub; p < ptr[i + 1]; p += 8) s += val[p] * x[col[p]];
#pragma unroll
    for (int o = 4; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (sub == 0 && i < n) y[i] = (beta == 0.0 ? 0.0 : beta * y[i]) + alpha * s;
}
void csr_gemv(long n, const int32_t *ptr, const int32_t *col, const double *val, const double *x,
              double alpha, double beta, double *y) {
    if (n <= 0) return;
    k_csr_gemv<<<(unsigned)((n + 31) / 32), 256, 0, ctx().stream>>>(n, ptr, col, val, x, alpha, beta, y);
    SKL_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// Sparse sketch SA = S A (A CSC n x d, S: k targets per source row).
// One workgroup per column: counting sort of the column's k*nnz contributions
// by target bucket (LDS histogram + scan), then each bucket's contributions
// re-ordered by position in the column and summed in that order -- the order
// of CXSparse cs_multiply's scatter (src/sketching.cpp:111).
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_sparse_sketch(long m, int k, const int32_t *__restrict__ srows,
                                                       const double *__restrict__ svals,
                                                       const int32_t *__restrict__ colptr,
                                                       const int32_t *__restrict__ rowind,
                                                       const double *__restrict__ val,
                                                       int32_t *__restrict__ slots,
                                                       T *__restrict__ SA, long ldsa) {
#pragma clang fp contract(off)
    extern __shared__ int32_t shi[];
    int32_t *cnt = shi;
    int32_t *cur = shi + m;
    __shared__ int32_t part[256];
    const long c = blockIdx.x;
    const int p0 = colptr[c], p1 = colptr[c + 1];
    const long K = (long)(p1 - p0) * k;
    for (long h = threadIdx.x; h < m; h += 256) cnt[h] = 0;
    __syncthreads();
    for (long e = threadIdx.x; e < K; e += 256) {
        int p = p0 + (int)(e / k), t = (int)(e % k);
        atomicAdd(&cnt[srows[(long)rowind[p] * k + t]], 1);
    }
    __syncthreads();
    // exclusive scan cnt -> cur over m entries
    const long per = (m + 255) / 256;
    const long hb = threadIdx.x * per, he = min(m, hb + per);
    int32_t loc = 0;
    for (long h = hb; h < he; ++h) loc += cnt[h];
    part[threadIdx.x] = loc;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t acc = 0;
        for (int i = 0; i < 256; ++i) {
            int32_t v = part[i];
            part[i] = acc;
            acc += v;
        }
    }
    __syncthreads();
    int32_t run = part[threadIdx.x];
    for (long h = hb; h < he; ++h) {
        cur[h] = run;
        run += cnt[h];
    }
    __syncthreads();
    int32_t *sl = slots + (long)p0 * k;
    for (long e = threadIdx.x; e < K; e += 256) {
        int p = p0 + (int)(e / k), t = (int)(e % k);
        int h = srows[(long)rowind[p] * k + t];
        int pos = atomicAdd(&cur[h], 1);
        sl[pos] = (int32_t)e;
    }
    __syncthreads();
    for (long h = threadIdx.x; h < m; h += 256) {
        int e1 = cur[h], e0 = e1 - cnt[h];
        for (int a = e0 + 1; a < e1; ++a) { // order by position in the column
            int32_t key = sl[a];
            int b = a - 1;
            while (b >= e0 && sl[b] > key) {
                sl[b + 1] = sl[b];
                --b;
            }
            sl[b + 1] = key;
        }
        T acc = 0;
        for (int a = e0; a < e1; ++a) {
            int e = sl[a];
            int p = p0 + e / k, t = e % k;
            acc = acc + (T)val[p] * (T)svals[(long)rowind[p] * k + t];
        }
        SA[h + c * ldsa] = acc;
    }
}

template <class T>
void sparse_sketch_t(long m, long d, long k, const int32_t *srows, const double *svals,
                     const int32_t *colptr, const int32_t *rowind, const double *val, long nnz,
                     T *SA, long ldsa, Arena &ar) {
    SKL_REQUIRE(m <= 16384, SKL_E_UNSUPPORTED, "sparse sketch supports at most 16384 sketch rows");
    int32_t *slots = ar.alloc_n<int32_t>((size_t)std::max<long>(nnz * k, 1));
    size_t sh = sizeof(int32_t) * 2 * (size_t)m;
    if (sh > 65536)
        SKL_HIP(hipFuncSetAttribute((const void *)k_sparse_sketch<T>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    if (d > 0)
        k_sparse_sketch<T><<<(unsigned)d, 256, sh, ctx().stream>>>(m, (int)k, srows, svals, colptr,
                                                                   rowind, val, slots, SA, ldsa);
    SKL_CHECK_LAUNCH();
}
template void sparse_sketch_t<double>(long, long, long, const int32_t *, const double *,
                                      const int32_t *, const int32_t *, const double *, long,
                                      double *, long, Arena &);
template void sparse_sketch_t<float>(long, long, long, const int32_t *, const double *,
                                     const int32_t *, const int32_t *, const double *, long,
                                     float *, long, Arena &);

} // namespace skl
