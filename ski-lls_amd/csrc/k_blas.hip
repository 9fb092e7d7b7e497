// k_blas.hip -- dense building blocks for gfx950: MFMA GEMM (FP64 / FP32),
// blocked upper-triangular inverse, norms and small copy kernels.
#include <algorithm>
#include <cmath>

#include "k_common.h"
#include "k_tri64.h"

namespace skl {

// ---------------------------------------------------------------------------
// GEMM on the matrix cores.  64x64 output tile per 256-thread workgroup; each
// wave owns a 32x32 quarter as 2x2 MFMA 16x16 blocks (v_mfma_f64_16x16x4f64 or
// v_mfma_f32_16x16x4f32, K = 4 per instruction).  Operand layout of the
// 16x16x4 instructions: A(i,k) in lane i + 16k, B(k,j) in lane j + 16k.
// Accumulator element r of lane l holds D(row(l, r), l % 16) with
// row = 4 (l/16) + r for f32 and row = l/16 + 4 r for f64 (measured on gfx950).
// ---------------------------------------------------------------------------
template <class T> struct Mfma;
template <> struct Mfma<double> {
    typedef double acc __attribute__((ext_vector_type(4)));
    static __device__ inline acc fma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    typedef float acc __attribute__((ext_vector_type(4)));
    static __device__ inline acc fma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int GBM = 64, GBN = 64, GBK = 16;
constexpr int GS = 8; // super-tile side (tiles) of the XCD-aware order

// kchunk > 0 selects split-K: blockIdx.z covers k in [z kchunk, (z+1) kchunk)
// and writes its raw partial product to C + z sC (alpha, beta applied by the
// reduction); kchunk == 0 is a strided batch over blockIdx.z.
template <class T, bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, T alpha, const T *__restrict__ A,
                                              long lda, const T *__restrict__ B, long ldb, T beta,
                                              T *__restrict__ C, long ldc, long sA, long sB,
                                              long sC, int kchunk, int upper, int stm,
                                              int stn) {
    // upper: only the tiles on or above the diagonal of a square C (A^T A
    // for a symmetric result whose lower triangle is never read)
    int bi = blockIdx.x, bj = blockIdx.y;
    if (stm > 0) {
        // tile order for large grids (stm x stn tiles, 1-D grid): 8 x 8
        // super-tiles of consecutive logical blocks, and the logical blocks of
        // one XCD (workgroup id mod 8) contiguous, so the tiles an XCD runs at
        // once share their A and B panels in its L2; upper: only the
        // super-tiles on or above the diagonal, column by column
        const long L = xcd_local_block(blockIdx.x, gridDim.x);
        const long sup = L / (GS * GS);
        const int w = (int)(L % (GS * GS));
        long SI, SJ;
        if (upper) {
            SJ = (long)((sqrt(8.0 * (double)sup + 1.0) - 1.0) * 0.5);
            while (SJ * (SJ + 1) / 2 > sup) --SJ;
            while ((SJ + 1) * (SJ + 2) / 2 <= sup) ++SJ;
            SI = sup - SJ * (SJ + 1) / 2;
        } else {
            const long sm = (stm + GS - 1) / GS;
            SI = sup % sm;
            SJ = sup / sm;
        }
        bi = (int)(SI * GS + w % GS);
        bj = (int)(SJ * GS + w / GS);
        if (bi >= stm || bj >= stn) return;
    }
    if (upper && bi > bj) return;
    typedef typename Mfma<T>::acc acc_t;
    __shared__ T As[2][GBK][GBM + 1];
    __shared__ T Bs[2][GBK][GBN + 1];
    const long bz = blockIdx.z;
    if (kchunk > 0) {
        const long k0 = bz * (long)kchunk;
        K = (int)min((long)kchunk, (long)K - k0);
        A += TA ? k0 : k0 * lda;
        B += TB ? k0 * ldb : k0;
        C += bz * sC;
    } else {
        A += bz * sA;
        B += bz * sB;
        C += bz * sC;
    }
    const int i0 = bi * GBM, j0 = bj * GBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    acc_t acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    // C is read early so its latency hides behind the products
    T cold[2][2][4];
    const bool rd_c = (beta != T(0));
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int gi = i0 + wm + mi * 16 + Mfma<T>::row(lane, r);
                int gj = j0 + wn + ni * 16 + (lane & 15);
                cold[mi][ni][r] = (rd_c && gi < M && gj < N) ? C[gi + (long)gj * ldc] : T(0);
            }

    T ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int e = tid + 256 * q;
            int ii, kk;
            if (!TA) { ii = e % GBM; kk = e / GBM; }
            else { kk = e % GBK; ii = e / GBK; }
            int gi = i0 + ii, gk = k0 + kk;
            T v = 0;
            if (gi < M && gk < K) v = TA ? A[gk + (long)gi * lda] : A[gi + (long)gk * lda];
            ra[q] = v;
            int jj;
            if (!TB) { kk = e % GBK; jj = e / GBK; }
            else { jj = e % GBN; kk = e / GBN; }
            int gj = j0 + jj;
            gk = k0 + kk;
            T w = 0;
            if (gk < K && gj < N) w = TB ? B[gj + (long)gk * ldb] : B[gk + (long)gj * ldb];
            rb[q] = w;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int e = tid + 256 * q;
            int ii, kk, jj;
            if (!TA) { ii = e % GBM; kk = e / GBM; }
            else { kk = e % GBK; ii = e / GBK; }
            As[buf][kk][ii] = ra[q];
            if (!TB) { kk = e % GBK; jj = e / GBK; }
            else { jj = e % GBN; kk = e / GBN; }
            Bs[buf][kk][jj] = rb[q];
        }
    };
    int nk = (K + GBK - 1) / GBK;
    if (nk > 0) {
        load(0);
        store(0);
    }
    __syncthreads();
    for (int t = 0; t < nk; ++t) {
        int buf = t & 1;
        if (t + 1 < nk) load((t + 1) * GBK);
#pragma unroll
        for (int ks = 0; ks < GBK; ks += 4) {
            T a0 = As[buf][ks + (lane >> 4)][wm + (lane & 15)];
            T a1 = As[buf][ks + (lane >> 4)][wm + 16 + (lane & 15)];
            T b0 = Bs[buf][ks + (lane >> 4)][wn + (lane & 15)];
            T b1 = Bs[buf][ks + (lane >> 4)][wn + 16 + (lane & 15)];
            acc[0][0] = Mfma<T>::fma(a0, b0, acc[0][0]);
            acc[0][1] = Mfma<T>::fma(a0, b1, acc[0][1]);
            acc[1][0] = Mfma<T>::fma(a1, b0, acc[1][0]);
            acc[1][1] = Mfma<T>::fma(a1, b1, acc[1][1]);
        }
        if (t + 1 < nk) store(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int gi = i0 + wm + mi * 16 + Mfma<T>::row(lane, r);
                int gj = j0 + wn + ni * 16 + (lane & 15);
                if (gi < M && gj < N) {
                    T v = alpha * acc[mi][ni][r];
                    C[gi + (long)gj * ldc] = rd_c ? v + beta * cold[mi][ni][r] : v;
                }
            }
}

// C = alpha sum_z W[z] + beta C over S split-K partials (fixed order)
// (the partials are loaded eight at a time, then added in order, so the
// loads of a group are in flight together)
template <class T>
__global__ void k_splitk_reduce(int M, int N, int S, const T *W, T alpha, T beta, T *C, long ldc) {
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)M * N) return;
    int i = (int)(e % M), j = (int)(e / M);
    const long MN = (long)M * N;
    T s = W[e];
    int z = 1;
    for (; z + 8 <= S; z += 8) {
        T w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) w[q] = W[e + (long)(z + q) * MN];
#pragma unroll
        for (int q = 0; q < 8; ++q) s += w[q];
    }
    for (; z < S; ++z) s += W[e + (long)z * MN];
    T *c = C + i + (long)j * ldc;
    *c = (beta == T(0)) ? alpha * s : alpha * s + beta * *c;
}

template <class T>
void gemm(char ta, char tb, long M, long N, long K, T alpha, const T *A, long lda, const T *B,
          long ldb, T beta, T *C, long ldc, long batch, long sA, long sB, long sC, bool upper) {
    if (M <= 0 || N <= 0 || batch <= 0) return;
    const long tm = (M + GBM - 1) / GBM, tn = (N + GBN - 1) / GBN;
    hipStream_t s = ctx().stream;
    bool TA = (ta == 'T' || ta == 't'), TB = (tb == 'T' || tb == 't');
    // Few output tiles over a long K (V^T C, G V, V^T V in the QR): split K so
    // the launch fills the chip; partials go to the context's split-K scratch
    // (stream-ordered, so the next GEMM may reuse it).
    // Target: about four workgroups per CU (a single workgroup per CU left
    // each one a long chain of dependent tile loads: V^T C at 6300 x 3456
    // read C at 1.3 TB/s), and no chunk shorter than 128 (eight k-steps).
    int S = 1;
    long kchunk = 0;
    if (batch == 1 && tm * tn < 4L * ctx().num_cu && K >= 256 && !ctx().no_alloc) {
        const long wgs = 4L * ctx().num_cu;
        long want = std::min<long>((wgs + tm * tn - 1) / (tm * tn), K / 128);
        long cap = (64L << 20) / ((long)sizeof(T) * M * N);
        S = (int)std::max<long>(1, std::min<long>({want, cap, 64}));
        if (S > 1) {
            kchunk = ((K + S - 1) / S + GBK - 1) / GBK * GBK;
            S = (int)((K + kchunk - 1) / kchunk);
        }
    }
    T *Cw = C;
    long ldw = ldc, sW = sC, nz = batch;
    T al = alpha, be = beta;
    if (S > 1) {
        Cw = static_cast<T *>(ctx().workspace(sizeof(T) * (size_t)S * M * N));
        ldw = M;
        sW = M * N;
        nz = S;
        al = T(1);
        be = T(0);
    }
    dim3 grid((unsigned)tm, (unsigned)tn, (unsigned)nz);
    const int kc = (int)kchunk;
    // large symmetric products (the Gram matrix, the Cholesky trailing
    // updates) take the XCD-aware tile order (1-D grid over super-tiles):
    // 2.18 -> 1.80 ms for the 4400 x 4000 FP32 Gram; the general products of
    // the FP64 QR measured no better with it and keep the 2-D grid
    int stm = 0, stn = 0;
    if (upper && tm >= GS && tm == tn) {
        const long sm = (tm + GS - 1) / GS, sn = (tn + GS - 1) / GS;
        const long nsup = upper ? sn * (sn + 1) / 2 : sm * sn;
        grid = dim3((unsigned)(nsup * GS * GS), 1, (unsigned)nz);
        stm = (int)tm;
        stn = (int)tn;
    }
    if (!TA && !TB)
        k_gemm<T, false, false><<<grid, 256, 0, s>>>((int)M, (int)N, (int)K, al, A, lda, B, ldb,
                                                     be, Cw, ldw, sA, sB, sW, kc, upper ? 1 : 0, stm, stn);
    else if (TA && !TB)
        k_gemm<T, true, false><<<grid, 256, 0, s>>>((int)M, (int)N, (int)K, al, A, lda, B, ldb,
                                                    be, Cw, ldw, sA, sB, sW, kc, upper ? 1 : 0, stm, stn);
    else if (!TA && TB)
        k_gemm<T, false, true><<<grid, 256, 0, s>>>((int)M, (int)N, (int)K, al, A, lda, B, ldb,
                                                    be, Cw, ldw, sA, sB, sW, kc, upper ? 1 : 0, stm, stn);
    else
        k_gemm<T, true, true><<<grid, 256, 0, s>>>((int)M, (int)N, (int)K, al, A, lda, B, ldb,
                                                   be, Cw, ldw, sA, sB, sW, kc, upper ? 1 : 0, stm, stn);
    SKL_CHECK_LAUNCH();
    if (S > 1) {
        long tot = M * N;
        k_splitk_reduce<T><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>((int)M, (int)N, S, Cw,
                                                                          alpha, beta, C, ldc);
        SKL_CHECK_LAUNCH();
    }
}

template void gemm<double>(char, char, long, long, long, double, const double *, long,
                           const double *, long, double, double *, long, long, long, long, long,
                           bool);
template void gemm<float>(char, char, long, long, long, float, const float *, long,
                          const float *, long, float, float *, long, long, long, long, long, bool);

// ---------------------------------------------------------------------------
// Small helpers
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_fill2d(long m, long n, T *A, long lda, T v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long j = blockIdx.y;
    for (; j < n; j += gridDim.y)
        if (i < m) A[i + j * lda] = v;
}

template <class Ts, class Td>
__global__ void k_copy2d(long m, long n, const Ts *A, long lda, Td *B, long ldb) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (long j = blockIdx.y; j < n; j += gridDim.y)
        if (i < m) B[i + j * ldb] = (Td)A[i + j * lda];
}

template <class T> void fill2d(long m, long n, T *A, long lda, T v) {
    if (m <= 0 || n <= 0) return;
    dim3 grid((unsigned)((m + 255) / 256), (unsigned)std::min<long>(n, 65535));
    k_fill2d<T><<<grid, 256, 0, ctx().stream>>>(m, n, A, lda, v);
    SKL_CHECK_LAUNCH();
}
template void fill2d<double>(long, long, double *, long, double);
template void fill2d<float>(long, long, float *, long, float);
template void fill2d<int32_t>(long, long, int32_t *, long, int32_t);

template <class Ts, class Td> void copy2d(long m, long n, const Ts *A, long lda, Td *B, long ldb) {
    if (m <= 0 || n <= 0) return;
    dim3 grid((unsigned)((m + 255) / 256), (unsigned)std::min<long>(n, 65535));
    k_copy2d<Ts, Td><<<grid, 256, 0, ctx().stream>>>(m, n, A, lda, B, ldb);
    SKL_CHECK_LAUNCH();
}
template void copy2d<double, double>(long, long, const double *, long, double *, long);
template void copy2d<double, float>(long, long, const double *, long, float *, long);
template void copy2d<float, double>(long, long, const float *, long, double *, long);
template void copy2d<float, float>(long, long, const float *, long, float *, long);

// upper triangle of R (T) into Rd (double), zeros below the diagonal
template <class T>
__global__ void k_upcast_upper(long r, const T *R, long ldr, double *Rd, long ldd) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (long j = blockIdx.y; j < r; j += gridDim.y)
        if (i < r) Rd[i + j * ldd] = (i <= j) ? (double)R[i + j * ldr] : 0.0;
}
template <class T> void upcast_upper(long r, const T *R, long ldr, double *Rd, long ldd) {
    if (r <= 0) return;
    dim3 grid((unsigned)((r + 255) / 256), (unsigned)std::min<long>(r, 65535));
    k_upcast_upper<T><<<grid, 256, 0, ctx().stream>>>(r, R, ldr, Rd, ldd);
    SKL_CHECK_LAUNCH();
}
template void upcast_upper<double>(long, const double *, long, double *, long);
template void upcast_upper<float>(long, const float *, long, double *, long);

// ---------------------------------------------------------------------------
// Deterministic reductions: per-block partials then one block for the total.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sumsq_partial(long n, const double *x, double *part) {
    double s = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        s += x[i] * x[i];
    s = block_sum<256>(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_sum_final(int np, const double *part, double *out,
                                                   int take_sqrt) {
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) s += part[i];
    s = block_sum<256>(s);
    if (threadIdx.x == 0) *out = take_sqrt ? sqrt(s) : s;
}

double dev_nrm2(long n, const double *x, Arena &ar) {
    int np = (int)std::min<long>(1024, std::max<long>(1, (n + 255) / 256));
    double *part = ar.alloc_n<double>((size_t)np + 1);
    hipStream_t s = ctx().stream;
    k_sumsq_partial<<<np, 256, 0, s>>>(n, x, part);
    k_sum_final<<<1, 256, 0, s>>>(np, part, part + np, 1);
    SKL_CHECK_LAUNCH();
    double h = 0;
    SKL_HIP(hipMemcpyAsync(&h, part + np, sizeof(double), hipMemcpyDeviceToHost, s));
    SKL_HIP(hipStreamSynchronize(s));
    return h;
}

// inf-norm of diag(rs) R diag(cs) for upper triangular R (rs, cs may be
// null): partial row sums over 256-column chunks, then per-row totals and the
// maximum, so every CU takes part.
constexpr int RA_CW = 256;
__global__ __launch_bounds__(256) void k_rowabs_part(long r, const double *R, long ldr,
                                                     const double *cs, double *part) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    const long c0 = (long)blockIdx.y * RA_CW, c1 = min(r, c0 + RA_CW);
    if (i >= r) return;
    double s = 0.0;
    for (long j = max(i, c0); j < c1; ++j) s += fabs(R[i + j * ldr]) * (cs ? cs[j] : 1.0);
    part[(long)blockIdx.y * r + i] = s;
}
__global__ __launch_bounds__(256) void k_rowabs_max(long r, int nch, const double *part,
                                                    const double *rs, double *bmax) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (i < r) {
        for (int c = 0; c < nch; ++c) s += part[(long)c * r + i];
        if (rs) s *= rs[i];
    }
    s = block_max<256>(s);
    if (threadIdx.x == 0) bmax[blockIdx.x] = s;
}
__global__ __launch_bounds__(256) void k_max_final(long n, const double *x, double *out) {
    double s = 0.0;
    for (long i = threadIdx.x; i < n; i += 256) s = fmax(s, x[i]);
    s = block_max<256>(s);
    if (threadIdx.x == 0) *out = s;
}
double norm_inf_upper(long r, const double *R, long ldr, Arena &ar, const double *rs,
                      const double *cs) {
    if (r <= 0) return 0.0;
    const int nch = (int)((r + RA_CW - 1) / RA_CW);
    const unsigned nb = (unsigned)((r + 255) / 256);
    const size_t mark = ar.used();
    double *part = ar.alloc_n<double>((size_t)nch * r);
    double *bm = ar.alloc_n<double>((size_t)nb + 1);
    hipStream_t s = ctx().stream;
    k_rowabs_part<<<dim3(nb, (unsigned)nch), 256, 0, s>>>(r, R, ldr, cs, part);
    k_rowabs_max<<<nb, 256, 0, s>>>(r, nch, part, rs, bm);
    k_max_final<<<1, 256, 0, s>>>(nb, bm, bm + nb);
    SKL_CHECK_LAUNCH();
    double h = 0;
    SKL_HIP(hipMemcpyAsync(&h, bm + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    SKL_HIP(hipStreamSynchronize(s));
    ar.rewind(mark);
    return h;
}

// D[j] = ||R(0:j, j)||_2 (one wave per column); invD[j] = 1 / D[j] (0 for a zero column)
__global__ __launch_bounds__(256) void k_colnorm_upper(long r, const double *R, long ldr, double *D,
                                                       double *invD) {
    const long j = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= r) return;
    double s = 0.0;
    for (long i = lane; i <= j; i += 64) {
        double v = R[i + j * ldr];
        s += v * v;
    }
    s = wave_sum(s);
    if (lane == 0) {
        D[j] = sqrt(s);
        invD[j] = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
    }
}
void colnorm_upper(long r, const double *R, long ldr, double *D, double *invD) {
    if (r <= 0) return;
    k_colnorm_upper<<<(unsigned)((r + 3) / 4), 256, 0, ctx().stream>>>(r, R, ldr, D, invD);
    SKL_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// Upper triangular inverse M = (R + p I)^{-1}.  Diagonal 32x32 blocks are
// inverted in LDS, one workgroup each, walking the superdiagonals; the
// off-diagonal blocks follow by doubling: for [M11 M12; 0 M22],
// M12 = -M11 R12 M22, two batched GEMMs per level.
// ---------------------------------------------------------------------------
constexpr int TIB = 32;

__global__ __launch_bounds__(1024) void k_trinv_diag(long r, const double *R, long ldr,
                                                     double p, double *M, long ldm) {
    __shared__ double Rs[TIB][TIB + 1];
    __shared__ double Xs[TIB][TIB + 1];
    long b0 = (long)blockIdx.x * TIB;
    int nb = (int)min((long)TIB, r - b0);
    int i = threadIdx.x % TIB, j = threadIdx.x / TIB;
    Rs[i][j] = (i < nb && j < nb && i <= j) ? R[(b0 + i) + (b0 + j) * ldr] : 0.0;
    if (i == j && i < nb) Rs[i][j] += p;
    Xs[i][j] = 0.0;
    __syncthreads();
    if (i == j && i < nb) Xs[i][i] = 1.0 / Rs[i][i];
    __syncthreads();
    for (int dist = 1; dist < nb; ++dist) {
        if (j - i == dist && j < nb) {
            double s = 0.0;
            for (int l = i + 1; l <= j; ++l) s += Rs[i][l] * Xs[l][j];
            Xs[i][j] = -s / Rs[i][i];
        }
        __syncthreads();
    }
    if (i < nb && j < nb) M[(b0 + i) + (b0 + j) * ldm] = Xs[i][j];
}

void trinv_upper(long r, const double *R, long ldr, double perturb, double *M, long ldm,
                 Arena &ar) {
    if (r <= 0) return;
    fill2d<double>(r, r, M, ldm, 0.0);
    long nblk = (r + TIB - 1) / TIB;
    k_trinv_diag<<<(unsigned)nblk, 1024, 0, ctx().stream>>>(r, R, ldr, perturb, M, ldm);
    SKL_CHECK_LAUNCH();
    // doubling over super-blocks [b0, b0+2s): M12 (rows b0.., cols b0+s..)
    double *tmp = ar.alloc_n<double>((size_t)r * std::min<long>(r, 2 * TIB) + 64);
    size_t tmp_cap = (size_t)r * std::min<long>(r, 2 * TIB) + 64;
    for (long s = TIB; s < r; s *= 2) {
        long nfull = r / (2 * s);         // super-blocks with a full M22
        long rem = r - nfull * 2 * s;     // tail: [b0, r) with b0 = nfull*2s
        size_t need = (size_t)s * s * (size_t)std::max<long>(nfull, 1);
        if (need > tmp_cap) {
            tmp = ar.alloc_n<double>(need);
            tmp_cap = need;
        }
        if (nfull > 0) {
            // tmp = R12 * M22 ; M12 = -M11 * tmp
            long str = 2 * s * (ldr + 1), stm = 2 * s * (ldm + 1);
            gemm<double>('N', 'N', s, s, s, 1.0, R + s * ldr, ldr, M + s + s * ldm, ldm, 0.0, tmp,
                         s, nfull, str, stm, s * s);
            gemm<double>('N', 'N', s, s, s, -1.0, M, ldm, tmp, s, 0.0, M + s * ldm, ldm, nfull,
                         stm, s * s, stm);
        }
        if (rem > s) {
            long b0 = nfull * 2 * s, s2 = rem - s;
            const double *R12 = R + b0 + (b0 + s) * ldr;
            const double *M22 = M + (b0 + s) + (b0 + s) * ldm;
            const double *M11 = M + b0 + b0 * ldm;
            double *M12 = M + b0 + (b0 + s) * ldm;
            gemm<double>('N', 'N', s, s2, s2, 1.0, R12, ldr, M22, ldm, 0.0, tmp, s);
            gemm<double>('N', 'N', s, s2, s, -1.0, M11, ldm, tmp, s, 0.0, M12, ldm);
        }
    }
}

// ---------------------------------------------------------------------------
// Cholesky factorization G = R^T R (upper R) of a symmetric positive definite
// d x d matrix, blocked right-looking with 64-column blocks: the diagonal
// block in one workgroup, the block row by forward substitution
// (one wave per four columns), the trailing matrix by a GEMM on the matrix
// cores (upper tiles only).
// ---------------------------------------------------------------------------
constexpr int CH_B = 64;

constexpr int CH_NT = 256;
// Cholesky of the b x b diagonal block (upper triangle of A, b <= 64) in
// place: R upper, zeros below; *fail = 1 when a pivot is not positive and
// finite (the step then goes on with a unit pivot).  Blocked by 16 steps with
// no barrier inside a block: wave 0 runs the steps, lane p for column p,
// reading column k of the Schur complement as it was at the block's start
// (LDS, the full symmetric matrix) minus the block's earlier rows of R (the
// pivot column's entries by v_readlane), with 1/sqrt of the pivot from the
// hardware estimate and two Newton steps -- a few dependent operations per
// step; between blocks the four waves update the Schur complement by the
// block's 16 rows on the matrix cores.
constexpr int CH_LD = CH_B + 1;
static_assert(CH_LD == CQ_LD, "the triangular inverse reads the block with ld CQ_LD");
__global__ __launch_bounds__(CH_NT) void k_chol_diag(int b, double *A, long lda, int *fail,
                                                   double *Uinv) {
    static_assert(CH_NT == 256 && CH_B == 64, "four waves over a 64 x 64 block");
    __shared__ double Ss[CH_B * CH_LD]; // Schur complement at the block's start
    __shared__ double Rb[CH_B * CH_LD]; // rows of R by column
    __shared__ double Ys[CH_B * CH_LD]; // scratch of the inverse
    const int tid = threadIdx.x, p = tid & 63, g = tid >> 6;
    for (int e = tid; e < CH_B * CH_B; e += CH_NT) {
        const int i = e & 63, c = e >> 6;
        const double v = (i < b && c < b) ? (i <= c ? A[i + (long)c * lda] : A[c + (long)i * lda])
                                          : (i == c ? 1.0 : 0.0);
        Ss[i + c * CH_LD] = v;
        Rb[i + c * CH_LD] = 0.0;
    }
    __syncthreads();
    bool bad = false;
    for (int k0 = 0; k0 < b; k0 += 16) {
        if (g == 0) {
            double rl[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) rl[t] = 0.0;
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int k = k0 + t;
                if (k < b) {
                    double sc = Ss[p + k * CH_LD];
#pragma unroll
                    for (int u = 0; u < t; ++u) sc = fma(-readlane(rl[u], k), rl[u], sc);
                    double d = readlane(sc, k);
                    const bool pb = !(d > 0.0) || !(d < INFINITY);
                    bad = bad || pb;
                    if (pb) d = 1.0;
                    const double inv = rsqrt_nr(d);
                    rl[t] = p == k ? d * inv : ((p > k && p < b) ? sc * inv : 0.0);
                }
            }
#pragma unroll
            for (int t = 0; t < 16; ++t)
                if (k0 + t < b) Rb[(k0 + t) + p * CH_LD] = rl[t];
        }
        __syncthreads();
        if (k0 + 16 >= b) break;
        // S -= R(k0:k0+16, :)^T R(k0:k0+16, :): 16 x 16 tiles, four per wave,
        // each read and written by its own wave only
        {
            typedef double acc4 __attribute__((ext_vector_type(4)));
            const int li = p & 15, lk = p >> 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int tile = g * 4 + q, ti = tile & 3, tj = tile >> 2;
                acc4 c;
#pragma unroll
                for (int r = 0; r < 4; ++r) c[r] = Ss[(16 * ti + lk + 4 * r) + (16 * tj + li) * CH_LD];
#pragma unroll
                for (int u0 = 0; u0 < 16; u0 += 4)
                    c = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        -Rb[(k0 + u0 + lk) + (16 * ti + li) * CH_LD],
                        Rb[(k0 + u0 + lk) + (16 * tj + li) * CH_LD], c, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) Ss[(16 * ti + lk + 4 * r) + (16 * tj + li) * CH_LD] = c[r];
            }
        }
        __syncthreads();
    }
    if (g == 0 && __ballot(bad) != 0 && p == 0) *fail = 1;
    for (int e = tid; e < CH_B * CH_B; e += CH_NT) {
        const int i = e & 63, c = e >> 6;
        if (i < b && c < b) A[i + (long)c * lda] = i <= c ? Rb[i + c * CH_LD] : 0.0;
    }
    if (!Uinv) return;
    // U^-1 (64 x 64, ld 64; identity past b) for the block row's product
    if (tid < CH_B && tid >= b) Rb[tid + tid * CH_LD] = 1.0;
    __syncthreads();
    tri_inv_upper_lds(Rb, Ss, Ys);
    for (int e = tid; e < CH_B * CH_B; e += CH_NT) Uinv[e] = Ss[(e & 63) + (e >> 6) * CH_LD];
}

constexpr int CH_TC = 4; // columns of B per wave in the block-row product

// X = U^{-T} B in place for the b x b upper factor U (b <= 64) and B (b x nc)
// from Uinv = U^{-1} (k_chol_diag; upper, ld 64): x_p = sum_{i <= p} Uinv(i, p) b_i, lane p = row p with column p of
// Uinv in registers, b staged through LDS into the row-broadcast layout so
// that every term is one fused DPP FMA (bcast16_fma) -- independent terms
// instead of a forward substitution's 64-step chain.
__global__ __launch_bounds__(64) void k_chol_xmul(int b, const double *Uinv, double *B, long ldb,
                                                  long nc) {
    __shared__ double sb[CH_TC][CH_B];
    const int p = threadIdx.x;
    double u[CH_B];
#pragma unroll
    for (int i = 0; i < CH_B; ++i) u[i] = (i < b && p < b && i <= p) ? Uinv[i + CH_B * p] : 0.0;
    const long c0 = (long)blockIdx.x * CH_TC;
#pragma unroll
    for (int t = 0; t < CH_TC; ++t) sb[t][p] = (p < b && c0 + t < nc) ? B[p + (c0 + t) * ldb] : 0.0;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < CH_TC; ++t) {
        double vb[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) vb[q] = sb[t][16 * q + (p & 15)];
        dpp_src_fence(vb);
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        static_for<0, CH_B>([&](auto rc) {
            constexpr int R = decltype(rc)::value;
            bcast16_fma<R>(acc[R & 3], vb[R >> 4], u[R]);
        });
        if (p < b && c0 + t < nc) B[p + (c0 + t) * ldb] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    }
}

// R (upper, zeros below, ld ldr) from G (overwritten); false when a pivot
// was not positive (G not numerically positive definite)
bool cholesky_upper(long d, double *G, long ldg, double *R, long ldr, Arena &ar) {
    hipStream_t s = ctx().stream;
    const size_t mark = ar.used();
    int *fail = ar.alloc_n<int>(1);
    SKL_HIP(hipMemsetAsync(fail, 0, sizeof(int), s));
    // the diagonal blocks' inverses, one slot per block
    double *uinv = ar.alloc_n<double>((size_t)((d + CH_B - 1) / CH_B) * CH_B * CH_B);
    // Look-ahead: after block j's row B_j is solved, the main stream updates
    // only the next block row (what block j + 1 factors and solves next); the
    // rest of the trailing update runs on the context's second stream beside
    // block j + 1's diagonal factor and solve, and the main stream waits for
    // it only before it updates the following block row (which that update
    // also touches).
    Context &c = ctx();
    c.make_side();
    bool pending = false; // a trailing update on the side stream not yet waited for
    for (long j = 0; j < d; j += CH_B) {
        const int b = (int)std::min<long>(CH_B, d - j);
        const long nc = d - j - b;
        double *Gjj = G + j + j * ldg;
        double *Uinv = nc > 0 ? uinv + (j / CH_B) * CH_B * CH_B : nullptr;
        k_chol_diag<<<1, CH_NT, 0, s>>>(b, Gjj, ldg, fail, Uinv);
        if (nc > 0) {
            double *Bp = G + j + (j + b) * ldg;
            k_chol_xmul<<<(unsigned)((nc + CH_TC - 1) / CH_TC), 64, 0, s>>>(b, Uinv, Bp, ldg, nc);
            if (pending) SKL_HIP(hipStreamWaitEvent(s, c.ev[1], 0));
            pending = false;
            // next block row: rows j + b .. j + b + b2 of the trailing matrix
            const long b2 = std::min<long>(CH_B, nc), nc2 = nc - b2;
            gemm<double>('T', 'N', b2, nc, b, -1.0, Bp, ldg, Bp, ldg, 1.0,
                         G + (j + b) + (j + b) * ldg, ldg, 1, 0, 0, 0, true);
            if (nc2 > 0) {
                SKL_HIP(hipEventRecord(c.ev[0], s));
                OnSideStream on_side;
                hipStream_t ss = ctx().stream;
                SKL_HIP(hipStreamWaitEvent(ss, c.ev[0], 0));
                const double *B2 = Bp + b2 * ldg;
                gemm<double>('T', 'N', nc2, nc2, b, -1.0, B2, ldg, B2, ldg, 1.0,
                             G + (j + b + b2) + (j + b + b2) * ldg, ldg, 1, 0, 0, 0, true);
                SKL_HIP(hipEventRecord(c.ev[1], ss));
                pending = true;
            }
        }
        SKL_CHECK_LAUNCH();
    }
    if (pending) SKL_HIP(hipStreamWaitEvent(s, c.ev[1], 0));
    upcast_upper<double>(d, G, ldg, R, ldr);
    int hf = 0;
    SKL_HIP(hipMemcpyAsync(&hf, fail, sizeof(int), hipMemcpyDeviceToHost, s));
    SKL_HIP(hipStreamSynchronize(s));
    ar.rewind(mark);
    return hf == 0;
}

} // namespace skl
