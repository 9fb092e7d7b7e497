// solver.hip -- orchestration of the three Ski-LLS solvers on one MI355X and
// the C ABI of include/skl.h.
//
// Dense (src/ls_dense_hashing_blendenpik.cpp:23-149):
//   signs + hashing pattern (host, libc-compatible draws) -> fused DHT +
//   bucket-sum sketch of [A | b] -> CPQR of [SA | c] (c receives Q^T c) ->
//   rank -> M = R^{-1} -> x0 = P M c -> r = b - A x0 -> LSQR on A(:,E) M ->
//   x = x0 + P M y.  A is never permuted in memory: E is folded into the
//   LSQR kernels' indexing.
// Dense without CPQR (src/ls_dense_hashing_blendenpik_noCPQR.cpp:23-60).
// Sparse (src/ls_sparse_spqr.cpp:79-214): hashing sketch of the CSC matrix,
//   Householder QR of [SA | S b] standing in for SuiteSparseQR, rank check
//   against SPQR's default tolerance (column-pivoted refactorisation when it
//   fails), exact inf-norm condition number of R11, LSQR with the pivot
//   perturbation when R11 is ill conditioned.
// Mixed precision (SKL_MIXED): the factorization above is replaced by a
//   CholeskyQR of the column-equilibrated sketch (mixed_cholqr), with a
//   fall-back to it when that factor loses rank or conditioning.
#include <algorithm>
#include <functional>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <map>
#include <memory>
#include <mutex>

#include "k_common.h"

namespace skl {

template <class T>
void dht_cols_from_double(long n, long ncols, const double *X, long ldx, T *Y, long ldy, Arena &ar);
template <class T>
void sparse_sketch_t(long m, long d, long k, const int32_t *srows, const double *svals,
                     const int32_t *colptr, const int32_t *rowind, const double *val, long nnz,
                     T *SA, long ldsa, Arena &ar);
void narrow_indices(long n, const long *in, int32_t *out);
void csc_to_csr(long n, long d, long nnz, const int32_t *colptr, const int32_t *rowind,
                const double *val, int32_t *rptr, int32_t *ccol, double *cval, Arena &ar);
void csr_gemv(long n, const int32_t *ptr, const int32_t *col, const double *val, const double *x,
              double alpha, double beta, double *y);
void set_device(int dev);

static thread_local std::string t_err;
void set_last_error(const std::string &msg) { t_err = msg; }
static thread_local skl_lsqr_step *t_hist = nullptr;
static thread_local long t_hist_len = 0;

// ---------------------------------------------------------------------------
// glue kernels
// ---------------------------------------------------------------------------
__global__ void k_scatter_E(long d, long r, const int32_t *E, const double *v, double *x) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < r) x[E ? E[i] : i] = v[i];
}
__global__ void k_fill_d(long n, double *x, double v) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}
__global__ void k_axpy_d(long n, const double *a, double *y) { // y = a + y
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a[i] + y[i];
}
__global__ void k_iota(long n, int32_t *x) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = (int32_t)i;
}
template <class T> __global__ void k_diag(long r, const T *W, long ldw, double *dg) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < r) dg[i] = (double)W[i + i * ldw];
}
template <class T> __global__ void k_col_to_double(long r, const T *c, double *o) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < r) o[i] = (double)c[i];
}
__global__ void k_add_diag(long r, double *R, long ldr, double p) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < r) R[i + i * ldr] += p;
}
template <class T>
__global__ __launch_bounds__(256) void k_colnorm2(long m, long d, const T *W, long ldw, double *out) {
    long j = blockIdx.x;
    double s = 0.0;
    for (long i = threadIdx.x; i < m; i += 256) {
        double v = (double)W[i + j * ldw];
        s += v * v;
    }
    s = block_sum<256>(s);
    if (threadIdx.x == 0) out[j] = s;
}

static inline unsigned gblocks(long n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

// M v (upper r x r) into out (double): out = M v
static void trmv_n(long r, const double *M, long ldm, const double *v, double *out, Arena &ar) {
    dev_gemv_n_dense(r, r, M, ldm, v, 1.0, 0.0, out, ar);
}

struct DevPattern {
    double *D = nullptr;
    int32_t *bptr = nullptr, *bsrc = nullptr, *rows = nullptr;
    double *coef_d = nullptr; // alpha * val in bucket order (double)
    float *coef_f = nullptr;
    double *vals = nullptr;   // n*k values (source order)
};

static DevPattern upload_pattern(const Pattern &P, double alpha, bool want_D, bool want_rows,
                                 bool want_float, Arena &ar) {
    hipStream_t s = ctx().stream;
    DevPattern dp;
    const long nk = (long)P.bsrc.size(); // bucket entries: n k, or m for row sampling
    if (want_D) {
        dp.D = ar.alloc_n<double>((size_t)P.n);
        SKL_HIP(hipMemcpyAsync(dp.D, P.D.data(), sizeof(double) * P.n, hipMemcpyHostToDevice, s));
    }
    dp.bptr = ar.alloc_n<int32_t>((size_t)P.m + 1);
    dp.bsrc = ar.alloc_n<int32_t>((size_t)std::max<long>(nk, 1));
    SKL_HIP(hipMemcpyAsync(dp.bptr, P.bptr.data(), sizeof(int32_t) * (P.m + 1), hipMemcpyHostToDevice, s));
    SKL_HIP(hipMemcpyAsync(dp.bsrc, P.bsrc.data(), sizeof(int32_t) * nk, hipMemcpyHostToDevice, s));
    // alpha * vals[idx] as the reference multiplies (SpMat.cpp:55)
    std::vector<double> cd((size_t)nk);
    for (long q = 0; q < nk; ++q) cd[(size_t)q] = alpha * P.vals[(size_t)P.bslot[(size_t)q]];
    dp.coef_d = ar.alloc_n<double>((size_t)std::max<long>(nk, 1));
    SKL_HIP(hipMemcpyAsync(dp.coef_d, cd.data(), sizeof(double) * nk, hipMemcpyHostToDevice, s));
    if (want_float) {
        std::vector<float> cf((size_t)nk);
        for (long q = 0; q < nk; ++q) cf[(size_t)q] = (float)cd[(size_t)q];
        dp.coef_f = ar.alloc_n<float>((size_t)std::max<long>(nk, 1));
        SKL_HIP(hipMemcpyAsync(dp.coef_f, cf.data(), sizeof(float) * nk, hipMemcpyHostToDevice, s));
        SKL_HIP(hipStreamSynchronize(s)); // cf goes out of scope
    }
    if (want_rows) {
        const long nr = (long)P.rows.size();
        dp.rows = ar.alloc_n<int32_t>((size_t)std::max<long>(nr, 1));
        SKL_HIP(hipMemcpyAsync(dp.rows, P.rows.data(), sizeof(int32_t) * nr, hipMemcpyHostToDevice, s));
        dp.vals = ar.alloc_n<double>((size_t)std::max<long>(nr, 1));
        SKL_HIP(hipMemcpyAsync(dp.vals, P.vals.data(), sizeof(double) * nr, hipMemcpyHostToDevice, s));
    }
    SKL_HIP(hipStreamSynchronize(s));
    return dp;
}

static const std::vector<double> &gaussian_for(long m) {
    static std::mutex mu;
    static std::map<long, std::vector<double>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(m);
    if (it != cache.end()) return it->second;
    std::vector<double> G;
    gen_gaussian(12, 74, m, G); // srand(12) in HQRRP (skint.c:390)
    return cache.emplace(m, std::move(G)).first->second;
}

static void sync() { SKL_HIP(hipStreamSynchronize(ctx().stream)); }

// At the end of a solve the arena folds any overflow blocks into one
// reservation of the solve's high-water size, so that a repeated solve of the
// same shape allocates nothing -- including the first repeat.
struct ArenaTrim {
    Arena &ar;
    explicit ArenaTrim(Arena &a) : ar(a) {}
    ~ArenaTrim() {
        (void)hipStreamSynchronize(ctx().stream);
        // an early exit can leave work on the other streams still running
        // (the row-major copy of A, a look-ahead update)
        if (ctx().aux) (void)hipStreamSynchronize(ctx().aux);
        if (ctx().side) (void)hipStreamSynchronize(ctx().side);
        try {
            ar.reset();
        } catch (...) {
            // a failed reallocation leaves the arena empty; the next solve
            // reserves again
        }
    }
};

// ---------------------------------------------------------------------------
// Dense solve (both variants), device pointers
// ---------------------------------------------------------------------------
// [SA | c] = S H D [A, b] into W (c present when db is)
template <class T>
static void dense_sketch(long n, long d, long m, const double *dA, long lda, const double *db,
                         const DevPattern &dp, T *W, Arena &ar) {
    const T *coef = std::is_same<T, float>::value ? (const T *)dp.coef_f : (const T *)dp.coef_d;
    double scale = 1 / std::sqrt((double)n);
    dht_sketch<T>(n, d, dA, lda, dp.D, scale, m, dp.bptr, dp.bsrc, coef, W, m, ar);
    if (db) dht_sketch<T>(n, 1, db, n, dp.D, scale, m, dp.bptr, dp.bsrc, coef, W + d * m, m, ar);
}

// QR of the first d columns of W (m x ncol), the rest receiving Q^T:
// column-pivoted when pivot is set (jpvt receives the order), plain
// Householder otherwise (jpvt = identity)
template <class T>
static void dense_factor(long m, long ncol, long d, bool pivot, T *W, int32_t *jpvt, T *tau,
                         Arena &ar) {
    hipLaunchKernelGGL(k_iota, dim3(gblocks(d)), dim3(256), 0, ctx().stream, d, jpvt);
    if (pivot) {
        if (std::min(m, d) < 1000)
            cpqr_greedy<T>(m, ncol, d, W, m, jpvt, tau, ar);
        else
            cpqr_hqrrp<T>(m, ncol, d, W, m, jpvt, tau, gaussian_for(m).data(), ar);
    } else {
        qr_householder<T>(m, ncol, d, W, m, tau, ar);
    }
}

template <class T>
static void dense_sketch_qr(long n, long d, long m, const double *dA, long lda, const double *db,
                            const DevPattern &dp, bool pivot, T *W, int32_t *jpvt, T *tau,
                            double *t_sketch, double *t_qr, Arena &ar,
                            const std::function<void()> &after_sketch = {}) {
    double t0 = wall();
    dense_sketch<T>(n, d, m, dA, lda, db, dp, W, ar);
    sync();
    double t1 = wall();
    if (after_sketch) after_sketch();
    dense_factor<T>(m, db ? d + 1 : d, d, pivot, W, jpvt, tau, ar);
    sync();
    double t2 = wall();
    *t_sketch += t1 - t0;
    *t_qr += t2 - t1;
}

// reciprocal inf-norm condition number from R (double, upper) and M = R^{-1}
static double rcond_from(long r, const double *Rd, long ldr, const double *M, long ldm, Arena &ar) {
    double a = norm_inf_upper(r, Rd, ldr, ar);
    double b = norm_inf_upper(r, M, ldm, ar);
    if (!(a > 0) || !(b > 0) || !std::isfinite(b)) return 0.0;
    return (1.0 / a) / b;
}

// inf-norm condition number of the column-equilibrated factor R D^{-1},
// D = diag(||R(:, j)||_2): ||R D^{-1}||_inf ||D M||_inf.  Householder QR in
// FP32 is columnwise backward stable (the computed R is the exact factor of
// SA + E with ||E(:, j)|| <= c u_32 ||SA(:, j)||), so the quality of an FP32
// preconditioner is governed by this number, not by the raw condition number
// that column scaling alone can inflate.
static double scaled_cond(long r, const double *Rd, long ldr, const double *M, long ldm,
                          Arena &ar) {
    if (r <= 0) return INFINITY;
    const size_t mark = ar.used();
    double *D = ar.alloc_n<double>((size_t)r);
    double *invD = ar.alloc_n<double>((size_t)r);
    colnorm_upper(r, Rd, ldr, D, invD);
    double a = norm_inf_upper(r, Rd, ldr, ar, nullptr, invD);
    double b = norm_inf_upper(r, M, ldm, ar, D, nullptr);
    ar.rewind(mark);
    if (!(a > 0) || !(b > 0) || !std::isfinite(a * b)) return INFINITY;
    return a * b;
}

// ---- mixed precision: CholeskyQR of the column-equilibrated sketch ----------
__global__ void k_norms_from_squares(long d, const double *sq, double *D, double *invD) {
    long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= d) return;
    const double v = sqrt(sq[j]);
    D[j] = v;
    invD[j] = v > 0.0 ? 1.0 / v : 0.0;
}
template <class T>
__global__ void k_equil_to_float(long m, long d, const T *SA, long ldsa, const double *invD,
                                 float *W) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    for (long j = blockIdx.y; j < d; j += gridDim.y)
        W[i + j * m] = (float)((double)SA[i + j * ldsa] * invD[j]);
}
template <class T>
__global__ void k_equil_to_double(long m, long d, const T *SA, long ldsa, const double *invD,
                                  double *W) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    for (long j = blockIdx.y; j < d; j += gridDim.y) W[i + j * m] = (double)SA[i + j * ldsa] * invD[j];
}
// R = R1 diag(D), M = diag(1/D) M1 (d x d, ld d)
__global__ void k_unequilibrate(long d, const double *R1, const double *M1, const double *D,
                                const double *invD, double *R, double *M) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d) return;
    for (long j = blockIdx.y; j < d; j += gridDim.y) {
        R[i + j * d] = R1[i + j * d] * D[j];
        M[i + j * d] = invD[i] * M1[i + j * d];
    }
}

// One Lanczos step on H = (SA M)^T (SA M), given h = H v_j:
// alpha_j = v_j^T h, h -= alpha_j v_j + beta_j v_{j-1}, beta_{j+1} = ||h||,
// v_{j+1} = h / beta_{j+1}.  One workgroup; ab[2j] = alpha_j,
// ab[2j+1] = beta_{j+1} (ab[-1], the beta_0 of the first step, is not read).
__global__ __launch_bounds__(1024) void k_lanczos_step(long d, int j, const double *v,
                                                        const double *vp, double *h, double *vn,
                                                        double *ab) {
    double a = 0.0;
    for (long i = threadIdx.x; i < d; i += 1024) a += v[i] * h[i];
    a = block_sum<1024>(a);
    const double b = j > 0 ? ab[2 * j - 1] : 0.0;
    double q = 0.0;
    for (long i = threadIdx.x; i < d; i += 1024) {
        const double t = h[i] - a * v[i] - (j > 0 ? b * vp[i] : 0.0);
        h[i] = t;
        q += t * t;
    }
    q = sqrt(block_sum<1024>(q));
    const double inv = q > 0.0 ? 1.0 / q : 0.0;
    for (long i = threadIdx.x; i < d; i += 1024) vn[i] = h[i] * inv;
    if (threadIdx.x == 0) {
        ab[2 * j] = a;
        ab[2 * j + 1] = q;
    }
}

// Extreme eigenvalues of the symmetric tridiagonal with diagonal a and
// off-diagonal b (b[i] couples i - 1 and i), by bisection on the Sturm count.
static void tridiag_extremes(const std::vector<double> &a, const std::vector<double> &b,
                             double &lmin, double &lmax) {
    const size_t k = a.size();
    double lo = INFINITY, hi = -INFINITY;
    for (size_t i = 0; i < k; ++i) {
        const double r = (i > 0 ? std::fabs(b[i]) : 0.0) + (i + 1 < k ? std::fabs(b[i + 1]) : 0.0);
        lo = std::min(lo, a[i] - r);
        hi = std::max(hi, a[i] + r);
    }
    auto below = [&](double x) { // eigenvalues < x
        size_t c = 0;
        double q = 1.0;
        for (size_t i = 0; i < k; ++i) {
            q = a[i] - x - (i > 0 ? b[i] * b[i] / q : 0.0);
            if (q == 0.0) q = -DBL_MIN;
            if (q < 0.0) ++c;
        }
        return c;
    };
    auto bisect = [&](size_t target) { // smallest x with `target` eigenvalues below
        double l = lo, h = hi;
        for (int it = 0; it < 200 && h - l > 4 * DBL_EPSILON * std::max(std::fabs(l), std::fabs(h));
             ++it) {
            const double mid = 0.5 * (l + h);
            if (below(mid) >= target)
                h = mid;
            else
                l = mid;
        }
        return 0.5 * (l + h);
    };
    lmin = bisect(1);
    lmax = bisect(k);
}

// Lanczos estimate of the extreme eigenvalues of (SA M)^T (SA M) from a fixed
// pseudo-random start (`steps` products with SA and M, one read-back).  For
// an exact R both are 1.  CholeskyQR from an FP32 Gram matrix cannot resolve
// a column of SA that depends on the others: its pivot stops at about
// sqrt(u_32) instead of 0 and the factor looks full rank, but SA M then has a
// singular value near 0, which the Ritz values find within a few steps when
// the rest of the spectrum clusters near 1.
static void precond_spectrum(long m, long d, const double *SA, long ldsa, const double *M,
                             long ldm, int steps, double &lmin, double &lmax, Arena &ar) {
    hipStream_t s = ctx().stream;
    const size_t mark = ar.used();
    const int K = (int)std::min<long>(steps, d);
    double *V = ar.alloc_n<double>((size_t)d * (K + 1));
    double *w = ar.alloc_n<double>((size_t)d);
    double *t = ar.alloc_n<double>((size_t)m);
    double *z = ar.alloc_n<double>((size_t)d);
    double *h = ar.alloc_n<double>((size_t)d);
    double *ab = ar.alloc_n<double>((size_t)2 * K);
    std::vector<double> v0((size_t)d);
    uint64_t st = 0x9E3779B97F4A7C15ull;
    double nn = 0.0;
    for (long i = 0; i < d; ++i) { // splitmix64, uniform in (-1, 1)
        uint64_t x = (st += 0x9E3779B97F4A7C15ull);
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        v0[(size_t)i] = ((double)(x >> 11) + 0.5) * 0x1.0p-52 - 1.0;
        nn += v0[(size_t)i] * v0[(size_t)i];
    }
    for (double &x : v0) x /= std::sqrt(nn);
    SKL_HIP(hipMemcpyAsync(V, v0.data(), sizeof(double) * d, hipMemcpyHostToDevice, s));
    for (int j = 0; j < K; ++j) {
        const double *vj = V + (size_t)j * d;
        dev_gemv_n_dense(d, d, M, ldm, vj, 1.0, 0.0, w, ar);
        dev_gemv_n_dense(m, d, SA, ldsa, w, 1.0, 0.0, t, ar);
        dev_gemv_t_dense(m, d, SA, ldsa, t, 1.0, 0.0, z, ar);
        dev_gemv_t_dense(d, d, M, ldm, z, 1.0, 0.0, h, ar);
        hipLaunchKernelGGL(k_lanczos_step, dim3(1), dim3(1024), 0, s, d, j, vj,
                           j > 0 ? vj - d : vj, h, V + (size_t)(j + 1) * d, ab);
    }
    SKL_CHECK_LAUNCH();
    std::vector<double> hab((size_t)2 * K);
    SKL_HIP(hipMemcpyAsync(hab.data(), ab, sizeof(double) * 2 * K, hipMemcpyDeviceToHost, s));
    SKL_HIP(hipStreamSynchronize(s));
    ar.rewind(mark);
    // T_k: alpha on the diagonal, beta_1 .. beta_{k-1} beside it; a beta at
    // rounding level ends the Krylov space (its Ritz values are exact)
    std::vector<double> a, b;
    for (int j = 0; j < K; ++j) {
        a.push_back(hab[(size_t)2 * j]);
        b.push_back(j > 0 ? hab[(size_t)2 * j - 1] : 0.0);
        if (!(hab[(size_t)2 * j + 1] > 1e-12 * std::fabs(hab[(size_t)2 * j]))) break;
    }
    tridiag_extremes(a, b, lmin, lmax);
}

// The optimistic factor of the mixed mode.  Only R (for M = R^{-1}) is
// needed -- the start is refined against the FP64 sketch and LSQR works on
// A M -- so any QR of SA serves, and CholeskyQR needs no per-column
// reduction: the Gram matrix of the column-equilibrated sketch W = SA D^{-1}
// (D = column norms) is one FP32 MFMA GEMM with FP32 accumulation, its
// Cholesky factor R1 (FP64, blocked) gives R = R1 D.  Equilibration keeps the
// Gram's condition number at cond(SA D^{-1})^2: FP32 resolves it while
// cond(SA D^{-1}) stays below about 1/sqrt(u_32) = 4e3; beyond that the
// Cholesky meets a non-positive pivot and the caller falls back to FP64.
// Returns false on that breakdown; dg (host) receives diag(R), maxn the
// largest column norm of SA.
template <class T>
// gram64: the Gram matrix of the equilibrated sketch in FP64 (the precision
// ablation SKL_CHOLQR64) instead of FP32; everything else alike.
static bool mixed_cholqr(long m, long d, const T *SA, long ldsa, double *Rd, double *Mi,
                         std::vector<double> &dg, double &maxn, skl_info &info, Arena &ar,
                         bool gram64 = false) {
    hipStream_t s = ctx().stream;
    const size_t mark = ar.used();
    double *sq = ar.alloc_n<double>((size_t)d);
    double *D = ar.alloc_n<double>((size_t)d);
    double *invD = ar.alloc_n<double>((size_t)d);
    hipLaunchKernelGGL(k_colnorm2<T>, dim3((unsigned)d), dim3(256), 0, s, m, d, SA, ldsa, sq);
    hipLaunchKernelGGL(k_norms_from_squares, dim3(gblocks(d)), dim3(256), 0, s, d, sq, D, invD);
    double *G64 = ar.alloc_n<double>((size_t)d * d);
    const dim3 eg(gblocks(m), (unsigned)std::min<long>(d, 65535));
    if (gram64) {
        double *W = ar.alloc_n<double>((size_t)m * d);
        hipLaunchKernelGGL(k_equil_to_double<T>, eg, dim3(256), 0, s, m, d, SA, ldsa, invD, W);
        gemm<double>('T', 'N', d, d, m, 1.0, W, m, W, m, 0.0, G64, d, 1, 0, 0, 0, true);
    } else {
        float *W = ar.alloc_n<float>((size_t)m * d);
        hipLaunchKernelGGL(k_equil_to_float<T>, eg, dim3(256), 0, s, m, d, SA, ldsa, invD, W);
        float *G = ar.alloc_n<float>((size_t)d * d);
        gemm<float>('T', 'N', d, d, m, 1.0f, W, m, W, m, 0.0f, G, d, 1, 0, 0, 0, true);
        copy2d<float, double>(d, d, G, d, G64, d);
    }
    double *R1 = ar.alloc_n<double>((size_t)d * d);
    bool ok = cholesky_upper(d, G64, d, R1, d, ar);
    std::vector<double> hD((size_t)d);
    SKL_HIP(hipMemcpyAsync(hD.data(), D, sizeof(double) * d, hipMemcpyDeviceToHost, s));
    if (ok) {
        double *M1 = G64; // the Gram is spent
        trinv_upper(d, R1, d, 0.0, M1, d, ar);
        hipLaunchKernelGGL(k_unequilibrate, dim3(gblocks(d), (unsigned)std::min<long>(d, 65535)),
                           dim3(256), 0, s, d, R1, M1, D, invD, Rd, Mi);
        double *ddg = ar.alloc_n<double>((size_t)d);
        hipLaunchKernelGGL(k_diag<double>, dim3(gblocks(d)), dim3(256), 0, s, d, Rd, d, ddg);
        dg.resize((size_t)d);
        SKL_HIP(hipMemcpyAsync(dg.data(), ddg, sizeof(double) * d, hipMemcpyDeviceToHost, s));
    }
    SKL_HIP(hipStreamSynchronize(s));
    maxn = 0.0;
    for (double v : hD) maxn = std::max(maxn, v);
    // Pivots of the equilibrated factor below sqrt(max(m, d) u_32) are where
    // a dependent column and a merely ill-conditioned one look alike; then
    // the Lanczos estimate of SA M decides (an eigenvalue of (SA M)^T SA M
    // below 0.1: a dependent column, or an FP32 factor too poor to keep).
    bool probe = false;
    const double amb = std::sqrt((double)std::max(m, d) * (double)FLT_EPSILON);
    for (long j = 0; ok && j < d; ++j)
        if (!(std::fabs(dg[(size_t)j]) > amb * hD[(size_t)j])) probe = true;
    if (probe) {
        const double *S64;
        long ld64 = ldsa;
        if constexpr (std::is_same<T, double>::value) {
            S64 = SA;
        } else {
            double *tmp = ar.alloc_n<double>((size_t)m * d);
            copy2d<T, double>(m, d, SA, ldsa, tmp, m);
            S64 = tmp;
            ld64 = m;
        }
        precond_spectrum(m, d, S64, ld64, Mi, d, 16, info.precond_lmin, info.precond_lmax, ar);
        if (!(info.precond_lmin > 0.1)) ok = false;
    }
    ar.rewind(mark);
    return ok;
}

__global__ void k_axpy_s(long n, double a, const double *x, double *y) { // y += a x
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}
__global__ void k_xpby(long n, const double *x, double b, double *y) { // y = x + b y
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] + b * y[i];
}

// The sketch-and-solve start from the FP64 sketch [SA | c] and the inverse M
// of the FP32 factor: CGLS on min ||c - (SA M) z|| from z = 0, y = M z.  SA M
// is nearly orthonormal (its normal matrix is I + E with ||E|| = O(u_32
// cond(SA D^-1)^2)), so the iteration converges in a few steps to the FP64
// sketched solution; it stops when the normal-equation residual ||M^T SA^T
// r|| has fallen by `tol` or after max_steps.  Without it the FP32 factor's
// own solution would leave A x0 - b far above the least-squares residual and
// LSQR would spend its first iterations closing the gap.  Returns the steps
// taken.
static int refine_sketched_solution(long m, long d, const double *SA, long ldsa,
                                    const double *c, const double *M, long ldm, double *y,
                                    int max_steps, double tol, Arena &ar) {
    hipStream_t s = ctx().stream;
    const size_t mark = ar.used();
    double *r = ar.alloc_n<double>((size_t)m);
    double *q = ar.alloc_n<double>((size_t)m);
    double *t = ar.alloc_n<double>((size_t)d);
    double *sv = ar.alloc_n<double>((size_t)d);
    double *p = ar.alloc_n<double>((size_t)d);
    double *h = ar.alloc_n<double>((size_t)d);
    double *z = ar.alloc_n<double>((size_t)d);
    const size_t mark2 = ar.used();
    SKL_HIP(hipMemcpyAsync(r, c, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_fill_d, dim3(gblocks(d)), dim3(256), 0, s, d, z, 0.0);
    dev_gemv_t_dense(m, d, SA, ldsa, r, 1.0, 0.0, t, ar);
    dev_gemv_t_dense(d, d, M, ldm, t, 1.0, 0.0, sv, ar);
    SKL_HIP(hipMemcpyAsync(p, sv, sizeof(double) * d, hipMemcpyDeviceToDevice, s));
    double gamma = dev_nrm2(d, sv, ar);
    gamma *= gamma;
    const double stop = tol * tol * gamma;
    int k = 0;
    while (k < max_steps && gamma > stop) {
        ar.rewind(mark2);
        dev_gemv_n_dense(d, d, M, ldm, p, 1.0, 0.0, h, ar);
        dev_gemv_n_dense(m, d, SA, ldsa, h, 1.0, 0.0, q, ar);
        double qn = dev_nrm2(m, q, ar);
        if (!(qn > 0.0)) break;
        const double alpha = gamma / (qn * qn);
        hipLaunchKernelGGL(k_axpy_s, dim3(gblocks(d)), dim3(256), 0, s, d, alpha, p, z);
        hipLaunchKernelGGL(k_axpy_s, dim3(gblocks(m)), dim3(256), 0, s, m, -alpha, q, r);
        dev_gemv_t_dense(m, d, SA, ldsa, r, 1.0, 0.0, t, ar);
        dev_gemv_t_dense(d, d, M, ldm, t, 1.0, 0.0, sv, ar);
        double gn = dev_nrm2(d, sv, ar);
        gn *= gn;
        hipLaunchKernelGGL(k_xpby, dim3(gblocks(d)), dim3(256), 0, s, d, sv, gn / gamma, p);
        gamma = gn;
        ++k;
    }
    ar.rewind(mark2);
    dev_gemv_n_dense(d, d, M, ldm, z, 1.0, 0.0, y, ar);
    SKL_CHECK_LAUNCH();
    ar.rewind(mark);
    return k;
}

// cpqr: ls_dense_hashing_blendenpik (sketch-and-solve start, CPQR, rank);
// otherwise LSQR from zero on the Householder-QR preconditioner, after the
// hashing sketch (ls_dense_hashing_blendenpik_noCPQR) or, with subsample,
// after uniform row sampling of the transform (ls_dense_blendenpik).
// AT_pre: a row-major copy of A (lsqr_make_rows) that outlives the call, or
// null (the copy is made per solve when the one-pass LSQR kernel is used)
static int dense_impl(long n, long d, const double *dA, long lda, const double *db, double *dx,
                      const skl_opts &o, skl_info &info, bool cpqr, bool subsample = false,
                      const double *AT_pre = nullptr) {
    Context &C = ctx();
    Arena &ar = C.cur_arena();
    hipStream_t s = C.stream;
    std::memset(&info, 0, sizeof info);
    if (n < d) throw Error(SKL_E_SHAPE, "Matrix is not overdetermined.");
    SKL_REQUIRE(d > 0 && n > 0 && lda >= n, SKL_E_ARG, "invalid dense shape");
    SKL_REQUIRE(o.k >= 1 && o.gamma > 0, SKL_E_ARG, "invalid sketch parameters");
    long m = (long)std::ceil((double)d * o.gamma);
    SKL_REQUIRE((subsample || m >= o.k) && m >= d, SKL_E_ARG, "sketch must have at least d rows");
    SKL_REQUIRE(!subsample || m <= n, SKL_E_ARG, "row sampling needs at most n sketch rows");
    info.m = m;
    ar.reset();
    ArenaTrim trim(ar);
    double T0 = wall();
    // ---- pattern (libc draw order of dense_helpers.cpp:92 / 12-14) ----
    Glibc g;
    g.seed(o.seed);
    Pattern P;
    gen_signs(g, n, P.D);
    if (subsample) {
        // random_sample seeds again from the same clock (dense_helpers.cpp:25),
        // so the row draws restart the stream the signs came from
        Glibc g2;
        g2.seed(o.seed);
        gen_subsample(g2, n, m, P);
    } else {
        gen_hashing(g, m, n, o.k, P);
        build_buckets(P);
    }
    const bool gram64 = (o.precision == SKL_CHOLQR64);
    bool mixed = (o.precision == SKL_MIXED) || gram64;
    DevPattern dp = upload_pattern(P, subsample ? 1.0 : 1 / std::sqrt((double)o.k), true, false,
                                   mixed, ar);
    double T1 = wall();
    info.t_pattern = T1 - T0;
    // ---- sketch + QR ----
    long ncol = cpqr ? d + 1 : d;
    int32_t *jpvt = ar.alloc_n<int32_t>((size_t)d);
    double *Rd = nullptr;
    long rank = d;
    double *cvec = ar.alloc_n<double>((size_t)d + 1); // Q^T c, first rank entries
    bool escalated = false;
    double cond_est = 0.0;
    double *Mi = ar.alloc_n<double>((size_t)d * d);
    double *W64 = nullptr; // mixed robust path: the FP64 sketch, kept for the x0 refinement
    // the row-major copy of A for LSQR's one-pass kernel is made on the aux
    // stream while the QR runs (after the sketch, which is bandwidth-bound too)
    LsqrProblem pre;
    pre.n = n;
    pre.d = d;
    pre.A = dA;
    pre.lda = lda;
    if (AT_pre) {
        pre.AT = AT_pre;
        pre.ldt = lsqr_rows_ld(d);
    }
    auto start_rows = [&]() {
        if (!pre.AT) lsqr_start_rows(pre, ar);
    };
    auto finish_qr = [&](auto *W) {
        // diagonal -> rank (first |R_ii| <= rcond), R to double, c
        std::vector<double> dg((size_t)d);
        double *ddg = ar.alloc_n<double>((size_t)d);
        hipLaunchKernelGGL(k_diag, dim3(gblocks(d)), dim3(256), 0, s, d, W, m, ddg);
        SKL_HIP(hipMemcpyAsync(dg.data(), ddg, sizeof(double) * d, hipMemcpyDeviceToHost, s));
        check_grid_fault(); // before R and the pivots are used
        rank = d;
        if (cpqr) {
            rank = 0;
            for (long i = 0; i < d; ++i) {
                if (std::fabs(dg[(size_t)i]) > o.rcond)
                    rank++;
                else
                    break;
            }
            hipLaunchKernelGGL(k_col_to_double, dim3(gblocks(d)), dim3(256), 0, s, d, W + d * m, cvec);
        }
        Rd = ar.alloc_n<double>((size_t)std::max<long>(rank, 1) * std::max<long>(rank, 1));
        upcast_upper(rank, W, m, Rd, rank);
        trinv_upper(rank, Rd, rank, 0.0, Mi, rank, ar);
    };
    if (!mixed) {
        double *W = ar.alloc_n<double>((size_t)m * ncol);
        double *tau = ar.alloc_n<double>((size_t)d);
        dense_sketch_qr<double>(n, d, m, dA, lda, cpqr ? db : nullptr, dp, cpqr, W, jpvt, tau,
                                &info.t_sketch, &info.t_qr, ar, start_rows);
        finish_qr(W);
    } else {
        // Optimistic FP32 preconditioner (mixed_cholqr): R from the FP32 Gram
        // matrix of the column-equilibrated sketch, no pivoting.  Pivoting
        // only matters when A is (nearly) rank deficient, and then the factor
        // shows it -- a Cholesky breakdown, a |R_ii| at or below rcond, or a
        // large condition number of the column-equilibrated R -- and the
        // column-pivoted factorization is redone in FP64.  For a
        // well-conditioned full-rank A the least-squares solution is unique,
        // whatever the preconditioner.  The robust path sketches in FP64: the
        // FP64 sketch refines the sketch-and-solve start and serves the
        // fallback without a second transform.  The fast path (no start)
        // sketches in FP32.
        float *W = nullptr;
        double t0 = wall();
        if (cpqr || gram64) {
            W64 = ar.alloc_n<double>((size_t)m * ncol);
            dense_sketch<double>(n, d, m, dA, lda, cpqr ? db : nullptr, dp, W64, ar);
        } else {
            W = ar.alloc_n<float>((size_t)m * d);
            dense_sketch<float>(n, d, m, dA, lda, nullptr, dp, W, ar);
        }
        sync();
        double t1 = wall();
        info.t_sketch += t1 - t0;
        start_rows();
        Rd = ar.alloc_n<double>((size_t)d * d);
        std::vector<double> dg;
        double maxn = 0.0;
        const bool ok = W64 ? mixed_cholqr<double>(m, d, W64, m, Rd, Mi, dg, maxn, info, ar, gram64)
                            : mixed_cholqr<float>(m, d, W, m, Rd, Mi, dg, maxn, info, ar);
        info.t_qr += wall() - t1;
        rank = d;
        if (ok && cpqr) {
            rank = 0;
            while (rank < d && std::fabs(dg[(size_t)rank]) > o.rcond) ++rank;
        }
        cond_est = ok ? scaled_cond(d, Rd, d, Mi, d, ar) : INFINITY;
        hipLaunchKernelGGL(k_iota, dim3(gblocks(d)), dim3(256), 0, s, d, jpvt);
        if (!ok || rank < d || !(cond_est <= o.escalate_cond)) {
            escalated = true;
            double *taud = ar.alloc_n<double>((size_t)d);
            if (W64) {
                double t2 = wall();
                dense_factor<double>(m, ncol, d, cpqr, W64, jpvt, taud, ar);
                sync();
                info.t_qr += wall() - t2;
                finish_qr(W64);
                W64 = nullptr; // factored in place
            } else {
                double *Wd = ar.alloc_n<double>((size_t)m * ncol);
                dense_sketch_qr<double>(n, d, m, dA, lda, nullptr, dp, cpqr, Wd, jpvt, taud,
                                        &info.t_sketch, &info.t_qr, ar);
                finish_qr(Wd);
            }
        }
    }
    info.escalated = escalated ? 1 : 0;
    info.cond_est = cond_est;
    info.rank = rank;
    sync();
    double T2 = wall();
    if (rank == 0) throw Error(SKL_E_RANK0, "rank 0: all |R_ii| <= rcond");
    LsqrProblem p;
    p.n = n;
    p.d = d;
    p.r = rank;
    p.A = dA;
    p.lda = lda;
    p.E = cpqr ? jpvt : nullptr;
    p.M = Mi;
    p.ldm = rank;
    p.AT = pre.AT;
    p.ldt = pre.ldt;
    p.AT_ready = pre.AT_ready;
    double *y = ar.alloc_n<double>((size_t)rank);
    if (!cpqr) {
        // LSQR on b directly, x = M y
        LsqrResult res = lsqr_solve(p, db, o.it_tol, o.max_it, y, o.use_graph != 0, o.graph_chunk,
                                    t_hist, t_hist_len, ar);
        double *my = ar.alloc_n<double>((size_t)rank);
        trmv_n(rank, Mi, rank, y, my, ar);
        hipLaunchKernelGGL(k_scatter_E, dim3(gblocks(rank)), dim3(256), 0, s, d, rank,
                           (const int32_t *)nullptr, my, dx);
        sync();
        info.flag = res.flag;
        info.it = res.it;
        info.lsqr_nrm_ar = res.nrm_ar;
        info.lsqr_nrm_a = res.nrm_a;
        info.lsqr_nrm_r = res.nrm_r;
        double T3 = wall();
        info.t_lsqr = T3 - T2;
        info.t_total = T3 - T0;
        return SKL_OK;
    }
    // x0 = P M c ; rs = b - A x0
    double *c1 = ar.alloc_n<double>((size_t)rank);
    if (W64) {
        // mixed, no fallback: the sketch-and-solve start from the FP64 sketch
        // and the optimistic factor, refined from zero
        refine_sketched_solution(m, d, W64, m, W64 + d * m, Mi, rank, c1, 40, 1e-11, ar);
    } else {
        trmv_n(rank, Mi, rank, cvec, c1, ar);
    }
    double *x0 = ar.alloc_n<double>((size_t)d);
    hipLaunchKernelGGL(k_fill_d, dim3(gblocks(d)), dim3(256), 0, s, d, x0, 0.0);
    hipLaunchKernelGGL(k_scatter_E, dim3(gblocks(rank)), dim3(256), 0, s, d, rank, jpvt, c1, x0);
    double *rs = ar.alloc_n<double>((size_t)n);
    SKL_HIP(hipMemcpyAsync(rs, db, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    dev_gemv_n_dense(n, d, dA, lda, x0, -1.0, 1.0, rs, ar);
    double r0 = dev_nrm2(n, rs, ar);
    info.x0_residual = r0;
    double T3 = wall();
    info.t_x0 = T3 - T2;
    if (r0 < o.abs_tol) {
        SKL_HIP(hipMemcpyAsync(dx, x0, sizeof(double) * d, hipMemcpyDeviceToDevice, s));
        sync();
        info.it = 0;
        info.flag = 0;
        info.t_total = wall() - T0;
        return SKL_OK;
    }
    LsqrResult res = lsqr_solve(p, rs, o.it_tol, o.max_it, y, o.use_graph != 0, o.graph_chunk,
                                t_hist, t_hist_len, ar);
    double *my = ar.alloc_n<double>((size_t)rank);
    trmv_n(rank, Mi, rank, y, my, ar);
    hipLaunchKernelGGL(k_fill_d, dim3(gblocks(d)), dim3(256), 0, s, d, dx, 0.0);
    hipLaunchKernelGGL(k_scatter_E, dim3(gblocks(rank)), dim3(256), 0, s, d, rank, jpvt, my, dx);
    hipLaunchKernelGGL(k_axpy_d, dim3(gblocks(d)), dim3(256), 0, s, d, x0, dx);
    sync();
    info.flag = res.flag;
    info.it = res.it;
    info.lsqr_nrm_ar = res.nrm_ar;
    info.lsqr_nrm_a = res.nrm_a;
    info.lsqr_nrm_r = res.nrm_r;
    double T4 = wall();
    info.t_lsqr = T4 - T3;
    info.t_total = T4 - T0;
    return SKL_OK;
}

// ---------------------------------------------------------------------------
// Sparse solve, device CSC with int64 indices
// ---------------------------------------------------------------------------
// [SA | S b] into W (m x (d + 1)): the CXSparse product orders of
// sketching.cpp:104-111
template <class T>
static void sparse_sketch_full(long n, long d, long m, long k, const DevPattern &dp,
                               const int32_t *cp, const int32_t *ri, const double *val, long nnz,
                               const double *db, T *W, Arena &ar) {
    sparse_sketch_t<T>(m, d, k, dp.rows, dp.vals, cp, ri, val, nnz, W, m, ar);
    // b_s = S b in cs_gaxpy order, as the last column
    const size_t mark = ar.used();
    double *bs = ar.alloc_n<double>((size_t)m);
    hash_gather<double>(m, 1, dp.bptr, dp.bsrc, dp.coef_d, db, n, bs, m);
    copy2d<double, T>(m, 1, bs, m, W + d * m, m);
    sync();
    ar.rewind(mark);
}

// Sketch [SA | S b] (or copy it from src, an FP64 sketch made before) and
// factor it as the reference does: natural-order Householder QR, and when any
// |R_jj| falls to SPQR's default tolerance 20 (m + d) eps max_j ||SA_j|| a
// column-pivoted refactorisation decides the rank.
template <class T>
static long sparse_factor(long n, long d, long m, long k, const DevPattern &dp, const int32_t *cp,
                          const int32_t *ri, const double *val, long nnz, const double *db, T *W,
                          T *tau, int32_t *jpvt, double *tol_out, Arena &ar,
                          const double *src = nullptr) {
    hipStream_t s = ctx().stream;
    auto sketch = [&] {
        if (src)
            copy2d<double, T>(m, d + 1, src, m, W, m);
        else
            sparse_sketch_full<T>(n, d, m, k, dp, cp, ri, val, nnz, db, W, ar);
    };
    sketch();
    // SPQR's default tolerance 20 (m + d) eps max_j ||SA_j||
    double *cn = ar.alloc_n<double>((size_t)d);
    hipLaunchKernelGGL(k_colnorm2<T>, dim3((unsigned)d), dim3(256), 0, s, m, d, W, m, cn);
    std::vector<double> hc((size_t)d);
    SKL_HIP(hipMemcpyAsync(hc.data(), cn, sizeof(double) * d, hipMemcpyDeviceToHost, s));
    sync();
    double maxn = 0;
    for (double v : hc) maxn = std::max(maxn, std::sqrt(v));
    double eps = std::is_same<T, float>::value ? (double)FLT_EPSILON : DBL_EPSILON;
    double tol = 20.0 * (double)(m + d) * eps * maxn;
    *tol_out = tol;
    qr_householder<T>(m, d + 1, d, W, m, tau, ar);
    long kmax = std::min(m, d);
    std::vector<double> dg((size_t)kmax);
    double *ddg = ar.alloc_n<double>((size_t)kmax);
    hipLaunchKernelGGL(k_diag<T>, dim3(gblocks(kmax)), dim3(256), 0, s, kmax, W, m, ddg);
    SKL_HIP(hipMemcpyAsync(dg.data(), ddg, sizeof(double) * kmax, hipMemcpyDeviceToHost, s));
    check_grid_fault();
    bool full = true;
    for (long i = 0; i < kmax; ++i)
        if (!(std::fabs(dg[(size_t)i]) > tol)) {
            full = false;
            break;
        }
    hipLaunchKernelGGL(k_iota, dim3(gblocks(d)), dim3(256), 0, s, d, jpvt);
    if (full) return kmax;
    // rank deficient: column-pivoted refactorisation of a fresh sketch (the
    // CPQR of the dense path: dgeqp3 pivoting below 1000 columns, HQRRP above)
    sketch();
    dense_factor<T>(m, d + 1, d, true, W, jpvt, tau, ar);
    hipLaunchKernelGGL(k_diag<T>, dim3(gblocks(kmax)), dim3(256), 0, s, kmax, W, m, ddg);
    SKL_HIP(hipMemcpyAsync(dg.data(), ddg, sizeof(double) * kmax, hipMemcpyDeviceToHost, s));
    check_grid_fault();
    long rank = 0;
    for (long i = 0; i < kmax; ++i) {
        if (std::fabs(dg[(size_t)i]) > tol)
            rank++;
        else
            break;
    }
    return rank;
}

// The int32 CSC indices and the CSR copy of a sparse A, when they outlive the
// call (a handle); otherwise sparse_impl builds them per solve.
struct SparsePrep {
    const int32_t *cp = nullptr, *ri = nullptr; // CSC column pointers / row indices
    const int32_t *rp = nullptr, *cc = nullptr; // CSR row pointers / column indices
    const double *cv = nullptr;                 // CSR values
};

static int sparse_impl(long n, long d, long nnz, const long *dcp, const long *dri,
                       const double *dval, const double *db, double *dx, const skl_opts &o,
                       skl_info &info, const SparsePrep *prep = nullptr) {
    Context &C = ctx();
    Arena &ar = C.arena;
    hipStream_t s = C.stream;
    std::memset(&info, 0, sizeof info);
    if (n < d) throw Error(SKL_E_SHAPE, "Matrix is not overdetermined.");
    SKL_REQUIRE(d > 0 && n > 0 && nnz >= 0, SKL_E_ARG, "invalid sparse shape");
    SKL_REQUIRE(nnz < 2147483647L && n < 2147483647L, SKL_E_UNSUPPORTED, "nnz must fit in int32");
    long m = (long)std::ceil((double)d * o.gamma);
    SKL_REQUIRE(m >= o.k && m >= 1, SKL_E_ARG, "invalid sketch parameters");
    info.m = m;
    ar.reset();
    ArenaTrim trim(ar);
    double T0 = wall();
    const int32_t *cp, *ri, *rp, *cc;
    const double *cv;
    if (prep) {
        cp = prep->cp;
        ri = prep->ri;
        rp = prep->rp;
        cc = prep->cc;
        cv = prep->cv;
    } else {
        int32_t *cp_ = ar.alloc_n<int32_t>((size_t)d + 1);
        int32_t *ri_ = ar.alloc_n<int32_t>((size_t)std::max<long>(nnz, 1));
        narrow_indices(d + 1, dcp, cp_);
        narrow_indices(nnz, dri, ri_);
        int32_t *rp_ = ar.alloc_n<int32_t>((size_t)n + 1);
        int32_t *cc_ = ar.alloc_n<int32_t>((size_t)std::max<long>(nnz, 1));
        double *cv_ = ar.alloc_n<double>((size_t)std::max<long>(nnz, 1));
        csc_to_csr(n, d, nnz, cp_, ri_, dval, rp_, cc_, cv_, ar);
        cp = cp_;
        ri = ri_;
        rp = rp_;
        cc = cc_;
        cv = cv_;
    }
    // pattern: no srand on the sparse path (sketching.cpp:100); seed = the
    // state of the process-wide generator, 1 for a fresh process
    Glibc g;
    g.seed(o.seed);
    Pattern P;
    gen_hashing(g, m, n, o.k, P);
    build_buckets(P);
    const bool gram64 = (o.precision == SKL_CHOLQR64);
    bool mixed = (o.precision == SKL_MIXED) || gram64;
    DevPattern dp = upload_pattern(P, 1.0, false, true, false, ar);
    double T1 = wall();
    info.t_pattern = T1 - T0;
    int32_t *jpvt = ar.alloc_n<int32_t>((size_t)d);
    double *cvec = ar.alloc_n<double>((size_t)d + 1);
    double *Rd = nullptr;
    long rank = 0;
    double tol = 0;
    auto extract = [&](auto *W) {
        Rd = ar.alloc_n<double>((size_t)std::max<long>(rank, 1) * std::max<long>(rank, 1));
        upcast_upper(rank, W, m, Rd, rank);
        hipLaunchKernelGGL(k_col_to_double, dim3(gblocks(rank)), dim3(256), 0, s, rank, W + d * m, cvec);
    };
    bool escalated = false;
    double cond_est = 0;
    double *Mi = ar.alloc_n<double>((size_t)d * d);
    double *W64 = nullptr; // mixed: the FP64 sketch, kept for the x0 refinement
    double tq0 = wall();
    if (!mixed) {
        double *W = ar.alloc_n<double>((size_t)m * (d + 1));
        double *tau = ar.alloc_n<double>((size_t)d + 1);
        rank = sparse_factor<double>(n, d, m, o.k, dp, cp, ri, dval, nnz, db, W, tau, jpvt, &tol, ar);
        if (rank == 0) throw Error(SKL_E_RANK0, "SPQR Factorization fails");
        extract(W);
        trinv_upper(rank, Rd, rank, 0.0, Mi, rank, ar);
    } else {
        // FP64 sketch, optimistic CholeskyQR factor (see dense_impl).  The
        // FP32 rank test: |R_jj| above max(m, d) u_32 max_j ||SA_j||, the
        // resolution of an FP32 factor, else the FP64 factorization decides.
        W64 = ar.alloc_n<double>((size_t)m * (d + 1));
        sparse_sketch_full<double>(n, d, m, o.k, dp, cp, ri, dval, nnz, db, W64, ar);
        Rd = ar.alloc_n<double>((size_t)d * d);
        std::vector<double> dg;
        double maxn = 0.0;
        const bool ok = mixed_cholqr<double>(m, d, W64, m, Rd, Mi, dg, maxn, info, ar, gram64);
        bool short_rank = !ok || m < d;
        if (!short_rank) {
            tol = (double)std::max(m, d) * (double)FLT_EPSILON * maxn;
            for (long j = 0; j < d; ++j)
                if (!(std::fabs(dg[(size_t)j]) > tol)) {
                    short_rank = true;
                    break;
                }
        }
        rank = d;
        hipLaunchKernelGGL(k_iota, dim3(gblocks(d)), dim3(256), 0, s, d, jpvt);
        cond_est = short_rank ? INFINITY : scaled_cond(d, Rd, d, Mi, d, ar);
        if (short_rank || !(cond_est <= o.escalate_cond)) {
            escalated = true;
            double *Wd = ar.alloc_n<double>((size_t)m * (d + 1));
            double *taud = ar.alloc_n<double>((size_t)d + 1);
            rank = sparse_factor<double>(n, d, m, o.k, dp, cp, ri, dval, nnz, db, Wd, taud, jpvt,
                                         &tol, ar, W64);
            if (rank == 0) throw Error(SKL_E_RANK0, "SPQR Factorization fails");
            extract(Wd);
            trinv_upper(rank, Rd, rank, 0.0, Mi, rank, ar);
            W64 = nullptr;
        }
    }
    info.escalated = escalated;
    info.cond_est = cond_est;
    info.rank = rank;
    double rc = rcond_from(rank, Rd, rank, Mi, rank, ar);
    info.rcond = rc;
    info.success = rc > o.rcond_threshold ? 1 : 0;
    double *Mp = Mi; // LSQR operator
    if (!info.success) {
        // LSQR operator uses (R11 + perturb I)^{-1} (sparse_helpers.cpp:138-147)
        Mp = ar.alloc_n<double>((size_t)rank * rank);
        trinv_upper(rank, Rd, rank, o.perturb, Mp, rank, ar);
    }
    sync();
    double T2 = wall();
    info.t_sketch = 0;
    info.t_qr = T2 - tq0;
    // x0 = P M c (no perturbation, ls_sparse_spqr.cpp:144)
    double *c1 = ar.alloc_n<double>((size_t)rank);
    if (W64) {
        // mixed, no fallback: the start from the FP64 sketch, refined from zero
        refine_sketched_solution(m, d, W64, m, W64 + d * m, Mi, rank, c1, 40, 1e-11, ar);
    } else {
        trmv_n(rank, Mi, rank, cvec, c1, ar);
    }
    double *x0 = ar.alloc_n<double>((size_t)d);
    hipLaunchKernelGGL(k_fill_d, dim3(gblocks(d)), dim3(256), 0, s, d, x0, 0.0);
    hipLaunchKernelGGL(k_scatter_E, dim3(gblocks(rank)), dim3(256), 0, s, d, rank, jpvt, c1, x0);
    double *rs = ar.alloc_n<double>((size_t)n);
    SKL_HIP(hipMemcpyAsync(rs, db, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    csr_gemv(n, rp, cc, cv, x0, -1.0, 1.0, rs);
    double r0 = dev_nrm2(n, rs, ar);
    info.x0_residual = r0;
    double T3 = wall();
    info.t_x0 = T3 - T2;
    if (r0 < o.abs_tol) {
        SKL_HIP(hipMemcpyAsync(dx, x0, sizeof(double) * d, hipMemcpyDeviceToDevice, s));
        sync();
        info.it = 0;
        info.flag = 0;
        info.t_total = wall() - T0;
        return SKL_OK;
    }
    LsqrProblem p;
    p.n = n;
    p.d = d;
    p.r = rank;
    p.csr_ptr = rp;
    p.csr_col = cc;
    p.csr_val = cv;
    p.csc_ptr = cp;
    p.csc_row = ri;
    p.csc_val = dval;
    p.nnz = nnz;
    p.E = jpvt;
    p.M = Mp;
    p.ldm = rank;
    double *y = ar.alloc_n<double>((size_t)rank);
    LsqrResult res = lsqr_solve(p, rs, o.it_tol, o.max_it, y, o.use_graph != 0, o.graph_chunk,
                                t_hist, t_hist_len, ar);
    // y <- R11^{-1} y (no perturbation, ls_sparse_spqr.cpp:188); y += c; x = P y
    double *my = ar.alloc_n<double>((size_t)rank);
    trmv_n(rank, Mi, rank, y, my, ar);
    hipLaunchKernelGGL(k_axpy_d, dim3(gblocks(rank)), dim3(256), 0, s, rank, c1, my);
    hipLaunchKernelGGL(k_fill_d, dim3(gblocks(d)), dim3(256), 0, s, d, dx, 0.0);
    hipLaunchKernelGGL(k_scatter_E, dim3(gblocks(rank)), dim3(256), 0, s, d, rank, jpvt, my, dx);
    sync();
    info.flag = res.flag;
    info.it = res.it;
    info.lsqr_nrm_ar = res.nrm_ar;
    info.lsqr_nrm_a = res.nrm_a;
    info.lsqr_nrm_r = res.nrm_r;
    double T4 = wall();
    info.t_lsqr = T4 - T3;
    info.t_total = T4 - T0;
    return SKL_OK;
}

} // namespace skl

// ===========================================================================
// C ABI
// ===========================================================================
using namespace skl;

template <class F> static int guarded(F &&f) {
    try {
        skl::t_err.clear();
        return f();
    } catch (const skl::Error &e) {
        skl::t_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        skl::t_err = "host allocation failed";
        return SKL_E_NOMEM;
    } catch (const std::exception &e) {
        skl::t_err = e.what();
        return SKL_E_HIP;
    }
}

static skl_opts resolve(const skl_opts *o) {
    skl_opts r;
    if (o)
        r = *o;
    else
        skl_default_opts(&r);
    if (r.graph_chunk <= 0) r.graph_chunk = 16;
    return r;
}

namespace skl {
// Host-pointer variants: inputs are staged in device buffers owned by the call.
struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) {
        AllocLock lk(capture_mutex());
        SKL_HIP(hipMalloc(&p, std::max<size_t>(bytes, 8)));
    }
    ~DevBuf() {
        if (p) {
            AllocLock lk(capture_mutex());
            (void)hipFree(p);
        }
    }
    template <class T> T *as() { return static_cast<T *>(p); }
};

static int dense_host(long n, long d, const double *A, long lda, const double *b, double *x,
                      const skl_opts *opts, skl_info *info, bool cpqr, bool subsample = false) {
    skl_info tmp;
    return guarded([&] {
        if (n < d) throw Error(SKL_E_SHAPE, "Matrix is not overdetermined.");
        SKL_REQUIRE(A && b && x && lda >= n && d > 0, SKL_E_ARG, "invalid arguments");
        skl_opts o = resolve(opts);
        ensure_init();
        hipStream_t s = ctx().stream;
        DevBuf dA(sizeof(double) * (size_t)n * d), db(sizeof(double) * n), dx(sizeof(double) * d);
        SKL_HIP(hipMemcpy2DAsync(dA.p, sizeof(double) * n, A, sizeof(double) * lda,
                                 sizeof(double) * n, d, hipMemcpyHostToDevice, s));
        SKL_HIP(hipMemcpyAsync(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
        SKL_HIP(hipStreamSynchronize(s));
        int rc = dense_impl(n, d, dA.as<double>(), n, db.as<double>(), dx.as<double>(), o,
                            info ? *info : tmp, cpqr, subsample);
        SKL_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * d, hipMemcpyDeviceToHost, s));
        SKL_HIP(hipStreamSynchronize(s));
        return rc;
    });
}

static Pattern pattern_from(long m, long n, long k, const long *rows, const double *vals) {
    Pattern P;
    P.m = m;
    P.n = n;
    P.k = k;
    P.rows.resize((size_t)(n * k));
    P.vals.assign(vals, vals + n * k);
    for (long q = 0; q < n * k; ++q) {
        SKL_REQUIRE(rows[q] >= 0 && rows[q] < m, SKL_E_ARG, "pattern row out of range");
        P.rows[(size_t)q] = (int32_t)rows[q];
    }
    build_buckets(P);
    return P;
}

template <class T>
static void download_as_double(const T *src, long rows, long cols, long lds, double *dst, long ldd) {
    std::vector<T> h((size_t)rows * cols);
    SKL_HIP(hipMemcpy2D(h.data(), sizeof(T) * rows, src, sizeof(T) * lds, sizeof(T) * rows, cols,
                        hipMemcpyDeviceToHost));
    for (long j = 0; j < cols; ++j)
        for (long i = 0; i < rows; ++i) dst[i + j * ldd] = (double)h[(size_t)(i + j * rows)];
}

template <class T>
static int qr_component(long m, long n, long nfac, double *A, long lda, double *tau, long *jpvt,
                        bool cpqr) {
    Arena &ar = ctx().arena;
    ar.reset();
    T *W = ar.alloc_n<T>((size_t)m * n);
    std::vector<T> h((size_t)m * n);
    for (long j = 0; j < n; ++j)
        for (long i = 0; i < m; ++i) h[(size_t)(i + j * m)] = (T)A[i + j * lda];
    SKL_HIP(hipMemcpy(W, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    T *dt = ar.alloc_n<T>((size_t)nfac + 1);
    int32_t *jp = ar.alloc_n<int32_t>((size_t)nfac + 1);
    if (cpqr) {
        hipLaunchKernelGGL(k_iota, dim3(gblocks(nfac)), dim3(256), 0, ctx().stream, nfac, jp);
        if (std::min(m, nfac) < 1000)
            cpqr_greedy<T>(m, n, nfac, W, m, jp, dt, ar);
        else
            cpqr_hqrrp<T>(m, n, nfac, W, m, jp, dt, gaussian_for(m).data(), ar);
    } else {
        qr_householder<T>(m, n, nfac, W, m, dt, ar);
    }
    check_grid_fault();
    download_as_double(W, m, n, m, A, lda);
    long kk = std::min(m, nfac);
    std::vector<T> ht((size_t)kk);
    SKL_HIP(hipMemcpy(ht.data(), dt, sizeof(T) * kk, hipMemcpyDeviceToHost));
    for (long i = 0; i < kk; ++i) tau[i] = (double)ht[(size_t)i];
    if (cpqr && jpvt) {
        std::vector<int32_t> hj((size_t)nfac);
        SKL_HIP(hipMemcpy(hj.data(), jp, sizeof(int32_t) * nfac, hipMemcpyDeviceToHost));
        for (long i = 0; i < nfac; ++i) jpvt[i] = hj[(size_t)i] + 1;
    }
    return SKL_OK;
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation entry point (skl_dense_solve_dev_noalloc): a
// static device array, one copy per GPU the library runs on, held until the
// process ends.  Host code never uses the array's name as an address: its
// device address comes from hipGetSymbolAddress at context set-up, once per
// device, and that pointer is what the solver carves its buffers from.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;
// the calls of this process share each device's array: one at a time
static std::mutex g_noalloc_call_mu;

void noalloc_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= NOALLOC_MAX_DEVICES || g_scratch_ptr[device]) return;
    void *p = nullptr;
    SKL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
}

static void *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return (device >= 0 && device < NOALLOC_MAX_DEVICES) ? g_scratch_ptr[device] : nullptr;
}

// Device scratch of the robust dense FP64 solve without the row-major copy
// of A (two-pass LSQR), a bound used to refuse a problem before any work
// is queued: the sketch [SA | c], M and R, the pattern, the CPQR's panel and
// update buffers, the LSQR vectors and partial sums, and the transform's
// column chunks.
static size_t noalloc_footprint(long n, long d, long m, long k) {
    const double b = 8.0 * ((double)m * (d + 1) + 4.0 * d * d + 6.0 * n + 2.0 * n * k + 80.0 * m) +
                     16.0 * d * ((double)n / 1024.0 + 64.0) + 32.0 * n + (4 << 20);
    return (size_t)std::min(b, 1e18);
}

} // namespace skl

extern "C" {

void skl_default_opts(skl_opts *o) {
    std::memset(o, 0, sizeof *o);
    // include/bench_config.hpp:7-20 defaults
    o->gamma = 1.1;
    o->k = 3;
    o->abs_tol = 1e-8;
    o->rcond = 1e-10;
    o->it_tol = 1e-6;
    o->max_it = 10000;
    o->rcond_threshold = 1e-10;
    o->perturb = 1e-10;
    o->ordering = 2;
    o->seed = 1;
    o->precision = SKL_FP64;
    o->escalate_cond = 1e6;
    o->use_graph = 1;
    o->graph_chunk = 16;
}

int skl_init(int device) {
    return guarded([&] {
        skl::set_device(device);
        skl::ensure_init();
        return SKL_OK;
    });
}

void skl_finalize(void) { skl::release_thread_context(); }

const char *skl_error_string(int code) {
    switch (code) {
    case SKL_OK: return "ok";
    case SKL_E_SHAPE: return "Matrix is not overdetermined.";
    case SKL_E_RANK0: return "SPQR Factorization fails";
    case SKL_E_NOMEM: return "out of device memory";
    case SKL_E_HIP: return "HIP runtime error";
    case SKL_E_ARG: return "invalid argument";
    case SKL_E_UNSUPPORTED: return "unsupported size";
    case SKL_E_NODEVICE: return "no GPU available";
    default: return "unknown error";
    }
}

const char *skl_last_error(void) { return skl::t_err.c_str(); }

int skl_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

void skl_set_barrier_spins(long spins) {
    try {
        skl::ctx().barrier_spins =
            spins < 0 ? (1u << 28) : (unsigned)std::min<long>(spins, 0xffffffffL);
    } catch (...) {
    }
}

void skl_set_history(skl_lsqr_step *steps, long max_steps) {
    skl::t_hist = steps;
    skl::t_hist_len = steps ? max_steps : 0;
}

int skl_dense_solve_dev(long n, long d, const double *dA, long lda, const double *db, double *dx,
                        const skl_opts *opts, skl_info *info) {
    skl_info tmp;
    return guarded([&] {
        skl_opts o = resolve(opts);
        return dense_impl(n, d, dA, lda, db, dx, o, info ? *info : tmp, true);
    });
}

int skl_dense_solve_dev_noalloc(long n, long d, const double *dA, long lda, const double *db,
                                double *dx, const skl_opts *opts, skl_info *info) {
    using namespace skl;
    skl_info tmp;
    return guarded([&] {
        if (!has_context())
            throw Error(SKL_E_ARG, "skl_dense_solve_dev_noalloc: call skl_init on this thread first");
        Context &C = ctx();
        skl_opts o = resolve(opts);
        SKL_REQUIRE(o.precision == SKL_FP64, SKL_E_UNSUPPORTED,
                    "skl_dense_solve_dev_noalloc runs the FP64 solver only");
        if (n < d) throw Error(SKL_E_SHAPE, "Matrix is not overdetermined.");
        SKL_REQUIRE(d > 0 && n > 0 && lda >= n && o.gamma > 0 && o.k >= 1, SKL_E_ARG,
                    "invalid dense shape or sketch parameters");
        const long m = (long)std::ceil((double)d * o.gamma);
        const size_t need = noalloc_footprint(n, d, m, o.k);
        if (need > NOALLOC_SCRATCH_BYTES)
            throw Error(SKL_E_UNSUPPORTED,
                        "skl_dense_solve_dev_noalloc: " + std::to_string(n) + " x " +
                            std::to_string(d) + " needs about " + std::to_string(need >> 20) +
                            " MiB of device scratch, above its " +
                            std::to_string(NOALLOC_SCRATCH_BYTES >> 20) + " MiB");
        void *p = noalloc_scratch(C.device);
        if (!p) throw Error(SKL_E_HIP, "no-allocation scratch not set up on this device");
        std::lock_guard<std::mutex> lk(g_noalloc_call_mu);
        struct Mode {
            Context &c;
            ~Mode() { c.no_alloc = false; }
        } mode{C};
        C.fixed.attach_fixed(p, NOALLOC_SCRATCH_BYTES);
        C.no_alloc = true;
        return dense_impl(n, d, dA, lda, db, dx, o, info ? *info : tmp, true);
    });
}

int skl_dense_solve_nocpqr_dev(long n, long d, const double *dA, long lda, const double *db,
                               double *dx, const skl_opts *opts, skl_info *info) {
    skl_info tmp;
    return guarded([&] {
        skl_opts o = resolve(opts);
        return dense_impl(n, d, dA, lda, db, dx, o, info ? *info : tmp, false);
    });
}

int skl_dense_solve_blendenpik_dev(long n, long d, const double *dA, long lda, const double *db,
                                   double *dx, const skl_opts *opts, skl_info *info) {
    skl_info tmp;
    return guarded([&] {
        skl_opts o = resolve(opts);
        return dense_impl(n, d, dA, lda, db, dx, o, info ? *info : tmp, false, true);
    });
}

int skl_sparse_solve_dev(long n, long d, long nnz, const long *dcp, const long *dri,
                         const double *dval, const double *db, double *dx, const skl_opts *opts,
                         skl_info *info) {
    skl_info tmp;
    return guarded([&] {
        skl_opts o = resolve(opts);
        return sparse_impl(n, d, nnz, dcp, dri, dval, db, dx, o, info ? *info : tmp);
    });
}

int skl_dense_solve(long n, long d, const double *A, long lda, const double *b, double *x,
                    const skl_opts *opts, skl_info *info) {
    return dense_host(n, d, A, lda, b, x, opts, info, true);
}
int skl_dense_solve_nocpqr(long n, long d, const double *A, long lda, const double *b, double *x,
                           const skl_opts *opts, skl_info *info) {
    return dense_host(n, d, A, lda, b, x, opts, info, false);
}
int skl_dense_solve_blendenpik(long n, long d, const double *A, long lda, const double *b,
                               double *x, const skl_opts *opts, skl_info *info) {
    return dense_host(n, d, A, lda, b, x, opts, info, false, true);
}

int skl_sparse_solve(long n, long d, long nnz, const long *colptr, const long *rowind,
                     const double *val, const double *b, double *x, const skl_opts *opts,
                     skl_info *info) {
    skl_info tmp;
    return guarded([&] {
        if (n < d) throw Error(SKL_E_SHAPE, "Matrix is not overdetermined.");
        SKL_REQUIRE(colptr && b && x && d > 0, SKL_E_ARG, "invalid arguments");
        SKL_REQUIRE(colptr[0] == 0 && colptr[d] == nnz, SKL_E_ARG, "colptr inconsistent with nnz");
        skl_opts o = resolve(opts);
        ensure_init();
        hipStream_t s = ctx().stream;
        DevBuf dcp(sizeof(long) * (d + 1)), dri(sizeof(long) * nnz), dv(sizeof(double) * nnz),
            db(sizeof(double) * n), dx(sizeof(double) * d);
        SKL_HIP(hipMemcpyAsync(dcp.p, colptr, sizeof(long) * (d + 1), hipMemcpyHostToDevice, s));
        if (nnz > 0) {
            SKL_HIP(hipMemcpyAsync(dri.p, rowind, sizeof(long) * nnz, hipMemcpyHostToDevice, s));
            SKL_HIP(hipMemcpyAsync(dv.p, val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
        }
        SKL_HIP(hipMemcpyAsync(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice, s));
        SKL_HIP(hipStreamSynchronize(s));
        int rc = sparse_impl(n, d, nnz, dcp.as<long>(), dri.as<long>(), dv.as<double>(),
                             db.as<double>(), dx.as<double>(), o, info ? *info : tmp);
        SKL_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * d, hipMemcpyDeviceToHost, s));
        SKL_HIP(hipStreamSynchronize(s));
        return rc;
    });
}

// ---- HBM-resident problems --------------------------------------------------
// A handle owns a device copy of A together with what every solve on it would
// otherwise rebuild: for dense A the row-major copy read by the one-pass LSQR
// kernel, for CSC A the int32 indices and the CSR copy.  Solves only read it,
// so any number may run on one handle at once (from server slots or threads).
struct skl_handle {
    int device = 0;
    int sparse = 0;
    long n = 0, d = 0, nnz = 0;
    double *A = nullptr;  // dense, n x d, ld n
    double *AT = nullptr; // dense, row-major copy (ld lsqr_rows_ld(d)) or null
    int32_t *cp = nullptr, *ri = nullptr, *rp = nullptr, *cc = nullptr;
    double *val = nullptr, *cv = nullptr;
    size_t bytes = 0;
    std::vector<void *> blocks;
    // a device block of `sz` bytes, freed with the handle
    void *take(size_t sz) {
        void *p = nullptr;
        sz = std::max<size_t>(sz, 8);
        {
            AllocLock lk(capture_mutex());
            SKL_HIP(hipMalloc(&p, sz));
        }
        blocks.push_back(p);
        bytes += sz;
        return p;
    }
    ~skl_handle() {
        if (blocks.empty()) return;
        AllocLock lk(capture_mutex());
        for (void *p : blocks) (void)hipFree(p);
    }
};

namespace skl {
// the dense handle's row-major copy, when LSQR would use it and it fits
static void handle_rows(skl_handle *h) {
    if (ctx().no_alloc || !lsqr_rows_wanted(h->n, h->d)) return;
    size_t fr = 0, tot = 0;
    SKL_HIP(hipMemGetInfo(&fr, &tot));
    const double need = 8.0 * (double)h->n * (double)lsqr_rows_ld(h->d);
    if (need * 1.25 + 256.0 * (1 << 20) > (double)fr) return;
    h->AT = (double *)h->take(sizeof(double) * (size_t)h->n * lsqr_rows_ld(h->d));
    lsqr_make_rows(h->n, h->d, h->A, h->n, h->AT, ctx().stream);
}

static int upload_dense(long n, long d, const double *A, long lda, skl_handle **out,
                        hipMemcpyKind kind) {
    return guarded([&] {
        SKL_REQUIRE(out && A && n > 0 && d > 0 && lda >= n, SKL_E_ARG, "invalid arguments");
        *out = nullptr;
        if (n < d) throw Error(SKL_E_SHAPE, "Matrix is not overdetermined.");
        ensure_init();
        hipStream_t s = ctx().stream;
        std::unique_ptr<skl_handle> h(new skl_handle());
        h->device = ctx().device;
        h->n = n;
        h->d = d;
        h->A = (double *)h->take(sizeof(double) * (size_t)n * d);
        SKL_HIP(hipMemcpy2DAsync(h->A, sizeof(double) * n, A, sizeof(double) * lda,
                                 sizeof(double) * n, d, kind, s));
        handle_rows(h.get());
        SKL_HIP(hipStreamSynchronize(s));
        *out = h.release();
        return SKL_OK;
    });
}

static int solve_on_handle(skl_handle *h, int method, const double *db, double *dx,
                           const skl_opts &o, skl_info &info) {
    SKL_REQUIRE(h->device == ctx().device, SKL_E_ARG,
                "the handle lives on device " + std::to_string(h->device) +
                    ", the calling thread uses device " + std::to_string(ctx().device));
    if (h->sparse) {
        SKL_REQUIRE(method == SKL_METHOD_SPARSE_SPQR, SKL_E_ARG,
                    "a CSC handle takes SKL_METHOD_SPARSE_SPQR");
        SparsePrep sp;
        sp.cp = h->cp;
        sp.ri = h->ri;
        sp.rp = h->rp;
        sp.cc = h->cc;
        sp.cv = h->cv;
        return sparse_impl(h->n, h->d, h->nnz, nullptr, nullptr, h->val, db, dx, o, info, &sp);
    }
    switch (method) {
    case SKL_METHOD_DENSE_CPQR:
        return dense_impl(h->n, h->d, h->A, h->n, db, dx, o, info, true, false, h->AT);
    case SKL_METHOD_DENSE_NOCPQR:
        return dense_impl(h->n, h->d, h->A, h->n, db, dx, o, info, false, false, h->AT);
    case SKL_METHOD_DENSE_BLENDENPIK:
        return dense_impl(h->n, h->d, h->A, h->n, db, dx, o, info, false, true, h->AT);
    default:
        throw Error(SKL_E_ARG, "a dense handle takes SKL_METHOD_DENSE_CPQR, _NOCPQR or _BLENDENPIK");
    }
}
} // namespace skl

int skl_upload_dense(long n, long d, const double *A, long lda, skl_handle **h) {
    return upload_dense(n, d, A, lda, h, hipMemcpyHostToDevice);
}

int skl_upload_dense_dev(long n, long d, const double *dA, long lda, skl_handle **h) {
    return upload_dense(n, d, dA, lda, h, hipMemcpyDeviceToDevice);
}

int skl_upload_csc(long n, long d, long nnz, const long *colptr, const long *rowind,
                   const double *val, skl_handle **out) {
    return guarded([&] {
        SKL_REQUIRE(out && colptr && n > 0 && d > 0 && nnz >= 0, SKL_E_ARG, "invalid arguments");
        *out = nullptr;
        if (n < d) throw Error(SKL_E_SHAPE, "Matrix is not overdetermined.");
        SKL_REQUIRE(nnz == 0 || (rowind && val), SKL_E_ARG, "invalid arguments");
        SKL_REQUIRE(nnz < 2147483647L && n < 2147483647L, SKL_E_UNSUPPORTED,
                    "nnz must fit in int32");
        SKL_REQUIRE(colptr[0] == 0 && colptr[d] == nnz, SKL_E_ARG, "colptr inconsistent with nnz");
        for (long j = 0; j < d; ++j)
            SKL_REQUIRE(colptr[j] <= colptr[j + 1], SKL_E_ARG, "colptr not monotone");
        for (long e = 0; e < nnz; ++e)
            SKL_REQUIRE(rowind[e] >= 0 && rowind[e] < n, SKL_E_ARG, "row index out of range");
        ensure_init();
        Context &C = ctx();
        hipStream_t s = C.stream;
        std::unique_ptr<skl_handle> h(new skl_handle());
        h->device = C.device;
        h->sparse = 1;
        h->n = n;
        h->d = d;
        h->nnz = nnz;
        std::vector<int32_t> cp32((size_t)d + 1), ri32((size_t)std::max<long>(nnz, 1));
        for (long j = 0; j <= d; ++j) cp32[(size_t)j] = (int32_t)colptr[j];
        for (long e = 0; e < nnz; ++e) ri32[(size_t)e] = (int32_t)rowind[e];
        h->cp = (int32_t *)h->take(sizeof(int32_t) * ((size_t)d + 1));
        h->ri = (int32_t *)h->take(sizeof(int32_t) * (size_t)std::max<long>(nnz, 1));
        h->val = (double *)h->take(sizeof(double) * (size_t)std::max<long>(nnz, 1));
        h->rp = (int32_t *)h->take(sizeof(int32_t) * ((size_t)n + 1));
        h->cc = (int32_t *)h->take(sizeof(int32_t) * (size_t)std::max<long>(nnz, 1));
        h->cv = (double *)h->take(sizeof(double) * (size_t)std::max<long>(nnz, 1));
        SKL_HIP(hipMemcpyAsync(h->cp, cp32.data(), sizeof(int32_t) * (d + 1),
                               hipMemcpyHostToDevice, s));
        if (nnz > 0) {
            SKL_HIP(hipMemcpyAsync(h->ri, ri32.data(), sizeof(int32_t) * nnz,
                                   hipMemcpyHostToDevice, s));
            SKL_HIP(hipMemcpyAsync(h->val, val, sizeof(double) * nnz, hipMemcpyHostToDevice, s));
        }
        Arena &ar = C.arena;
        ar.reset();
        ArenaTrim trim(ar);
        csc_to_csr(n, d, nnz, h->cp, h->ri, h->val, h->rp, h->cc, h->cv, ar);
        SKL_HIP(hipStreamSynchronize(s));
        *out = h.release();
        return SKL_OK;
    });
}

int skl_handle_query(const skl_handle *h, long *n, long *d, long *nnz, size_t *device_bytes) {
    if (!h) return SKL_E_ARG;
    if (n) *n = h->n;
    if (d) *d = h->d;
    if (nnz) *nnz = h->sparse ? h->nnz : -1;
    if (device_bytes) *device_bytes = h->bytes;
    return SKL_OK;
}

int skl_solve_handle_dev(skl_handle *h, int method, const double *db, double *dx,
                         const skl_opts *opts, skl_info *info) {
    skl_info tmp;
    return guarded([&] {
        SKL_REQUIRE(h && db && dx, SKL_E_ARG, "invalid arguments");
        skl_opts o = resolve(opts);
        ensure_init();
        return solve_on_handle(h, method, db, dx, o, info ? *info : tmp);
    });
}

int skl_solve_handle(skl_handle *h, int method, const double *b, double *x, const skl_opts *opts,
                     skl_info *info) {
    skl_info tmp;
    return guarded([&] {
        SKL_REQUIRE(h && b && x, SKL_E_ARG, "invalid arguments");
        skl_opts o = resolve(opts);
        ensure_init();
        hipStream_t s = ctx().stream;
        DevBuf db(sizeof(double) * h->n), dx(sizeof(double) * h->d);
        SKL_HIP(hipMemcpyAsync(db.p, b, sizeof(double) * h->n, hipMemcpyHostToDevice, s));
        int rc = solve_on_handle(h, method, db.as<double>(), dx.as<double>(), o,
                                 info ? *info : tmp);
        SKL_HIP(hipMemcpyAsync(x, dx.p, sizeof(double) * h->d, hipMemcpyDeviceToHost, s));
        SKL_HIP(hipStreamSynchronize(s));
        return rc;
    });
}

int skl_release(skl_handle *h) {
    if (!h) return SKL_E_ARG;
    return guarded([&] {
        delete h;
        return SKL_OK;
    });
}

// ---- LSQR launch-shape autotuner --------------------------------------------
int skl_autotune_lsqr(int sparse, long n, long d, long nnz, double *best_ms, double *default_ms) {
    return guarded([&] {
        ensure_init();
        ArenaTrim trim(ctx().arena);
        double def = 0.0;
        const double best = lsqr_autotune(sparse != 0, n, d, nnz, 32, &def);
        if (best_ms) *best_ms = best;
        if (default_ms) *default_ms = def;
        return SKL_OK;
    });
}

long skl_autotune_entries(void) { return lsqr_shape_entries(); }

void skl_autotune_clear(void) { lsqr_shape_clear(); }

// ---- component entry points ------------------------------------------------
int skl_hash_pattern(unsigned seed, int dense, long m, long n, long k, double *D, long *rows,
                     double *vals) {
    return guarded([&] {
        SKL_REQUIRE(m >= k && k >= 1 && n >= 0, SKL_E_ARG, "invalid pattern shape");
        Glibc g;
        g.seed(seed);
        Pattern P;
        if (dense) {
            gen_signs(g, n, P.D);
            if (D) std::memcpy(D, P.D.data(), sizeof(double) * n);
        }
        gen_hashing(g, m, n, k, P);
        for (long q = 0; q < n * k; ++q) {
            if (rows) rows[q] = P.rows[(size_t)q];
            if (vals) vals[q] = P.vals[(size_t)q];
        }
        return SKL_OK;
    });
}

int skl_subsample_pattern(unsigned seed, long m, long n, double *D, long *idx) {
    return guarded([&] {
        SKL_REQUIRE(m >= 0 && m <= n, SKL_E_ARG, "invalid sampling shape");
        Glibc g;
        g.seed(seed);
        Pattern P;
        gen_signs(g, n, P.D);
        if (D) std::memcpy(D, P.D.data(), sizeof(double) * n);
        Glibc g2;
        g2.seed(seed);
        gen_subsample(g2, n, m, P);
        if (idx)
            for (long i = 0; i < m; ++i) idx[i] = P.bsrc[(size_t)i];
        return SKL_OK;
    });
}

int skl_dht(long n, long ncols, const double *X, long ldx, double *Y, long ldy, int precision) {
    return guarded([&] {
        ensure_init();
        Arena &ar = ctx().arena;
        ar.reset();
        double *dX = ar.alloc_n<double>((size_t)n * ncols);
        SKL_HIP(hipMemcpy2D(dX, sizeof(double) * n, X, sizeof(double) * ldx, sizeof(double) * n,
                            ncols, hipMemcpyHostToDevice));
        if (precision == SKL_MIXED) {
            float *dY = ar.alloc_n<float>((size_t)n * ncols);
            dht_cols_from_double<float>(n, ncols, dX, n, dY, n, ar);
            sync();
            download_as_double(dY, n, ncols, n, Y, ldy);
        } else {
            double *dY = ar.alloc_n<double>((size_t)n * ncols);
            dht_cols_from_double<double>(n, ncols, dX, n, dY, n, ar);
            sync();
            download_as_double(dY, n, ncols, n, Y, ldy);
        }
        return SKL_OK;
    });
}

int skl_hash_apply(long m, long n, long k, const long *rows, const double *vals, double alpha,
                   long ncols, const double *X, long ldx, double *Y, long ldy, int precision) {
    return guarded([&] {
        ensure_init();
        Arena &ar = ctx().arena;
        ar.reset();
        Pattern P = pattern_from(m, n, k, rows, vals);
        DevPattern dp = upload_pattern(P, alpha, false, false, precision == SKL_MIXED, ar);
        double *dX = ar.alloc_n<double>((size_t)n * ncols);
        SKL_HIP(hipMemcpy2D(dX, sizeof(double) * n, X, sizeof(double) * ldx, sizeof(double) * n,
                            ncols, hipMemcpyHostToDevice));
        if (precision == SKL_MIXED) {
            float *fX = ar.alloc_n<float>((size_t)n * ncols);
            copy2d<double, float>(n, ncols, dX, n, fX, n);
            float *dY = ar.alloc_n<float>((size_t)m * ncols);
            hash_gather<float>(m, ncols, dp.bptr, dp.bsrc, dp.coef_f, fX, n, dY, m);
            sync();
            download_as_double(dY, m, ncols, m, Y, ldy);
        } else {
            double *dY = ar.alloc_n<double>((size_t)m * ncols);
            hash_gather<double>(m, ncols, dp.bptr, dp.bsrc, dp.coef_d, dX, n, dY, m);
            sync();
            download_as_double(dY, m, ncols, m, Y, ldy);
        }
        return SKL_OK;
    });
}

int skl_dense_sketch(long n, long d, const double *A, long lda, const double *b, long m, long k,
                     const double *D, const long *rows, const double *vals, double *SA, long ldsa,
                     int precision) {
    return guarded([&] {
        ensure_init();
        Arena &ar = ctx().arena;
        ar.reset();
        Pattern P = pattern_from(m, n, k, rows, vals);
        P.D.assign(D, D + n);
        bool mixed = precision == SKL_MIXED;
        DevPattern dp = upload_pattern(P, 1 / std::sqrt((double)k), true, false, mixed, ar);
        double *dA = ar.alloc_n<double>((size_t)n * d);
        SKL_HIP(hipMemcpy2D(dA, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, d,
                            hipMemcpyHostToDevice));
        double *db = nullptr;
        if (b) {
            db = ar.alloc_n<double>((size_t)n);
            SKL_HIP(hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice));
        }
        long nc = d + (b ? 1 : 0);
        double scale = 1 / std::sqrt((double)n);
        if (mixed) {
            float *W = ar.alloc_n<float>((size_t)m * nc);
            dht_sketch<float>(n, d, dA, n, dp.D, scale, m, dp.bptr, dp.bsrc, dp.coef_f, W, m, ar);
            if (b) dht_sketch<float>(n, 1, db, n, dp.D, scale, m, dp.bptr, dp.bsrc, dp.coef_f, W + d * m, m, ar);
            sync();
            download_as_double(W, m, nc, m, SA, ldsa);
        } else {
            double *W = ar.alloc_n<double>((size_t)m * nc);
            dht_sketch<double>(n, d, dA, n, dp.D, scale, m, dp.bptr, dp.bsrc, dp.coef_d, W, m, ar);
            if (b) dht_sketch<double>(n, 1, db, n, dp.D, scale, m, dp.bptr, dp.bsrc, dp.coef_d, W + d * m, m, ar);
            sync();
            download_as_double(W, m, nc, m, SA, ldsa);
        }
        return SKL_OK;
    });
}

int skl_sparse_sketch(long m, long n, long d, long k, const long *rows, const double *vals,
                      const long *colptr, const long *rowind, const double *val, const double *b,
                      double *SA, long ldsa, double *bs) {
    return guarded([&] {
        ensure_init();
        Arena &ar = ctx().arena;
        ar.reset();
        Pattern P = pattern_from(m, n, k, rows, vals);
        DevPattern dp = upload_pattern(P, 1.0, false, true, false, ar);
        long nnz = colptr[d];
        int32_t *cp = ar.alloc_n<int32_t>((size_t)d + 1);
        int32_t *ri = ar.alloc_n<int32_t>((size_t)std::max<long>(nnz, 1));
        double *dv = ar.alloc_n<double>((size_t)std::max<long>(nnz, 1));
        std::vector<int32_t> hcp(colptr, colptr + d + 1), hri(rowind, rowind + nnz);
        SKL_HIP(hipMemcpy(cp, hcp.data(), sizeof(int32_t) * (d + 1), hipMemcpyHostToDevice));
        if (nnz > 0) {
            SKL_HIP(hipMemcpy(ri, hri.data(), sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
            SKL_HIP(hipMemcpy(dv, val, sizeof(double) * nnz, hipMemcpyHostToDevice));
        }
        double *W = ar.alloc_n<double>((size_t)m * d);
        sparse_sketch_t<double>(m, d, k, dp.rows, dp.vals, cp, ri, dv, nnz, W, m, ar);
        if (b && bs) {
            double *db = ar.alloc_n<double>((size_t)n);
            double *dbs = ar.alloc_n<double>((size_t)m);
            SKL_HIP(hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice));
            hash_gather<double>(m, 1, dp.bptr, dp.bsrc, dp.coef_d, db, n, dbs, m);
            sync();
            SKL_HIP(hipMemcpy(bs, dbs, sizeof(double) * m, hipMemcpyDeviceToHost));
        }
        sync();
        download_as_double(W, m, d, m, SA, ldsa);
        return SKL_OK;
    });
}

int skl_qr(long m, long n, long nfac, double *A, long lda, double *tau, int precision) {
    return guarded([&] {
        ensure_init();
        SKL_REQUIRE(nfac <= n && m > 0 && lda >= m, SKL_E_ARG, "invalid QR shape");
        return precision == SKL_MIXED ? qr_component<float>(m, n, nfac, A, lda, tau, nullptr, false)
                                      : qr_component<double>(m, n, nfac, A, lda, tau, nullptr, false);
    });
}

int skl_cpqr(long m, long n, long nfac, double *A, long lda, long *jpvt, double *tau,
             int precision) {
    return guarded([&] {
        ensure_init();
        SKL_REQUIRE(nfac <= n && m > 0 && lda >= m, SKL_E_ARG, "invalid QR shape");
        return precision == SKL_MIXED ? qr_component<float>(m, n, nfac, A, lda, tau, jpvt, true)
                                      : qr_component<double>(m, n, nfac, A, lda, tau, jpvt, true);
    });
}

int skl_trinv(long r, const double *R, long ldr, double perturb, double *Minv, long ldm) {
    return guarded([&] {
        ensure_init();
        Arena &ar = ctx().arena;
        ar.reset();
        double *dR = ar.alloc_n<double>((size_t)r * r);
        double *dM = ar.alloc_n<double>((size_t)r * r);
        SKL_HIP(hipMemcpy2D(dR, sizeof(double) * r, R, sizeof(double) * ldr, sizeof(double) * r, r,
                            hipMemcpyHostToDevice));
        trinv_upper(r, dR, r, perturb, dM, r, ar);
        sync();
        SKL_HIP(hipMemcpy2D(Minv, sizeof(double) * ldm, dM, sizeof(double) * r, sizeof(double) * r,
                            r, hipMemcpyDeviceToHost));
        return SKL_OK;
    });
}

int skl_lsqr_dense(long n, long d, const double *A, long lda, long r, const long *E,
                   const double *M, long ldm, const double *b, double tol, long max_it, double *y,
                   int *flag, long *it, skl_lsqr_step *hist) {
    return guarded([&] {
        ensure_init();
        Arena &ar = ctx().arena;
        ar.reset();
        double *dA = ar.alloc_n<double>((size_t)n * d);
        double *dM = ar.alloc_n<double>((size_t)r * r);
        double *db = ar.alloc_n<double>((size_t)n);
        double *dy = ar.alloc_n<double>((size_t)r);
        SKL_HIP(hipMemcpy2D(dA, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, d,
                            hipMemcpyHostToDevice));
        // M: upper triangle, zeros below (the triangular GEMVs read the square)
        std::vector<double> hm((size_t)r * r, 0.0);
        for (long j = 0; j < r; ++j)
            for (long i = 0; i <= j; ++i) hm[(size_t)(i + j * r)] = M[i + j * ldm];
        SKL_HIP(hipMemcpy(dM, hm.data(), sizeof(double) * r * r, hipMemcpyHostToDevice));
        SKL_HIP(hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice));
        int32_t *dE = nullptr;
        if (E) {
            std::vector<int32_t> he(E, E + r);
            dE = ar.alloc_n<int32_t>((size_t)r);
            SKL_HIP(hipMemcpy(dE, he.data(), sizeof(int32_t) * r, hipMemcpyHostToDevice));
        }
        LsqrProblem p;
        p.n = n;
        p.d = d;
        p.r = r;
        p.A = dA;
        p.lda = n;
        p.E = dE;
        p.M = dM;
        p.ldm = r;
        LsqrResult res = lsqr_solve(p, db, tol, max_it, dy, true, 16, hist, hist ? max_it : 0, ar);
        SKL_HIP(hipMemcpy(y, dy, sizeof(double) * r, hipMemcpyDeviceToHost));
        if (flag) *flag = res.flag;
        if (it) *it = res.it;
        return SKL_OK;
    });
}

} // extern "C"
