/*
 * skl_oracle.c -- CPU restatement of the Ski-LLS hot path (FP64, plain C).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Built with -ffp-contract=off so that every product and sum is rounded the
 * way the reference's scalar loops round them (the reference's x86-64 build
 * has no FMA in the baseline ISA for the C sources).
 */
#include "skl_oracle.h"

/* ORC_PLANT (default 0) compiles a deliberate mistake into the HQRRP Y
 * downdate (1: wrong sign of the (G V) T U11^T term, 2: G not updated by the
 * block reflector, 3: the last row of R12 dropped).  Only the tests build
 * these variants, to show that the per-block pivot check in
 * tests/test_oracle.py fails on each of them. */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_sec(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ======================================================================== */
/* libc rand(): glibc srandom_r / random_r, TYPE_3 (degree 31, separation 3). */
/* ======================================================================== */
void orc_srand(orc_rng *g, unsigned seed) {
    int32_t word; /* glibc keeps the LCG word in an int32_t */
    int i;
    if (seed == 0) seed = 1;
    g->state[0] = (int32_t)seed;
    word = (int32_t)seed;
    for (i = 1; i < 31; ++i) {
        long hi = word / 127773;
        long lo = word % 127773;
        word = (int32_t)(16807 * lo - 2836 * hi);
        if (word < 0) word += 2147483647;
        g->state[i] = word;
    }
    g->f = 3;
    g->r = 0;
    for (i = 0; i < 310; ++i) (void)orc_rand(g);
    g->gauss_toggle = 0;
    g->gauss_b2 = 0.0;
}

int orc_rand(orc_rng *g) {
    uint32_t val = (uint32_t)g->state[g->f] + (uint32_t)g->state[g->r];
    g->state[g->f] = (int32_t)val;
    if (++g->f >= 31) g->f = 0;
    if (++g->r >= 31) g->r = 0;
    return (int)(val >> 1);
}

/* dense_helpers.cpp:7-16 */
void orc_gen_signs(orc_rng *g, long n, double scale, double *D) {
    for (long i = 0; i < n; ++i) D[i] = scale * (double)((orc_rand(g) % 2) * 2 - 1);
}

/* sparse_helpers.cpp:90-116, with select_k_from_n (:74-87) acting on an index
 * array that persists across columns. */
void orc_gen_hashing(orc_rng *g, long m, long n, long k, long *rows, long *colptr,
                     double *vals) {
    long *arr = (long *)malloc((size_t)m * sizeof(long));
    for (long i = 0; i < m; ++i) arr[i] = i;
    for (long i = 0; i < n; ++i) {
        for (long t = 0; t < k; ++t) {
            long j = ((long)orc_rand(g) % (m - t)) + t;
            long tmp = arr[j];
            arr[j] = arr[t];
            arr[t] = tmp;
        }
        for (long t = 0; t < k; ++t) rows[k * i + t] = arr[t];
    }
    free(arr);
    if (colptr)
        for (long i = 0; i < n + 1; ++i) colptr[i] = k * i;
    double scaling = sqrt((double)k);
    for (long i = 0; i < n; ++i)
        for (long t = 0; t < k; ++t)
            vals[i * k + t] = 1 / scaling * (double)((orc_rand(g) % 2) * 2 - 1);
}

void orc_select_k_from_n(orc_rng *g, long n, long k, long *idx) {
    long *arr = (long *)malloc((size_t)(n > 0 ? n : 1) * sizeof(long));
    for (long i = 0; i < n; ++i) arr[i] = i; /* gen_0_to_n, sparse_helpers.cpp:66-71 */
    for (long i = 0; i < k; ++i) {
        long j = ((long)orc_rand(g) % (n - i)) + i;
        long tmp = arr[j];
        arr[j] = arr[i];
        arr[i] = tmp;
    }
    for (long i = 0; i < k; ++i) idx[i] = arr[i];
    free(arr);
}

/* NoFLA_Normal_random_number (skint.c:603-624): Marsaglia polar method on
 * rand()/RAND_MAX with an alternate-call cache. */
static double orc_normal(orc_rng *g) {
    double c1, c2, a, factor;
    if (g->gauss_toggle == 1) {
        g->gauss_toggle = 0;
        return 0.0 + 1.0 * g->gauss_b2;
    }
    do {
        c1 = -1.0 + 2.0 * ((double)orc_rand(g) / 2147483647);
        c2 = -1.0 + 2.0 * ((double)orc_rand(g) / 2147483647);
        a = c1 * c1 + c2 * c2;
    } while ((a == 0) || (a >= 1));
    factor = sqrt((-2 * log(a)) / a);
    g->gauss_b2 = c2 * factor;
    g->gauss_toggle = 1;
    return 0.0 + 1.0 * (c1 * factor);
}

/* skint.c:585-600: filled column by column */
void orc_gaussian_matrix(orc_rng *g, long rows, long cols, double *G, long ldg) {
    for (long j = 0; j < cols; ++j)
        for (long i = 0; i < rows; ++i) G[i + j * ldg] = orc_normal(g);
}

/* ======================================================================== */
/* FFT: mixed-radix Stockham (radices 4,2,3,5,7,...,31) or Bluestein.        */
/* ======================================================================== */
static int orc_factor(long n, int *f) {
    int nf = 0;
    while (n % 4 == 0) { f[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { f[nf++] = 2; n /= 2; }
    for (long p = 3; p <= 31; p += 2)
        while (n % p == 0) { f[nf++] = (int)p; n /= p; }
    return n == 1 ? nf : -1;
}

/* twiddle table w[t] = exp(-2 pi i t / n), t < n (interleaved) */
static double *orc_twiddles(long n) {
    double *w = (double *)malloc((size_t)2 * n * sizeof(double));
    for (long t = 0; t < n; ++t) {
        double ang = -2.0 * M_PI * (double)t / (double)n;
        w[2 * t] = cos(ang);
        w[2 * t + 1] = sin(ang);
    }
    return w;
}

/* Stockham DIF: stage with radix r on current length len (stride s):
 *   y[q + s(r p + k)] = W_len^{pk} * sum_j x[q + s(p + j m)] W_r^{jk},  m = len/r */
static void orc_fft_smooth(long n, double *z, const int *f, int nf) {
    double *w = orc_twiddles(n);
    double *buf = (double *)malloc((size_t)2 * n * sizeof(double));
    double *x = z, *y = buf;
    long len = n, s = 1;
    for (int st = 0; st < nf; ++st) {
        int r = f[st];
        long m = len / r;
        long tstep = n / len; /* W_len^e = W_n^{e * tstep} */
        long rstep = n / r;   /* W_r^e   = W_n^{e * rstep} */
        for (long p = 0; p < m; ++p) {
            for (long q = 0; q < s; ++q) {
                double ar[32], ai[32];
                for (int j = 0; j < r; ++j) {
                    long idx = q + s * (p + j * m);
                    ar[j] = x[2 * idx];
                    ai[j] = x[2 * idx + 1];
                }
                for (int kk = 0; kk < r; ++kk) {
                    double sr = 0.0, si = 0.0;
                    for (int j = 0; j < r; ++j) {
                        long e = ((long)j * kk) % r;
                        double cr = w[2 * (e * rstep)], ci = w[2 * (e * rstep) + 1];
                        sr += ar[j] * cr - ai[j] * ci;
                        si += ar[j] * ci + ai[j] * cr;
                    }
                    long e2 = (p * kk) % len;
                    double tr = w[2 * (e2 * tstep)], ti = w[2 * (e2 * tstep) + 1];
                    long o = q + s * (r * p + kk);
                    y[2 * o] = sr * tr - si * ti;
                    y[2 * o + 1] = sr * ti + si * tr;
                }
            }
        }
        double *t = x;
        x = y;
        y = t;
        len = m;
        s *= r;
    }
    if (x != z) memcpy(z, x, (size_t)2 * n * sizeof(double));
    free(buf);
    free(w);
}

static void orc_fft_pow2(long n, double *z) {
    int f[64];
    int nf = orc_factor(n, f);
    orc_fft_smooth(n, z, f, nf);
}

/* Bluestein: X_k = w_k sum_j (x_j w_j) conj(w_{k-j}),  w_j = exp(-i pi j^2/n) */
static void orc_fft_bluestein(long n, double *z) {
    long L = 1;
    while (L < 2 * n - 1) L <<= 1;
    double *wr = (double *)malloc((size_t)2 * n * sizeof(double));
    for (long j = 0; j < n; ++j) {
        long e = (long)(((__int128)j * j) % (2 * n));
        double ang = -M_PI * (double)e / (double)n;
        wr[2 * j] = cos(ang);
        wr[2 * j + 1] = sin(ang);
    }
    double *a = (double *)calloc((size_t)2 * L, sizeof(double));
    double *bb = (double *)calloc((size_t)2 * L, sizeof(double));
    for (long j = 0; j < n; ++j) {
        double xr = z[2 * j], xi = z[2 * j + 1];
        a[2 * j] = xr * wr[2 * j] - xi * wr[2 * j + 1];
        a[2 * j + 1] = xr * wr[2 * j + 1] + xi * wr[2 * j];
        bb[2 * j] = wr[2 * j];
        bb[2 * j + 1] = -wr[2 * j + 1];
        if (j > 0) {
            bb[2 * (L - j)] = wr[2 * j];
            bb[2 * (L - j) + 1] = -wr[2 * j + 1];
        }
    }
    orc_fft_pow2(L, a);
    orc_fft_pow2(L, bb);
    for (long t = 0; t < L; ++t) { /* c = conj(A .* B) for the inverse via forward FFT */
        double pr = a[2 * t] * bb[2 * t] - a[2 * t + 1] * bb[2 * t + 1];
        double pi = a[2 * t] * bb[2 * t + 1] + a[2 * t + 1] * bb[2 * t];
        a[2 * t] = pr;
        a[2 * t + 1] = -pi;
    }
    orc_fft_pow2(L, a);
    for (long k = 0; k < n; ++k) {
        double cr = a[2 * k] / (double)L, ci = -a[2 * k + 1] / (double)L;
        z[2 * k] = cr * wr[2 * k] - ci * wr[2 * k + 1];
        z[2 * k + 1] = cr * wr[2 * k + 1] + ci * wr[2 * k];
    }
    free(a);
    free(bb);
    free(wr);
}

void orc_fft(long n, double *z) {
    int f[64];
    if (n <= 1) return;
    int nf = orc_factor(n, f);
    if (nf >= 0)
        orc_fft_smooth(n, z, f, nf);
    else
        orc_fft_bluestein(n, z);
}

/* DHT via one complex FFT per pair of real columns: with z = x1 + i x2,
 * Z_k = a + ib, Z_{n-k} = c + id:  H1_k = (a + c - b + d)/2,
 * H2_k = (b + d + a - c)/2.  (fftw_r2r.cpp:17-54 with kind FFTW_DHT) */
void orc_dht_cols(long n, long ncols, double *Y, long ldy) {
    double *z = (double *)malloc((size_t)2 * n * sizeof(double));
    for (long c = 0; c < ncols; c += 2) {
        double *y1 = Y + c * ldy;
        double *y2 = (c + 1 < ncols) ? Y + (c + 1) * ldy : NULL;
        for (long j = 0; j < n; ++j) {
            z[2 * j] = y1[j];
            z[2 * j + 1] = y2 ? y2[j] : 0.0;
        }
        orc_fft(n, z);
        for (long kk = 0; kk < n; ++kk) {
            long nk = kk == 0 ? 0 : n - kk;
            double a = z[2 * kk], b = z[2 * kk + 1], cc = z[2 * nk], dd = z[2 * nk + 1];
            y1[kk] = 0.5 * ((a + cc) - (b - dd));
            if (y2) y2[kk] = 0.5 * ((b + dd) + (a - cc));
        }
    }
    free(z);
}

/* fast_unitary_transform.cpp:88-100 (wisdom=0: padded length == n):
 * y = (scale * D[j]) * A(j, c), scale = 1/sqrt(n); then the unnormalised DHT. */
void orc_fut(long n, long d, const double *A, long lda, const double *D, double *Y,
             long ldy) {
    double scale = 1 / sqrt((double)n);
    for (long c = 0; c < d; ++c)
        for (long j = 0; j < n; ++j) Y[j + c * ldy] = scale * D[j] * A[j + c * lda];
    orc_dht_cols(n, d, Y, ldy);
}

/* ======================================================================== */
/* Hashing sketch                                                            */
/* ======================================================================== */
void orc_hash_apply(long m, long n, long k, const long *rows, const double *vals,
                    double alpha, long ncols, const double *X, long ldx, double *Y,
                    long ldy) {
    for (long c = 0; c < ncols; ++c) {
        double *y = Y + c * ldy;
        const double *x = X + c * ldx;
        for (long i = 0; i < m; ++i) y[i] = 0.0;
        for (long j = 0; j < n; ++j)
            for (long idx = k * j; idx < k * j + k; ++idx)
                y[rows[idx]] += alpha * vals[idx] * x[j];
    }
}

void orc_sketch_sparse(long m, long n, long k, const long *Srows, const double *Svals,
                       long d, const long *Ap, const long *Ai, const double *Ax,
                       double *SA, long ldsa) {
    long *mark = (long *)malloc((size_t)m * sizeof(long));
    (void)n;
    for (long i = 0; i < m; ++i) mark[i] = -1;
    for (long c = 0; c < d; ++c) {
        double *x = SA + c * ldsa;
        for (long i = 0; i < m; ++i) x[i] = 0.0;
        for (long p = Ap[c]; p < Ap[c + 1]; ++p) {
            long i = Ai[p];
            double beta = Ax[p];
            for (long q = k * i; q < k * i + k; ++q) {
                long h = Srows[q];
                if (mark[h] < c) {
                    mark[h] = c;
                    x[h] = beta * Svals[q];
                } else {
                    x[h] += beta * Svals[q];
                }
            }
        }
    }
    free(mark);
}

void orc_sketch_vec(long m, long n, long k, const long *Srows, const double *Svals,
                    const double *b, double *y) {
    for (long i = 0; i < m; ++i) y[i] = 0.0;
    for (long j = 0; j < n; ++j)
        for (long q = k * j; q < k * j + k; ++q) y[Srows[q]] += Svals[q] * b[j];
}

/* ======================================================================== */
/* Householder QR family                                                     */
/* ======================================================================== */
static double orc_nrm2(long n, const double *x, long incx) {
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += x[i * incx] * x[i * incx];
    return sqrt(s);
}

static double orc_dlapy2(double x, double y) {
    double xa = fabs(x), ya = fabs(y);
    double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
    if (z == 0.0) return w;
    return w * sqrt(1.0 + (z / w) * (z / w));
}

/* LAPACK dlarfg */
void orc_dlarfg(long n, double *alpha, double *x, long incx, double *tau) {
    if (n <= 1) {
        *tau = 0.0;
        return;
    }
    double xnorm = orc_nrm2(n - 1, x, incx);
    if (xnorm == 0.0) {
        *tau = 0.0;
        return;
    }
    double beta = -copysign(orc_dlapy2(*alpha, xnorm), *alpha);
    double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
    int knt = 0;
    if (fabs(beta) < safmin) {
        double rsafmn = 1.0 / safmin;
        do {
            ++knt;
            for (long i = 0; i < n - 1; ++i) x[i * incx] *= rsafmn;
            beta *= rsafmn;
            *alpha *= rsafmn;
        } while (fabs(beta) < safmin && knt < 20);
        xnorm = orc_nrm2(n - 1, x, incx);
        beta = -copysign(orc_dlapy2(*alpha, xnorm), *alpha);
    }
    *tau = (beta - *alpha) / beta;
    double sc = 1.0 / (*alpha - beta);
    for (long i = 0; i < n - 1; ++i) x[i * incx] *= sc;
    for (int j = 0; j < knt; ++j) beta *= safmin;
    *alpha = beta;
}

/* H = I - tau v v^T applied from the left to C (mrows x ncols), v[0] = 1
 * implied by the caller having stored 1 there (dlarf 'L': w = C^T v, C -= tau v w^T) */
static void orc_dlarf_left(long mrows, long ncols, const double *v, double tau,
                           double *C, long ldc) {
    if (tau == 0.0) return;
    for (long j = 0; j < ncols; ++j) {
        double *cj = C + j * ldc;
        double w = 0.0;
        for (long i = 0; i < mrows; ++i) w += cj[i] * v[i];
        double t = -tau * w;
        for (long i = 0; i < mrows; ++i) cj[i] = cj[i] + v[i] * t;
    }
}

/* dgeqr2: A = QR in place, reflectors below the diagonal */
void orc_geqrf(long m, long n, double *A, long lda, double *tau) {
    long kmax = m < n ? m : n;
    for (long i = 0; i < kmax; ++i) {
        double *aii = A + i + i * lda;
        orc_dlarfg(m - i, aii, aii + 1, 1, &tau[i]);
        if (i < n - 1) {
            double diag = *aii;
            *aii = 1.0;
            orc_dlarf_left(m - i, n - i - 1, aii, tau[i], A + i + (i + 1) * lda, lda);
            *aii = diag;
        }
    }
}

static void orc_swap_cols(long m, double *A, long lda, long a, long b) {
    if (a == b) return;
    double *x = A + a * lda, *y = A + b * lda;
    for (long i = 0; i < m; ++i) {
        double t = x[i];
        x[i] = y[i];
        y[i] = t;
    }
}

static long orc_idamax(long n, const double *x) {
    long best = 0;
    double bv = fabs(x[0]);
    for (long i = 1; i < n; ++i)
        if (fabs(x[i]) > bv) {
            bv = fabs(x[i]);
            best = i;
        }
    return best;
}

/* dgeqp3 / dlaqp2 (LAPACK 3.9): greedy column pivoting with LAWN 176 norm
 * downdating; jpvt is 1-based on output. */
void orc_geqp3(long m, long n, double *A, long lda, long *jpvt, double *tau) {
    long kmax = m < n ? m : n;
    double *vn1 = (double *)malloc((size_t)n * sizeof(double));
    double *vn2 = (double *)malloc((size_t)n * sizeof(double));
    double tol3z = sqrt(DBL_EPSILON * 0.5);
    for (long j = 0; j < n; ++j) {
        jpvt[j] = j + 1;
        vn1[j] = orc_nrm2(m, A + j * lda, 1);
        vn2[j] = vn1[j];
    }
    for (long i = 0; i < kmax; ++i) {
        long pvt = i + orc_idamax(n - i, vn1 + i);
        if (pvt != i) {
            orc_swap_cols(m, A, lda, pvt, i);
            long t = jpvt[pvt];
            jpvt[pvt] = jpvt[i];
            jpvt[i] = t;
            vn1[pvt] = vn1[i];
            vn2[pvt] = vn2[i];
        }
        double *aii = A + i + i * lda;
        if (i < m - 1)
            orc_dlarfg(m - i, aii, aii + 1, 1, &tau[i]);
        else
            orc_dlarfg(1, aii, aii, 1, &tau[i]);
        if (i < n - 1) {
            double diag = *aii;
            *aii = 1.0;
            orc_dlarf_left(m - i, n - i - 1, aii, tau[i], A + i + (i + 1) * lda, lda);
            *aii = diag;
        }
        for (long j = i + 1; j < n; ++j) {
            if (vn1[j] != 0.0) {
                double r = fabs(A[i + j * lda]) / vn1[j];
                double temp = 1.0 - r * r;
                if (temp < 0.0) temp = 0.0;
                double q = vn1[j] / vn2[j];
                double temp2 = temp * q * q;
                if (temp2 <= tol3z) {
                    if (i < m - 1) {
                        vn1[j] = orc_nrm2(m - i - 1, A + (i + 1) + j * lda, 1);
                        vn2[j] = vn1[j];
                    } else {
                        vn1[j] = 0.0;
                        vn2[j] = 0.0;
                    }
                } else {
                    vn1[j] = vn1[j] * sqrt(temp);
                }
            }
        }
    }
    free(vn1);
    free(vn2);
}

/* ---- HQRRP pieces ------------------------------------------------------- */

/* NoFLA_QRPmod_WY_unb_var4 (skint.c:784-896): unblocked QR with optional
 * pivoting of A (ma x na); B (mb rows) and C (mc rows) are pivoted alongside;
 * T (nstages x nstages, ldt) is built by dlarft when Tout != NULL. */
static void orc_qrpmod(int pivoting, long nstages, long ma, long na, double *A, long lda,
                       long *p, double *t, long mb, double *B, long ldb, long mc,
                       double *C, long ldc, double *T, long ldt) {
    long mn = ma < na ? ma : na;
    if (nstages < 0) nstages = mn;
    double *d = (double *)malloc((size_t)(na > 0 ? na : 1) * sizeof(double));
    double *e = (double *)malloc((size_t)(na > 0 ? na : 1) * sizeof(double));
    double tol3z = sqrt(DBL_EPSILON * 0.5);
    if (pivoting)
        for (long j = 0; j < na; ++j) {
            d[j] = orc_nrm2(ma, A + j * lda, 1);
            e[j] = d[j];
        }
    for (long j = 0; j < nstages; ++j) {
        if (pivoting) {
            long jm = orc_idamax(na - j, d + j); /* relative to j */
            if (jm != 0) {
                orc_swap_cols(ma, A + j * lda, lda, 0, jm);
                if (mb > 0) orc_swap_cols(mb, B + j * ldb, ldb, 0, jm);
                if (mc > 0) orc_swap_cols(mc, C + j * ldc, ldc, 0, jm);
                long iv = p[j + jm];
                p[j + jm] = p[j];
                p[j] = iv;
                d[j + jm] = d[j];
                e[j + jm] = e[j];
            }
        }
        double *ajj = A + j + j * lda;
        long nh = ma - j;
        orc_dlarfg(nh, ajj, A + (ma - 1 < j + 1 ? ma - 1 : j + 1) + j * lda, 1, &t[j]);
        double diag = *ajj;
        *ajj = 1.0;
        orc_dlarf_left(nh, na - j - 1, ajj, t[j], A + j + (j + 1) * lda, lda);
        *ajj = diag;
        if (pivoting) {
            /* NoFLA_QRP_downdate_partial_norms (skint.c:920-990) */
            long mrest = ma - j - 1;
            for (long jj = j + 1; jj < na; ++jj) {
                if (d[jj] != 0.0) {
                    double temp = fabs(A[j + jj * lda]) / d[jj];
                    temp = (1.0 + temp) * (1 - temp);
                    if (temp < 0.0) temp = 0.0;
                    double temp5 = d[jj] / e[jj];
                    double temp2 = temp * temp5 * temp5;
                    if (temp2 <= tol3z) {
                        if (mrest > 0) {
                            d[jj] = orc_nrm2(mrest, A + (j + 1) + jj * lda, 1);
                            e[jj] = d[jj];
                        } else {
                            d[jj] = 0.0;
                            e[jj] = 0.0;
                        }
                    } else {
                        d[jj] = d[jj] * sqrt(temp);
                    }
                }
            }
        }
    }
    if (T) {
        /* dlarft('Forward','Columnwise'): T(0:i,i) = -tau_i T(0:i,0:i) V(:,0:i)^T v_i */
        for (long i = 0; i < nstages; ++i) {
            double ti = t[i];
            for (long r = 0; r < i; ++r) {
                /* V(:,r)^T v_i, v_i = e_i + A(i+1:ma, i), V(:,r) unit at r */
                double s = A[i + r * lda]; /* V(i, r) * 1 */
                for (long q = i + 1; q < ma; ++q) s += A[q + r * lda] * A[q + i * lda];
                T[r + i * ldt] = -ti * s;
            }
            /* T(0:i,i) = T(0:i,0:i) * T(0:i,i) (upper triangular multiply) */
            for (long r = 0; r < i; ++r) {
                double s = 0.0;
                for (long q = r; q < i; ++q) s += T[r + q * ldt] * T[q + i * ldt];
                T[r + i * ldt] = s;
            }
            T[i + i * ldt] = ti;
        }
    }
    free(d);
    free(e);
}

/* V unit lower (mv x kb, ldv), T upper (kb x kb): C := (I - V T V^T)^T C  (left)
 * or C := C (I - V T V^T) (right, C is mc x mv). dlarfb semantics. */
static void orc_larfb_left_t(long mv, long kb, const double *V, long ldv, const double *T,
                             long ldt, long nc, double *C, long ldc) {
    /* W = V^T C (kb x nc); W = T^T W; C -= V W */
    double *W = (double *)malloc((size_t)(kb * nc > 0 ? kb * nc : 1) * sizeof(double));
    for (long j = 0; j < nc; ++j)
        for (long r = 0; r < kb; ++r) {
            double s = C[r + j * ldc];
            for (long q = r + 1; q < mv; ++q) s += V[q + r * ldv] * C[q + j * ldc];
            W[r + j * kb] = s;
        }
    for (long j = 0; j < nc; ++j)
        for (long r = kb - 1; r >= 0; --r) {
            double s = 0.0;
            for (long q = 0; q <= r; ++q) s += T[q + r * ldt] * W[q + j * kb];
            W[r + j * kb] = s;
        }
    for (long j = 0; j < nc; ++j)
        for (long q = 0; q < mv; ++q) {
            double s = 0.0;
            long top = q < kb - 1 ? q : kb - 1;
            for (long r = 0; r <= top; ++r) s += (q == r ? 1.0 : V[q + r * ldv]) * W[r + j * kb];
            C[q + j * ldc] -= s;
        }
    free(W);
}

static void orc_larfb_right_n(long mv, long kb, const double *V, long ldv, const double *T,
                              long ldt, long mc, double *C, long ldc) {
    /* W = C V (mc x kb); W = W T; C -= W V^T */
    double *W = (double *)malloc((size_t)(mc * kb > 0 ? mc * kb : 1) * sizeof(double));
    for (long r = 0; r < kb; ++r)
        for (long i = 0; i < mc; ++i) {
            double s = C[i + r * ldc];
            for (long q = r + 1; q < mv; ++q) s += C[i + q * ldc] * V[q + r * ldv];
            W[i + r * mc] = s;
        }
    for (long i = 0; i < mc; ++i)
        for (long r = kb - 1; r >= 0; --r) {
            double s = 0.0;
            for (long q = 0; q <= r; ++q) s += W[i + q * mc] * T[q + r * ldt];
            W[i + r * mc] = s;
        }
    for (long q = 0; q < mv; ++q)
        for (long i = 0; i < mc; ++i) {
            double s = 0.0;
            long top = q < kb - 1 ? q : kb - 1;
            for (long r = 0; r <= top; ++r) s += W[i + r * mc] * (q == r ? 1.0 : V[q + r * ldv]);
            C[i + q * ldc] -= s;
        }
    free(W);
}

/* NoFLA_Downdate_Y (skint.c:627-718):
 * Y2 -= (G1 - (G1 U11 + G2 U21) T U11^T) R12;  then G(:, j:m) := G(:, j:m) Q */
static void orc_downdate_y(long b, const double *U, long ldu, long mu, const double *R12,
                           long ldr, long nr, const double *T, long ldt, double *Y2,
                           long ldy, long my, double *G1, long ldg, long mg) {
    /* U = [U11; U21] is mu x b (unit lower), G1 points at G(:, j), G2 = G1 + b*ldg */
    double *B = (double *)malloc((size_t)my * b * sizeof(double));
    const double *G2 = G1 + b * ldg;
    long mu21 = mu - b;
    for (long r = 0; r < b; ++r)
        for (long i = 0; i < my; ++i) {
            /* (G1 U11)(i, r) = G1(i, r) + sum_{q>r, q<b} G1(i,q) U11(q,r) */
            double s = G1[i + r * ldg];
            for (long q = r + 1; q < b; ++q) s += G1[i + q * ldg] * U[q + r * ldu];
            for (long q = 0; q < mu21; ++q) s += G2[i + q * ldg] * U[b + q + r * ldu];
            B[i + r * my] = s;
        }
    /* B = B T (upper) */
    for (long i = 0; i < my; ++i)
        for (long r = b - 1; r >= 0; --r) {
            double s = 0.0;
            for (long q = 0; q <= r; ++q) s += B[i + q * my] * T[q + r * ldt];
            B[i + r * my] = s;
        }
    /* B = -B U11^T: (B U11^T)(i,r) = B(i,r) + sum_{q<r} B(i,q) U11(r,q) */
    for (long i = 0; i < my; ++i)
        for (long r = b - 1; r >= 0; --r) {
            double s = B[i + r * my];
            for (long q = 0; q < r; ++q) s += B[i + q * my] * U[r + q * ldu];
#if ORC_PLANT == 1
            B[i + r * my] = s; /* planted: wrong sign of the (G V) T U11^T term */
#else
            B[i + r * my] = -s;
#endif
        }
    for (long r = 0; r < b; ++r)
        for (long i = 0; i < my; ++i) B[i + r * my] += G1[i + r * ldg];
    for (long c = 0; c < nr; ++c)
        for (long i = 0; i < my; ++i) {
            double s = 0.0;
#if ORC_PLANT == 3
            for (long r = 0; r + 1 < b; ++r) s += B[i + r * my] * R12[r + c * ldr]; /* planted: last row of R12 dropped */
#else
            for (long r = 0; r < b; ++r) s += B[i + r * my] * R12[r + c * ldr];
#endif
            Y2[i + c * ldy] -= s;
        }
    free(B);
#if ORC_PLANT != 2
    orc_larfb_right_n(mu, b, U, ldu, T, ldt, mg, G1, ldg);
#endif /* ORC_PLANT == 2 (planted): G is not carried along with Q */
}

void orc_hqrrp(long m, long n, double *A, long lda, long *jpvt, double *tau, long nb,
               long pp) {
    long mn = m < n ? m : n;
    if (mn == 0) return;
    orc_rng g;
    orc_srand(&g, 12);
    long mY = nb + pp;
    double *Y = (double *)malloc((size_t)mY * n * sizeof(double));
    double *V = (double *)malloc((size_t)mY * n * sizeof(double));
    double *W = (double *)calloc((size_t)nb * n, sizeof(double));
    double *G = (double *)malloc((size_t)mY * m * sizeof(double));
    orc_gaussian_matrix(&g, mY, m, G, mY);
    for (long j = 0; j < n; ++j) {
        jpvt[j] = j + 1;
        for (long i = 0; i < mY; ++i) {
            double s = 0.0;
            for (long q = 0; q < m; ++q) s += G[i + q * mY] * A[q + j * lda];
            Y[i + j * mY] = s;
        }
    }
    for (long j = 0; j < mn; j += nb) {
        long b = nb;
        if (n - j < b) b = n - j;
        if (m - j < b) b = m - j;
        int last = (j + nb >= m) || (j + nb >= n);
        if (!last) {
            long nVR = n - j;
            for (long c = 0; c < nVR; ++c)
                memcpy(V + (j + c) * mY, Y + (j + c) * mY, (size_t)mY * sizeof(double));
            double *tt = (double *)malloc((size_t)(b > 0 ? b : 1) * sizeof(double));
            orc_qrpmod(1, b, mY, nVR, V + j * mY, mY, jpvt + j, tt, m, A + j * lda, lda,
                       mY, Y + j * mY, mY, NULL, 0);
            free(tt);
        }
        /* panel QRP of [A11; A21] = A(j:m, j:j+b), pivoting A01 and Y1, building T */
        double *T = W + j * nb;
        orc_qrpmod(1, -1, m - j, b, A + j + j * lda, lda, jpvt + j, tau + j, j, A + j * lda,
                   lda, mY, Y + j * mY, mY, T, nb);
        if (j + b < n)
            orc_larfb_left_t(m - j, b, A + j + j * lda, lda, T, nb, n - j - b,
                             A + j + (j + b) * lda, lda);
        if (!last)
            orc_downdate_y(b, A + j + j * lda, lda, m - j, A + j + (j + b) * lda, lda,
                           n - j - b, T, nb, Y + (j + b) * mY, mY, mY, G + j * mY, mY, mY);
    }
    free(Y);
    free(V);
    free(W);
    free(G);
}

void orc_cpqr(long m, long n, double *A, long lda, long *jpvt, double *tau) {
    long mn = m < n ? m : n;
    if (mn < 1000)
        orc_geqp3(m, n, A, lda, jpvt, tau);
    else
        orc_hqrrp(m, n, A, lda, jpvt, tau, 64, 10);
}

void orc_ormqr_lt(long m, long k, const double *A, long lda, const double *tau,
                  double *c) {
    for (long i = 0; i < k; ++i) {
        if (tau[i] == 0.0) continue;
        const double *v = A + i + i * lda;
        double w = c[i];
        for (long q = i + 1; q < m; ++q) w += v[q - i] * c[q];
        double t = -tau[i] * w;
        c[i] = c[i] + t;
        for (long q = i + 1; q < m; ++q) c[q] = c[q] + v[q - i] * t;
    }
}

void orc_trsv_un(long n, const double *R, long ldr, double *x, double perturb) {
    for (long j = n - 1; j >= 0; --j) {
        if (x[j] != 0.0) {
            x[j] = x[j] / (R[j + j * ldr] + perturb);
            double t = x[j];
            for (long i = j - 1; i >= 0; --i) x[i] = x[i] - t * R[i + j * ldr];
        }
    }
}

void orc_trsv_ut(long n, const double *R, long ldr, double *x, double perturb) {
    for (long j = 0; j < n; ++j) {
        double t = x[j];
        for (long i = 0; i < j; ++i) t = t - R[i + j * ldr] * x[i];
        t = t / (R[j + j * ldr] + perturb);
        x[j] = t;
    }
}

double orc_rcond_inf(long n, const double *R, long ldr) {
    if (n == 0) return 1.0;
    double anorm = 0.0;
    for (long i = 0; i < n; ++i) {
        double s = 0.0;
        for (long j = i; j < n; ++j) s += fabs(R[i + j * ldr]);
        if (s > anorm) anorm = s;
    }
    /* row sums of |R^{-1}|: column j of R^{-1} by back substitution of e_j */
    double *rs = (double *)calloc((size_t)n, sizeof(double));
    double *e = (double *)malloc((size_t)n * sizeof(double));
    for (long j = 0; j < n; ++j) {
        for (long i = 0; i <= j; ++i) e[i] = (i == j) ? 1.0 : 0.0;
        orc_trsv_un(j + 1, R, ldr, e, 0.0);
        for (long i = 0; i <= j; ++i) rs[i] += fabs(e[i]);
    }
    double ainv = 0.0;
    for (long i = 0; i < n; ++i)
        if (rs[i] > ainv) ainv = rs[i];
    free(rs);
    free(e);
    if (anorm == 0.0 || ainv == 0.0 || !isfinite(ainv)) return 0.0;
    return (1.0 / anorm) / ainv;
}

/* ======================================================================== */
/* LSQR (IterSolver.cpp:160-284)                                             */
/* ======================================================================== */
typedef struct {
    int sparse;
    long n, r;
    const double *A;
    long lda;
    const long *Ap, *Ai;
    const double *Ax;
    const long *E;
    const double *R;
    long ldr;
    double perturb;
    double *tmp; /* r */
} orc_op;

static long col_of(const orc_op *op, long j) { return op->E ? op->E[j] : j; }

/* y = op*v + beta_y*y   (dense: dgemv 'N' order; sparse: scatter then combine) */
static void orc_op_n(orc_op *op, const double *v, double beta_y, double *y) {
    long n = op->n, r = op->r;
    memcpy(op->tmp, v, (size_t)r * sizeof(double));
    orc_trsv_un(r, op->R, op->ldr, op->tmp, op->perturb);
    if (!op->sparse) {
        for (long i = 0; i < n; ++i) y[i] = beta_y * y[i];
        for (long j = 0; j < r; ++j) {
            double t = 1.0 * op->tmp[j];
            const double *a = op->A + col_of(op, j) * op->lda;
            for (long i = 0; i < n; ++i) y[i] = y[i] + t * a[i];
        }
    } else {
        double *W = (double *)calloc((size_t)n, sizeof(double));
        for (long j = 0; j < r; ++j) {
            long c = col_of(op, j);
            for (long p = op->Ap[c]; p < op->Ap[c + 1]; ++p)
                W[op->Ai[p]] += 1.0 * op->Ax[p] * op->tmp[j];
        }
        for (long i = 0; i < n; ++i) y[i] = beta_y * y[i] + 1.0 * W[i];
        free(W);
    }
}

/* v = op^T u + beta_v v */
static void orc_op_t(orc_op *op, const double *u, double beta_v, double *v) {
    long n = op->n, r = op->r;
    for (long j = 0; j < r; ++j) {
        long c = col_of(op, j);
        double t = 0.0;
        if (!op->sparse) {
            const double *a = op->A + c * op->lda;
            for (long i = 0; i < n; ++i) t = t + a[i] * u[i];
        } else {
            for (long p = op->Ap[c]; p < op->Ap[c + 1]; ++p)
                t += 1.0 * op->Ax[p] * u[op->Ai[p]];
        }
        op->tmp[j] = t;
    }
    orc_trsv_ut(r, op->R, op->ldr, op->tmp, op->perturb);
    for (long j = 0; j < r; ++j) v[j] = beta_v * v[j];
    for (long j = 0; j < r; ++j) v[j] = v[j] + 1.0 * op->tmp[j];
}

static void orc_lsqr_core(orc_op *op, const double *b, double tol, long max_it, double *x,
                          int *flag, long *it, orc_lsqr_step *hist, orc_info *info) {
    long n = op->n, r = op->r;
    double *u = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double *v = (double *)calloc((size_t)(r > 0 ? r : 1), sizeof(double));
    double *w = (double *)calloc((size_t)(r > 0 ? r : 1), sizeof(double));
    double alpha = 0.0;
    memcpy(u, b, (size_t)n * sizeof(double));
    double beta = orc_nrm2(n, u, 1);
    if (beta > 0) {
        double sc = 1.0 / beta;
        for (long i = 0; i < n; ++i) u[i] = sc * u[i];
        orc_op_t(op, u, 0.0, v);
        alpha = orc_nrm2(r, v, 1);
    }
    if (alpha > 0) {
        double sc = 1.0 / alpha;
        for (long j = 0; j < r; ++j) v[j] = sc * v[j];
        memcpy(w, v, (size_t)r * sizeof(double));
    }
    for (long j = 0; j < r; ++j) x[j] = 0.0;
    double nrm_ar = alpha * beta;
    double nrm_a = 0.0, nrm_r = beta;
    if (nrm_ar == 0) {
        *it = 0;
        *flag = 0;
    } else {
        double phi, rho, cs, sn, theta;
        double phibar = beta, rhobar = alpha;
        *flag = 1;
        *it = 0;
        for (long kk = 0; kk < max_it; ++kk) {
            (*it)++;
            orc_op_n(op, v, -alpha, u);
            beta = orc_nrm2(n, u, 1);
            if (beta > 0) {
                double sc = 1.0 / beta;
                for (long i = 0; i < n; ++i) u[i] = sc * u[i];
                nrm_a = sqrt(nrm_a * nrm_a + alpha * alpha + beta * beta);
                orc_op_t(op, u, -beta, v);
                alpha = orc_nrm2(r, v, 1);
                if (alpha > 0) {
                    double sc2 = 1.0 / alpha;
                    for (long j = 0; j < r; ++j) v[j] = sc2 * v[j];
                }
            }
            rho = sqrt(rhobar * rhobar + beta * beta);
            cs = rhobar / rho;
            sn = beta / rho;
            theta = sn * alpha;
            rhobar = -cs * alpha;
            phi = cs * phibar;
            phibar = sn * phibar;
            double c1 = phi / rho;
            for (long j = 0; j < r; ++j) x[j] = x[j] + c1 * w[j];
            double c2 = -theta / rho;
            for (long j = 0; j < r; ++j) w[j] = c2 * w[j];
            for (long j = 0; j < r; ++j) w[j] = w[j] + 1.0 * v[j];
            nrm_r = phibar;
            nrm_ar = phibar * alpha * fabs(cs);
            if (hist) {
                hist[kk].alpha = alpha;
                hist[kk].beta = beta;
                hist[kk].rhobar = rhobar;
                hist[kk].phibar = phibar;
                hist[kk].nrm_a = nrm_a;
            }
            if (nrm_ar < tol * nrm_a * nrm_r) {
                *flag = 0;
                break;
            }
        }
    }
    if (info) {
        info->lsqr_nrm_ar = nrm_ar;
        info->lsqr_nrm_a = nrm_a;
        info->lsqr_nrm_r = nrm_r;
    }
    free(u);
    free(v);
    free(w);
}

void orc_lsqr_dense(long n, long r, const double *A, long lda, const long *E,
                    const double *R, long ldr, const double *b, double tol, long max_it,
                    double *y, int *flag, long *it, orc_lsqr_step *hist, orc_info *info) {
    orc_op op;
    memset(&op, 0, sizeof op);
    op.n = n;
    op.r = r;
    op.A = A;
    op.lda = lda;
    op.E = E;
    op.R = R;
    op.ldr = ldr;
    op.tmp = (double *)malloc((size_t)(r > 0 ? r : 1) * sizeof(double));
    orc_lsqr_core(&op, b, tol, max_it, y, flag, it, hist, info);
    free(op.tmp);
}

void orc_lsqr_sparse(long n, long d, const long *Ap, const long *Ai, const double *Ax,
                     long r, const long *E, const double *R, long ldr, double perturb,
                     const double *b, double tol, long max_it, double *y, int *flag,
                     long *it, orc_lsqr_step *hist, orc_info *info) {
    orc_op op;
    (void)d;
    memset(&op, 0, sizeof op);
    op.sparse = 1;
    op.n = n;
    op.r = r;
    op.Ap = Ap;
    op.Ai = Ai;
    op.Ax = Ax;
    op.E = E;
    op.R = R;
    op.ldr = ldr;
    op.perturb = perturb;
    op.tmp = (double *)malloc((size_t)(r > 0 ? r : 1) * sizeof(double));
    orc_lsqr_core(&op, b, tol, max_it, y, flag, it, hist, info);
    free(op.tmp);
}

/* ======================================================================== */
/* Solvers                                                                   */
/* ======================================================================== */
static long ceil_gamma(long d, double gamma) { return (long)ceil((double)d * gamma); }

/* rand_hashing_dct (dense_helpers.cpp:62-112) + sketch (sketching.cpp:17-62):
 * SA = (1/sqrt k) S H D A and, if b, c = (1/sqrt k) S H D b. */
static void orc_dense_sketch(long n, long d, const double *A, long lda, const double *b,
                             double gamma, long k, unsigned seed, double *SA, double *c) {
    long m = ceil_gamma(d, gamma);
    orc_rng g;
    orc_srand(&g, seed);
    double *D = (double *)malloc((size_t)n * sizeof(double));
    orc_gen_signs(&g, n, 1.0, D);
    double *HDA = (double *)malloc((size_t)n * d * sizeof(double));
    orc_fut(n, d, A, lda, D, HDA, n);
    double *HDb = NULL;
    if (b) {
        HDb = (double *)malloc((size_t)n * sizeof(double));
        orc_fut(n, 1, b, n, D, HDb, n);
    }
    long *rows = (long *)malloc((size_t)n * k * sizeof(long));
    double *vals = (double *)malloc((size_t)n * k * sizeof(double));
    orc_gen_hashing(&g, m, n, k, rows, NULL, vals);
    double alpha = 1 / sqrt((double)k);
    orc_hash_apply(m, n, k, rows, vals, alpha, d, HDA, n, SA, m);
    if (b) orc_hash_apply(m, n, k, rows, vals, alpha, 1, HDb, n, c, m);
    free(D);
    free(HDA);
    free(HDb);
    free(rows);
    free(vals);
}

int orc_ls_dense(long n, long d, double *A, long lda, const double *b, double *x,
                 double gamma, long k, double abs_tol, double rcond, double it_tol,
                 long max_it, unsigned seed, orc_lsqr_step *hist, orc_info *info) {
    orc_info loc;
    if (!info) info = &loc;
    memset(info, 0, sizeof *info);
    if (n < d) return ORC_E_SHAPE;
    double t0 = now_sec();
    long m = ceil_gamma(d, gamma);
    double *SA = (double *)malloc((size_t)m * d * sizeof(double));
    double *c = (double *)malloc((size_t)m * sizeof(double));
    orc_dense_sketch(n, d, A, lda, b, gamma, k, seed, SA, c);
    double t1 = now_sec();
    long *E = (long *)calloc((size_t)d, sizeof(long));
    double *tau = (double *)calloc((size_t)d, sizeof(double));
    orc_cpqr(m, d, SA, m, E, tau);
    orc_ormqr_lt(m, d < m ? d : m, SA, m, tau, c);
    long rank = 0;
    for (long i = 0; i < d; ++i) {
        if (fabs(SA[i + i * m]) > rcond)
            rank++;
        else
            break;
    }
    double t2 = now_sec();
    orc_trsv_un(rank, SA, m, c, 0.0);
    for (long i = 0; i < rank; ++i) x[E[i] - 1] = c[i];
    for (long i = rank; i < d; ++i) x[E[i] - 1] = 0.0;
    double *rs = (double *)malloc((size_t)n * sizeof(double));
    memcpy(rs, b, (size_t)n * sizeof(double));
    for (long j = 0; j < d; ++j) {
        double t = -1.0 * x[j];
        for (long i = 0; i < n; ++i) rs[i] = rs[i] + t * A[i + j * lda];
    }
    double r0 = orc_nrm2(n, rs, 1);
    info->x0_residual = r0;
    info->rank = rank;
    double t3 = now_sec();
    if (r0 < abs_tol) {
        info->it = 0;
        info->flag = 0;
    } else {
        double *xs = (double *)malloc((size_t)d * sizeof(double));
        memcpy(xs, x, (size_t)d * sizeof(double));
        long *E0 = (long *)malloc((size_t)d * sizeof(long));
        for (long i = 0; i < d; ++i) E0[i] = E[i] - 1;
        double *y = (double *)calloc((size_t)(rank > 0 ? rank : 1), sizeof(double));
        orc_lsqr_dense(n, rank, A, lda, E0, SA, m, rs, it_tol, max_it, y, &info->flag,
                       &info->it, hist, info);
        orc_trsv_un(rank, SA, m, y, 0.0);
        for (long i = 0; i < rank; ++i) x[E0[i]] = y[i];
        for (long i = rank; i < d; ++i) x[E0[i]] = 0.0;
        for (long i = 0; i < d; ++i) x[i] = xs[i] + x[i];
        free(xs);
        free(E0);
        free(y);
    }
    double t4 = now_sec();
    info->t_sketch = t1 - t0;
    info->t_qr = t2 - t1;
    info->t_x0 = t3 - t2;
    info->t_lsqr = t4 - t3;
    info->t_total = t4 - t0;
    free(SA);
    free(c);
    free(E);
    free(tau);
    free(rs);
    return ORC_OK;
}

int orc_ls_dense_nocpqr(long n, long d, const double *A, long lda, const double *b,
                        double *x, double gamma, long k, double it_tol, long max_it,
                        unsigned seed, orc_lsqr_step *hist, orc_info *info) {
    orc_info loc;
    if (!info) info = &loc;
    memset(info, 0, sizeof *info);
    if (n < d) return ORC_E_SHAPE;
    double t0 = now_sec();
    long m = ceil_gamma(d, gamma);
    double *SA = (double *)malloc((size_t)m * d * sizeof(double));
    orc_dense_sketch(n, d, A, lda, NULL, gamma, k, seed, SA, NULL);
    double t1 = now_sec();
    double *tau = (double *)calloc((size_t)d, sizeof(double));
    orc_geqrf(m, d, SA, m, tau);
    double t2 = now_sec();
    info->rank = d;
    orc_lsqr_dense(n, d, A, lda, NULL, SA, m, b, it_tol, max_it, x, &info->flag, &info->it,
                   hist, info);
    orc_trsv_un(d, SA, m, x, 0.0);
    double t3 = now_sec();
    info->t_sketch = t1 - t0;
    info->t_qr = t2 - t1;
    info->t_lsqr = t3 - t2;
    info->t_total = t3 - t0;
    free(SA);
    free(tau);
    return ORC_OK;
}

int orc_ls_dense_blendenpik(long n, long d, const double *A, long lda, const double *b,
                            double *x, double gamma, double it_tol, long max_it,
                            unsigned seed, orc_lsqr_step *hist, orc_info *info) {
    orc_info loc;
    if (!info) info = &loc;
    memset(info, 0, sizeof *info);
    if (n < d) return ORC_E_SHAPE;
    double t0 = now_sec();
    long m = ceil_gamma(d, gamma);
    if (m > n) return ORC_E_SHAPE; /* select_k_from_n asserts n >= k */
    /* gen_scaled_rand_signs seeds from the clock, and random_sample seeds
     * again from the same clock (dense_helpers.cpp:12,25) */
    orc_rng g;
    orc_srand(&g, seed);
    double *D = (double *)malloc((size_t)n * sizeof(double));
    orc_gen_signs(&g, n, 1.0, D);
    double *HDA = (double *)malloc((size_t)n * d * sizeof(double));
    orc_fut(n, d, A, lda, D, HDA, n);
    long *idx = (long *)malloc((size_t)m * sizeof(long));
    orc_srand(&g, seed);
    orc_select_k_from_n(&g, n, m, idx);
    double *SA = (double *)malloc((size_t)m * d * sizeof(double));
    for (long j = 0; j < d; ++j)
        for (long i = 0; i < m; ++i) SA[i + j * m] = HDA[idx[i] + j * n];
    double t1 = now_sec();
    double *tau = (double *)calloc((size_t)d, sizeof(double));
    orc_geqrf(m, d, SA, m, tau); /* dense_qr */
    double t2 = now_sec();
    info->rank = d;
    orc_lsqr_dense(n, d, A, lda, NULL, SA, m, b, it_tol, max_it, x, &info->flag, &info->it,
                   hist, info);
    orc_trsv_un(d, SA, m, x, 0.0);
    double t3 = now_sec();
    info->t_sketch = t1 - t0;
    info->t_qr = t2 - t1;
    info->t_lsqr = t3 - t2;
    info->t_total = t3 - t0;
    free(D);
    free(HDA);
    free(idx);
    free(SA);
    free(tau);
    return ORC_OK;
}

long orc_sparse_precond(long m, long d, double *SA, long ldsa, double *bs, long *E,
                        double *tol_out) {
    double maxn = 0.0;
    for (long j = 0; j < d; ++j) {
        double s = orc_nrm2(m, SA + j * ldsa, 1);
        if (s > maxn) maxn = s;
    }
    double tol = 20.0 * (double)(m + d) * DBL_EPSILON * maxn;
    if (tol_out) *tol_out = tol;
    double *W = (double *)malloc((size_t)m * (d + 1) * sizeof(double));
    for (long j = 0; j < d; ++j) memcpy(W + j * m, SA + j * ldsa, (size_t)m * sizeof(double));
    memcpy(W + d * m, bs, (size_t)m * sizeof(double));
    double *tau = (double *)calloc((size_t)(d + 1), sizeof(double));
    long kmax = m < d ? m : d;
    /* Householder QR of [SA | b_s] in natural order: the reflectors only act on
     * the first d columns, the last column receives Q^T b_s. */
    for (long i = 0; i < kmax; ++i) {
        double *aii = W + i + i * m;
        orc_dlarfg(m - i, aii, aii + 1, 1, &tau[i]);
        double diag = *aii;
        *aii = 1.0;
        orc_dlarf_left(m - i, d - i, aii, tau[i], W + i + (i + 1) * m, m);
        *aii = diag;
    }
    long rank = kmax;
    for (long i = 0; i < kmax; ++i)
        if (!(fabs(W[i + i * m]) > tol)) {
            rank = -1;
            break;
        }
    if (rank == kmax) {
        for (long j = 0; j < d; ++j) E[j] = j;
    } else {
        /* rank deficiency detected: column-pivoted QR on a fresh copy */
        long *jp = (long *)malloc((size_t)d * sizeof(long));
        for (long j = 0; j < d; ++j) memcpy(W + j * m, SA + j * ldsa, (size_t)m * sizeof(double));
        orc_geqp3(m, d, W, m, jp, tau);
        memcpy(W + d * m, bs, (size_t)m * sizeof(double));
        orc_ormqr_lt(m, kmax, W, m, tau, W + d * m);
        rank = 0;
        for (long i = 0; i < kmax; ++i) {
            if (fabs(W[i + i * m]) > tol)
                rank++;
            else
                break;
        }
        for (long j = 0; j < d; ++j) E[j] = jp[j] - 1;
        free(jp);
    }
    for (long j = 0; j < d; ++j) memcpy(SA + j * ldsa, W + j * m, (size_t)m * sizeof(double));
    memcpy(bs, W + d * m, (size_t)m * sizeof(double));
    free(W);
    free(tau);
    return rank;
}

int orc_ls_sparse(long n, long d, const long *Ap, const long *Ai, const double *Ax,
                  const double *b, double *x, double gamma, long k, double abs_tol,
                  double it_tol, long max_it, double rcond_threshold, double perturb,
                  unsigned seed, orc_lsqr_step *hist, orc_info *info) {
    orc_info loc;
    if (!info) info = &loc;
    memset(info, 0, sizeof *info);
    if (n < d) return ORC_E_SHAPE;
    double t0 = now_sec();
    long m = ceil_gamma(d, gamma);
    orc_rng g;
    orc_srand(&g, seed);
    long *Srows = (long *)malloc((size_t)n * k * sizeof(long));
    double *Svals = (double *)malloc((size_t)n * k * sizeof(double));
    orc_gen_hashing(&g, m, n, k, Srows, NULL, Svals);
    double *bs = (double *)malloc((size_t)m * sizeof(double));
    orc_sketch_vec(m, n, k, Srows, Svals, b, bs);
    double *SA = (double *)malloc((size_t)m * d * sizeof(double));
    orc_sketch_sparse(m, n, k, Srows, Svals, d, Ap, Ai, Ax, SA, m);
    double t1 = now_sec();
    long *E = (long *)malloc((size_t)d * sizeof(long));
    long rank = orc_sparse_precond(m, d, SA, m, bs, E, NULL);
    free(Srows);
    free(Svals);
    if (rank == 0) {
        free(bs);
        free(SA);
        free(E);
        return ORC_E_RANK0;
    }
    double rc = orc_rcond_inf(rank, SA, m);
    info->rcond = rc;
    info->success = rc > rcond_threshold ? 1 : 0;
    info->rank = rank;
    double t2 = now_sec();
    /* x0 = P R11^{-1} c (ls_sparse_spqr.cpp:144-151, no perturbation) */
    orc_trsv_un(rank, SA, m, bs, 0.0);
    for (long i = 0; i < rank; ++i) x[E[i]] = bs[i];
    for (long i = rank; i < d; ++i) x[E[i]] = 0.0;
    double *rs = (double *)malloc((size_t)n * sizeof(double));
    memcpy(rs, b, (size_t)n * sizeof(double));
    {
        /* r_s = b - A x : CSC_Mat::mv('n', -1, x, 1, r_s) */
        double *W = (double *)calloc((size_t)n, sizeof(double));
        for (long j = 0; j < d; ++j)
            for (long p = Ap[j]; p < Ap[j + 1]; ++p) W[Ai[p]] += -1.0 * Ax[p] * x[j];
        for (long i = 0; i < n; ++i) rs[i] = 1.0 * rs[i] + 1.0 * W[i];
        free(W);
    }
    double r0 = orc_nrm2(n, rs, 1);
    info->x0_residual = r0;
    double t3 = now_sec();
    if (r0 < abs_tol) {
        info->it = 0;
        info->flag = 0;
    } else {
        double *y = (double *)calloc((size_t)rank, sizeof(double));
        orc_lsqr_sparse(n, d, Ap, Ai, Ax, rank, E, SA, m,
                        info->success ? 0.0 : perturb, rs, it_tol, max_it, y, &info->flag,
                        &info->it, hist, info);
        orc_trsv_un(rank, SA, m, y, 0.0); /* ls_sparse_spqr.cpp:188, no perturbation */
        for (long i = 0; i < rank; ++i) y[i] = y[i] + bs[i];
        for (long i = 0; i < d; ++i) x[i] = 0.0;
        for (long i = 0; i < rank; ++i) x[E[i]] = 1.0 * y[i];
        free(y);
    }
    double t4 = now_sec();
    info->t_sketch = t1 - t0;
    info->t_qr = t2 - t1;
    info->t_x0 = t3 - t2;
    info->t_lsqr = t4 - t3;
    info->t_total = t4 - t0;
    free(bs);
    free(SA);
    free(E);
    free(rs);
    return ORC_OK;
}
