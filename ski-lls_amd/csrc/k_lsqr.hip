// k_lsqr.hip -- LSQR (src/IterSolver.cpp:160-284) on the preconditioned
// operator A(:, E[0:r]) M with M = R^{-1}, in FP64 on the device.
//
// Each iteration is a handful of kernels that form A z and A^T u, apply M
// and M^T as triangular GEMVs, and keep every scalar of the Golub-Kahan /
// Givens recurrences in device memory.  Dense A (d <= 8192, when the device
// has room for a second copy) is read once per iteration, from a row-major
// copy made at the start of the solve, by one kernel that forms u = A z and
// A^T u together; otherwise A z and A^T u are two passes over the
// column-major A (coalesced 16-byte loads) or over CSR and CSC for sparse A.  Partial sums are reduced in a fixed
// order, so a solve is bitwise reproducible run to run.  The iteration
// sequence is captured once in a hipGraph and replayed until the device-side
// convergence flag is set; kernels of iterations past convergence exit at
// entry.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>
#include <type_traits>
#include <utility>

#include <hip/hip_ext.h>

#include <cstdio>

#include <array>
#include <cmath>
#include <map>
#include <mutex>

#include "k_common.h"

namespace skl {

struct LsqrState {
    double alpha, beta, rhobar, phibar, nrm_a, nrm_r, nrm_ar, coef, tol;
    long it, max_it;
    int flag, done;
};

// ---- A z: column-chunked partials, 2 rows per thread ----------------------
constexpr int GN_THREADS = 256, GN_ROWS = 2 * GN_THREADS, GN_MAXCW = 512;

// P[c * n + i] = sum_{j in chunk c} A[i, j] z[j]; with upper != 0 only the
// entries j >= i contribute (A upper triangular, zeros not read).
// ROLE only names the instantiation (0: LSQR A z, 1: triangular M v, 2: other)
// so that per-kernel profiles do not mix launch shapes.
template <int ROLE>
__global__ __launch_bounds__(GN_THREADS) void k_gemv_n_part(
    long n, long d, const double *__restrict__ A, long lda, const double *__restrict__ z, int cw,
    double *__restrict__ P, int upper, const LsqrState *__restrict__ st) {
    if (st && st->done) return;
    __shared__ double zs[GN_MAXCW];
    const long j0 = (long)blockIdx.y * cw;
    const long j1 = min(d, j0 + cw);
    for (long t = threadIdx.x; t < j1 - j0; t += GN_THREADS) zs[t] = z[j0 + t];
    __syncthreads();
    const long i = (long)blockIdx.x * GN_ROWS + 2 * threadIdx.x;
    if (i >= n) return;
    double a0 = 0.0, a1 = 0.0;
    long jb = j0;
    if (upper && jb < i) jb = i; // rows below the diagonal of this chunk see zeros
    const double *a = A + i + jb * lda;
    if (i + 1 < n && ((lda & 1) == 0)) {
        long j = jb;
        // eight columns' loads in flight before their products (the sums stay
        // in column order)
        for (; j + 8 <= j1; j += 8) {
            double2 v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = *(const double2 *)(a + q * lda);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const double zq = zs[j - j0 + q];
                a0 += v[q].x * zq;
                a1 += v[q].y * zq;
            }
            a += 8 * lda;
        }
        for (; j + 4 <= j1; j += 4) {
            double2 v0 = *(const double2 *)(a);
            double2 v1 = *(const double2 *)(a + lda);
            double2 v2 = *(const double2 *)(a + 2 * lda);
            double2 v3 = *(const double2 *)(a + 3 * lda);
            double z0 = zs[j - j0], z1 = zs[j - j0 + 1], z2 = zs[j - j0 + 2], z3 = zs[j - j0 + 3];
            a0 += v0.x * z0;
            a1 += v0.y * z0;
            a0 += v1.x * z1;
            a1 += v1.y * z1;
            a0 += v2.x * z2;
            a1 += v2.y * z2;
            a0 += v3.x * z3;
            a1 += v3.y * z3;
            a += 4 * lda;
        }
        for (; j < j1; ++j) {
            double2 v0 = *(const double2 *)(a);
            double z0 = zs[j - j0];
            a0 += v0.x * z0;
            a1 += v0.y * z0;
            a += lda;
        }
    } else {
        for (long j = jb; j < j1; ++j) {
            double zz = zs[j - j0];
            a0 += a[0] * zz;
            if (i + 1 < n) a1 += a[1] * zz;
            a += lda;
        }
    }
    P[(long)blockIdx.y * n + i] = a0;
    if (i + 1 < n) P[(long)blockIdx.y * n + i + 1] = a1;
}

// ---- A^T u: row-blocked partials, one wave per column -------------------------
constexpr int GT_THREADS = 256, GT_RB = 1024;

// P[rb * d + j] = sum_{i in row block rb} A[i, j] u[i]; with upper != 0 only
// rows i <= j contribute.
template <int ROLE>
__global__ __launch_bounds__(GT_THREADS) void k_gemv_t_part(
    long n, long d, const double *__restrict__ A, long lda, const double *__restrict__ u, int cpw,
    double *__restrict__ P, int upper, const LsqrState *__restrict__ st) {
    if (st && st->done) return;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long i0 = (long)blockIdx.x * GT_RB;
    double2 uu[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        long ii = i0 + 128 * t + 2 * lane;
        uu[t].x = (ii < n) ? u[ii] : 0.0;
        uu[t].y = (ii + 1 < n) ? u[ii + 1] : 0.0;
    }
    const bool vec = ((lda & 1) == 0) && (i0 + GT_RB <= n);
    // the LSQR product walks the columns from the last: the A z pass before it
    // ends there, so its first reads find that pass's last columns still in
    // the memory-side cache (and the next A z pass starts where this one ends)
    const long yb = ROLE == 0 ? (long)(gridDim.y - 1 - blockIdx.y) : (long)blockIdx.y;
    const long jbeg = yb * cpw, jend = min(d, jbeg + cpw);
    for (long j = jbeg + wave; j < jend; j += 4) {
        double s = 0.0;
        if (!(upper && i0 > j)) {
            const double *a = A + i0 + j * lda;
            if (vec) {
                double2 v[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) v[t] = *(const double2 *)(a + 128 * t + 2 * lane);
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    s += v[t].x * uu[t].x;
                    s += v[t].y * uu[t].y;
                }
            } else {
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    long ii = i0 + 128 * t + 2 * lane;
                    if (ii < n) s += a[128 * t + 2 * lane] * uu[t].x;
                    if (ii + 1 < n) s += a[128 * t + 2 * lane + 1] * uu[t].y;
                }
            }
        }
        s = wave_sum(s);
        if (lane == 0) P[(long)blockIdx.x * d + j] = s;
    }
}

// ---- M^T t for upper-triangular M: row-block partials ---------------------------
// P[rb * r + j] = sum over rows i in block rb (GT_RB rows), i <= j, of
// M[i, j] t[i], for the columns j >= rb * GT_RB only; entries with
// j < rb * GT_RB are never written (the caller zeroes them once).
// A workgroup per (row block rb, TC_COLS consecutive columns j >= rb GT_RB):
// the block's t in registers once (lane l holds t[i0 + l + 64 u]), each wave
// over TC_COLS / 4 columns two at a time (32 loads in flight); per column the
// products are summed by halving over the lane's values, then over the wave.
// Eight columns per workgroup (two per wave, one round of loads): 4000 x 4000
// M^T t 18.6 -> 16.2 us against 32 columns (four rounds per wave, 1.2
// workgroups per CU); TC_COLS is a multiple of 8.
constexpr int TC_COLS = 8;
static_assert(TC_COLS % 8 == 0, "two columns per wave step");
__global__ __launch_bounds__(256) void k_trmv_t_chunks(long r, const double *__restrict__ M, long ldm,
                                                       const double *__restrict__ t,
                                                       double *__restrict__ P, int nrb,
                                                       const LsqrState *__restrict__ st) {
    if (st && st->done) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long chunk = blockIdx.x;
    int rb = 0;
    for (; rb < nrb; ++rb) {
        const long cnt = (r - (long)rb * GT_RB + TC_COLS - 1) / TC_COLS;
        if (chunk < cnt) break;
        chunk -= cnt;
    }
    if (rb >= nrb) return;
    const long i0 = (long)rb * GT_RB, iend = min(r, i0 + GT_RB);
    constexpr int U = GT_RB / 64;
    double tv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const long i = i0 + lane + 64 * u;
        tv[u] = i < iend ? t[i] : 0.0;
    }
    constexpr int CPW = TC_COLS / 4;
    const long jb = i0 + chunk * TC_COLS + (long)w * CPW;
    for (int c = 0; c < CPW; c += 2) {
        double s0[U], s1[U];
        const long j0 = jb + c, j1 = jb + c + 1;
        const double *c0 = M + min(j0, r - 1) * ldm, *c1 = M + min(j1, r - 1) * ldm;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long i = i0 + lane + 64 * u;
            s0[u] = (j0 < r && i <= j0 && i < iend) ? c0[i] : 0.0;
            s1[u] = (j1 < r && i <= j1 && i < iend) ? c1[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            s0[u] = s0[u] * tv[u];
            s1[u] = s1[u] * tv[u];
        }
#pragma unroll
        for (int h = U / 2; h > 0; h >>= 1)
#pragma unroll
            for (int u = 0; u < h; ++u) {
                s0[u] += s0[u + h];
                s1[u] += s1[u + h];
            }
        const double a0 = wave_sum(s0[0]), a1 = wave_sum(s1[0]);
        if (lane == 0) {
            if (j0 < r) P[(long)rb * r + j0] = a0;
            if (j1 < r) P[(long)rb * r + j1] = a1;
        }
    }
}
static long trmv_t_chunks(long r) {
    long n = 0;
    for (long i0 = 0; i0 < r; i0 += GT_RB) n += (r - i0 + TC_COLS - 1) / TC_COLS;
    return n;
}

__global__ void k_zero_d(long n, double *__restrict__ p) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) p[i] = 0.0;
}


// ---- one pass over A: u <- A z - coef u and A^T u together ---------------------
// The pass reads a row-major copy AT of A (rows of ldt doubles, ldt even, the
// padding column zero), made once per solve by k_to_rows.  A 256-thread
// workgroup takes the rows blockIdx.x + k G; its four waves share each row:
// thread t holds the column pairs 2t + 512c (c < NC) of the row in registers,
// so one load instruction of a wave reads 1 KB of the row, and the next row is
// loaded while this one is used.  Per row the dot with z (z in LDS, zero past
// d) is reduced over the wave by shuffles and over the four waves through LDS
// (one barrier, two alternating slots); every thread then forms u of the row
// and adds u times its columns to its register sums of A^T u.  Per workgroup
// the partial A^T u goes to Q[wg][:] and the sum of squares of u to Sb[wg].
constexpr int FR_T = 256, FR_W = FR_T / 64, FR_COLS = 2 * FR_T;

template <int NC>
__global__ __launch_bounds__(FR_T) void k_fused_rows(long n, long d, const double *__restrict__ AT,
                                                     long ldt, const double *__restrict__ z,
                                                     double *__restrict__ ut,
                                                     double *__restrict__ Sb,
                                                     double *__restrict__ Q,
                                                     const LsqrState *__restrict__ st) {
    if (st->done) return;
    extern __shared__ double fsm[];
    double *zs = fsm;                // NC * FR_COLS
    double *red = fsm + NC * FR_COLS; // [2][FR_W]
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    for (int j = tid; j < NC * FR_COLS; j += FR_T) zs[j] = j < d ? z[j] : 0.0;
    const double coef = st->coef;
    // loads are unconditional (a load under a branch makes every wait on a
    // row wait for the prefetched one too): column pairs past d read pair 0
    // and meet zeros of z; rows past n read row n - 1 and are never used
    int off[NC];
    double2 q[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        off[c] = 2L * tid + (long)FR_COLS * c < d ? tid + FR_T * c : 0;
        q[c] = make_double2(0.0, 0.0);
    }
    double sq = 0.0;
    __syncthreads();
    const long G = gridDim.x;
    auto load = [&](long i, double2(&a)[NC], double &uo) {
        const long ic = i < n ? i : n - 1;
        const double2 *ap = (const double2 *)(AT + ic * ldt);
#pragma unroll
        for (int c = 0; c < NC; ++c) a[c] = ap[off[c]];
        uo = ut[ic];
    };
    int par = 0;
    auto row = [&](long i, const double2(&a)[NC], double uo) {
        const double2 *z2 = (const double2 *)zs + tid;
        double p0 = 0.0, p1 = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double2 zz = z2[FR_T * c];
            p0 = fma(a[c].x, zz.x, p0);
            p1 = fma(a[c].y, zz.y, p1);
        }
        const double p = wave_sum(p0 + p1);
        double *rb = red + par * FR_W;
        if (lane == 0) rb[w] = p;
        __syncthreads();
        const double s = (rb[0] + rb[1]) + (rb[2] + rb[3]);
        const double uv = i < n ? s - coef * uo : 0.0;
        if (tid == 0 && i < n) {
            ut[i] = uv;
            sq += uv * uv;
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            q[c].x = fma(a[c].x, uv, q[c].x);
            q[c].y = fma(a[c].y, uv, q[c].y);
        }
        par ^= 1;
    };
    double2 a0[NC], a1[NC];
    double u0, u1;
    long i = blockIdx.x;
    load(i, a0, u0);
    for (; i < n; i += 2 * G) {
        load(i + G, a1, u1);
        row(i, a0, u0);
        if (i + G >= n) break;
        load(i + 2 * G, a0, u0);
        row(i + G, a1, u1);
    }
    double *qo = Q + (long)blockIdx.x * d;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const long j = 2L * tid + (long)FR_COLS * c;
        if (j < d) qo[j] = q[c].x;
        if (j + 1 < d) qo[j + 1] = q[c].y;
    }
    if (tid == 0) Sb[blockIdx.x] = sq;
}

// AT[i * ldt + j] = A[i + j * lda] for i < n, j < ldt (zero for d <= j < ldt),
// through 64 x 64 tiles in LDS: the reads run down columns, the writes along rows
__global__ __launch_bounds__(256) void k_to_rows(long n, long d, const double *__restrict__ A,
                                                 long lda, double *__restrict__ AT, long ldt) {
    __shared__ double tile[64][65];
    const long i0 = (long)blockIdx.x * 64, j0 = (long)blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int jj = ty; jj < 64; jj += 4) {
        const long i = i0 + tx, j = j0 + jj;
        tile[jj][tx] = (i < n && j < d) ? A[i + j * lda] : 0.0;
    }
    __syncthreads();
    for (int ii = ty; ii < 64; ii += 4) {
        const long i = i0 + ii, j = j0 + tx;
        if (i < n && j < ldt) AT[i * ldt + j] = tile[tx][ii];
    }
}

// ---- reductions / scatter ----------------------------------------------------
// sum over q = q0, q0 + step, ... < qn of p[q * stride], added in increasing q
// (the order of a plain loop), with eight loads in flight at a time: a loop
// that adds each load before issuing the next waits the memory latency per
// term.  Past qn the loads repeat entry q0 and are multiplied by 0 (an exact
// no-op on the finite partial sums; a product by 0 is not folded into a
// select, which would let the compiler put the load behind a branch).
__device__ inline double ordered_sum(const double *__restrict__ p, long stride, int q0, int qn,
                                     int step) {
    double s = 0.0;
    for (int q = q0; q < qn; q += 8 * step) {
        double w[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int qk = q + k * step;
            w[k] = p[(long)(qk < qn ? qk : q0) * stride];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) s += w[k] * (q + k * step < qn ? 1.0 : 0.0);
    }
    return s;
}
// Sums of partial products: 64 outputs per 256-thread block, each summed by
// four threads over every fourth partial, then combined in a fixed order.
constexpr int RED_OUT = 64, RED_GRP = 4;

// z[j] = sum_c Pz[c * r + Einv[j]]  (0 when column j is outside E[0:r])
__global__ __launch_bounds__(256) void k_z_reduce(long d, long r, int nck,
                                                  const double *__restrict__ Pz,
                                                  const int32_t *__restrict__ Einv,
                                                  double *__restrict__ z,
                                                  const LsqrState *__restrict__ st) {
    if (st && st->done) return;
    __shared__ double sh[RED_GRP][RED_OUT];
    const int e = threadIdx.x % RED_OUT, g = threadIdx.x / RED_OUT;
    const long j = (long)blockIdx.x * RED_OUT + e;
    double s = 0.0;
    if (j < d) {
        const long i = Einv ? (long)Einv[j] : j;
        if (i >= 0 && i < r)
            s = ordered_sum(Pz + i, r, g, nck, RED_GRP);
    }
    sh[g][e] = s;
    __syncthreads();
    if (g == 0 && j < d) z[j] = (sh[0][e] + sh[1][e]) + (sh[2][e] + sh[3][e]);
}

// ut[i] = sum_c Pn[c * n + i] - coef * ut[i];  Sb[block] = sum of squares
constexpr int RN_PER = 256; // rows per block, one per thread
__global__ __launch_bounds__(256) void k_reduce_n(long n, int nck, const double *__restrict__ Pn,
                                                  double *__restrict__ ut, double *__restrict__ Sb,
                                                  LsqrState *__restrict__ st, int init) {
    if (st->done) return;
    const double coef = init ? 0.0 : st->coef;
    double sq = 0.0;
    for (int q = 0; q < RN_PER / 256; ++q) {
        long i = (long)blockIdx.x * RN_PER + q * 256 + threadIdx.x;
        if (i < n) {
            double s = 0.0;
            s = ordered_sum(Pn + i, n, 0, nck, 1);
            double v = s - coef * ut[i];
            ut[i] = v;
            sq += v * v;
        }
    }
    sq = block_sum<256>(sq);
    if (threadIdx.x == 0) Sb[blockIdx.x] = sq;
}

// squares of b into Sb (init; one partial per block of rpb rows), ut = b
__global__ __launch_bounds__(256) void k_init_u(long n, int rpb, const double *__restrict__ b,
                                                double *__restrict__ ut, double *__restrict__ Sb) {
    double sq = 0.0;
    const long i0 = (long)blockIdx.x * rpb, i1 = min(n, i0 + rpb);
    for (long i = i0 + threadIdx.x; i < i1; i += 256) {
        {
            double v = b[i];
            ut[i] = v;
            sq += v * v;
        }
    }
    sq = block_sum<256>(sq);
    if (threadIdx.x == 0) Sb[blockIdx.x] = sq;
}

// ||u|| from the nsb partial sums of squares, summed by the whole block (BS
// threads) in a fixed order; the result is valid in every thread
template <int BS> __device__ inline double beta_from(const double *Sb, int nsb) {
    double s = 0.0;
    for (int q = threadIdx.x; q < nsb; q += BS) s += Sb[q];
    return sqrt(block_sum<BS>(s));
}

// t[i] = (sum_rb Pt[rb * d + E[i]]) / beta
__global__ __launch_bounds__(256) void k_reduce_t(long d, long r, int nrb,
                                                  const double *__restrict__ Pt,
                                                  const int32_t *__restrict__ E,
                                                  const double *__restrict__ Sb, int nsb,
                                                  double *__restrict__ t,
                                                  const LsqrState *__restrict__ st) {
    if (st->done) return;
    __shared__ double sh[RED_GRP][RED_OUT];
    const double sbeta = beta_from<256>(Sb, nsb);
    const int e = threadIdx.x % RED_OUT, g = threadIdx.x / RED_OUT;
    const long i = (long)blockIdx.x * RED_OUT + e;
    double s = 0.0;
    if (i < r) {
        const long j = E ? (long)E[i] : i;
        s = ordered_sum(Pt + j, d, g, nrb, RED_GRP);
    }
    sh[g][e] = s;
    __syncthreads();
    if (g == 0 && i < r) t[i] = ((sh[0][e] + sh[1][e]) + (sh[2][e] + sh[3][e])) * (1.0 / sbeta);
}

// The same over many partials (the one-pass kernel's, one per workgroup): 16
// outputs per block, each summed by 16 threads over every 16th partial, then
// combined in a fixed order.
constexpr int RW_OUT = 16, RW_GRP = 256 / RW_OUT;
__global__ __launch_bounds__(256) void k_reduce_t_wide(long d, long r, int np,
                                                       const double *__restrict__ P,
                                                       const int32_t *__restrict__ E,
                                                       const double *__restrict__ Sb, int nsb,
                                                       double *__restrict__ t,
                                                       const LsqrState *__restrict__ st) {
    if (st->done) return;
    __shared__ double sh[RW_GRP][RW_OUT + 1];
    const int e = threadIdx.x % RW_OUT, g = threadIdx.x / RW_OUT;
    const long i = (long)blockIdx.x * RW_OUT + e;
    double s = 0.0;
    if (i < r) {
        const long j = E ? (long)E[i] : i;
        s = ordered_sum(P + j, d, g, np, RW_GRP);
    }
    sh[g][e] = s;
    // beta after the partial sums (their loads in flight meanwhile)
    const double sbeta = beta_from<256>(Sb, nsb);
    if (g == 0 && i < r) {
        double tot = 0.0;
#pragma unroll
        for (int k = 0; k < RW_GRP; ++k) tot += sh[k][e];
        t[i] = tot * (1.0 / sbeta);
    }
}

// ---- scalar recurrences (one workgroup) ---------------------------------------
constexpr int UPD_THREADS = 1024;

__global__ __launch_bounds__(UPD_THREADS) void k_update(
    long r, int nrbm, const double *__restrict__ Pm, const double *__restrict__ Sb, int nsb,
    double *__restrict__ v, double *__restrict__ w, double *__restrict__ x,
    LsqrState *__restrict__ st, double *__restrict__ hist, long hist_cap, int init) {
    if (st->done) return;
    __shared__ double sh[4];
    const double beta = beta_from<UPD_THREADS>(Sb, nsb);
    double alpha_old = init ? 0.0 : st->alpha;
    double alpha = alpha_old;
    double nrm_a = init ? 0.0 : st->nrm_a;
    bool upd = beta > 0;
    if (upd) {
        // v = A^T u - beta v ;  alpha = ||v|| ; v /= alpha
        double sq = 0.0;
        for (long j = threadIdx.x; j < r; j += UPD_THREADS) {
            double s = 0.0;
            s = ordered_sum(Pm + j, r, 0, nrbm, 1);
            double vv = init ? s : ((-beta) * v[j] + s);
            v[j] = vv;
            sq += vv * vv;
        }
        sq = block_sum<UPD_THREADS>(sq);
        alpha = sqrt(sq);
        if (!init) nrm_a = sqrt(nrm_a * nrm_a + alpha_old * alpha_old + beta * beta);
        if (alpha > 0) {
            double sc = 1.0 / alpha;
            for (long j = threadIdx.x; j < r; j += UPD_THREADS) v[j] = sc * v[j];
        }
    } else if (init) {
        alpha = 0.0;
        for (long j = threadIdx.x; j < r; j += UPD_THREADS) v[j] = 0.0;
    }
    __syncthreads();
    if (init) {
        for (long j = threadIdx.x; j < r; j += UPD_THREADS) {
            w[j] = v[j];
            x[j] = 0.0;
        }
        if (threadIdx.x == 0) {
            st->alpha = alpha;
            st->beta = beta;
            st->rhobar = alpha;
            st->phibar = beta;
            st->nrm_a = 0.0;
            st->nrm_r = beta;
            st->nrm_ar = alpha * beta;
            st->coef = (beta > 0) ? alpha / beta : 0.0;
            st->it = 0;
            if (alpha * beta == 0) {
                st->flag = 0;
                st->done = 1;
            } else {
                st->flag = 1;
                st->done = (st->max_it <= 0) ? 1 : 0;
            }
        }
        return;
    }
    double rhobar = st->rhobar, phibar = st->phibar;
    double rho = sqrt(rhobar * rhobar + beta * beta);
    double cs = rhobar / rho;
    double sn = beta / rho;
    double theta = sn * alpha;
    rhobar = -cs * alpha;
    double phi = cs * phibar;
    phibar = sn * phibar;
    double c1 = phi / rho, c2 = -theta / rho;
    for (long j = threadIdx.x; j < r; j += UPD_THREADS) {
        double wj = w[j];
        x[j] = x[j] + c1 * wj;
        w[j] = c2 * wj + 1.0 * v[j];
    }
    if (threadIdx.x == 0) {
        long it = st->it + 1;
        double nrm_r = phibar;
        double nrm_ar = phibar * alpha * fabs(cs);
        st->alpha = alpha;
        st->beta = beta;
        st->rhobar = rhobar;
        st->phibar = phibar;
        st->nrm_a = nrm_a;
        st->nrm_r = nrm_r;
        st->nrm_ar = nrm_ar;
        st->coef = (beta > 0) ? alpha / beta : 0.0;
        st->it = it;
        if (hist && it <= hist_cap) {
            double *h = hist + 5 * (it - 1);
            h[0] = alpha;
            h[1] = beta;
            h[2] = rhobar;
            h[3] = phibar;
            h[4] = nrm_a;
        }
        if (nrm_ar < st->tol * nrm_a * nrm_r) {
            st->flag = 0;
            st->done = 1;
        } else if (it >= st->max_it) {
            st->flag = 1;
            st->done = 1;
        }
    }
}

// The same iteration step (init = 0) over ceil(r / 256) workgroups in two
// kernels.  k_update_v: v = A^T u - beta v (unnormalised) and one partial sum
// of squares per workgroup in Sv.  k_update_wx: every workgroup forms alpha
// from Sv and the Givens scalars from the state (the same values everywhere),
// normalises its part of v and updates w and x; the last workgroup to finish
// (a counter in device memory, reset by that workgroup) writes the new state,
// when every other workgroup has read the old one.
constexpr int UV_T = 256;

__global__ __launch_bounds__(UV_T) void k_update_v(long r, int nrbm, const double *__restrict__ Pm,
                                                   const double *__restrict__ Sb, int nsb,
                                                   double *__restrict__ v,
                                                   double *__restrict__ Sv,
                                                   const LsqrState *__restrict__ st) {
    if (st->done) return;
    const double beta = beta_from<UV_T>(Sb, nsb);
    const long j = (long)blockIdx.x * UV_T + threadIdx.x;
    double sq = 0.0;
    if (beta > 0 && j < r) {
        double s = 0.0;
        s = ordered_sum(Pm + j, r, 0, nrbm, 1);
        const double vv = (-beta) * v[j] + s;
        v[j] = vv;
        sq = vv * vv;
    }
    sq = block_sum<UV_T>(sq);
    if (threadIdx.x == 0) Sv[blockIdx.x] = sq;
}

__global__ __launch_bounds__(UV_T) void k_update_wx(long r, int nsv, const double *__restrict__ Sv,
                                                    const double *__restrict__ Sb, int nsb,
                                                    double *__restrict__ v, double *__restrict__ w,
                                                    double *__restrict__ x,
                                                    LsqrState *__restrict__ st, int *cnt,
                                                    double *__restrict__ hist, long hist_cap) {
    if (st->done) return;
    // this thread's entries first, so that their loads overlap the sums
    const long j = (long)blockIdx.x * UV_T + threadIdx.x;
    double vj = 0.0, wj = 0.0, xj = 0.0;
    if (j < r) {
        vj = v[j];
        wj = w[j];
        xj = x[j];
    }
    const double beta = beta_from<UV_T>(Sb, nsb);
    const double alpha_old = st->alpha;
    double alpha = alpha_old, nrm_a = st->nrm_a;
    const bool upd = beta > 0;
    if (upd) {
        double s = 0.0;
        for (int q = threadIdx.x; q < nsv; q += UV_T) s += Sv[q];
        alpha = sqrt(block_sum<UV_T>(s));
        nrm_a = sqrt(nrm_a * nrm_a + alpha_old * alpha_old + beta * beta);
    }
    const bool scale = upd && alpha > 0;
    const double sc = scale ? 1.0 / alpha : 1.0;
    double rhobar = st->rhobar, phibar = st->phibar;
    const double rho = sqrt(rhobar * rhobar + beta * beta);
    const double cs = rhobar / rho;
    const double sn = beta / rho;
    const double theta = sn * alpha;
    rhobar = -cs * alpha;
    const double phi = cs * phibar;
    phibar = sn * phibar;
    const double c1 = phi / rho, c2 = -theta / rho;
    if (j < r) {
        if (scale) {
            vj = sc * vj;
            v[j] = vj;
        }
        x[j] = xj + c1 * wj;
        w[j] = c2 * wj + 1.0 * vj;
    }
    __shared__ int last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(cnt, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    *cnt = 0;
    const long it = st->it + 1;
    const double nrm_r = phibar;
    const double nrm_ar = phibar * alpha * fabs(cs);
    st->alpha = alpha;
    st->beta = beta;
    st->rhobar = rhobar;
    st->phibar = phibar;
    st->nrm_a = nrm_a;
    st->nrm_r = nrm_r;
    st->nrm_ar = nrm_ar;
    st->coef = (beta > 0) ? alpha / beta : 0.0;
    st->it = it;
    if (hist && it <= hist_cap) {
        double *h = hist + 5 * (it - 1);
        h[0] = alpha;
        h[1] = beta;
        h[2] = rhobar;
        h[3] = phibar;
        h[4] = nrm_a;
    }
    if (nrm_ar < st->tol * nrm_a * nrm_r) {
        st->flag = 0;
        st->done = 1;
    } else if (it >= st->max_it) {
        st->flag = 1;
        st->done = 1;
    }
}

// ---- persistent LSQR (small problems) ----------------------------------------------
// For problems whose preconditioned operator B = A(:, E) M (n x r) fits in
// the registers of at most half the CUs (r <= 256, at most 128 rows per
// workgroup), every iteration of a solve runs in one launch of P workgroups
// with ONE grid barrier per iteration:
//   each workgroup holds its rows of B in registers (lane = row, wave w =
//   columns w, w + 16, ...) and its rows of ut; before the barrier it forms
//   ut = B v - coef ut for its rows, their sum of squares (Sb) and its
//   partial B^T ut over its rows (Pt); after the barrier every workgroup sums
//   the P partials in the same order and so holds the same beta, t = B^T u,
//   v' = t - beta v, alpha = ||v'||, the Givens step and x, w, v (kept in
//   LDS, identical bits in every workgroup; workgroup 0 writes x and the
//   state at the end).  Pt and Sb alternate between two halves by iteration
//   parity: a workgroup writing the next iteration's partials never touches
//   those that a slower one may still read.
// B is formed once per solve (k_precond_op), so the iteration reads no M and,
// beyond the P x r partials, no global memory.  The recurrences, the stopping
// test and the history are those of the launch-per-kernel iteration; the
// operator is applied as one product with B in place of M then A, which moves
// the rounding by about one unit of B's entries.
constexpr int PS_T = 1024, PS_W = PS_T / 64, PS_KPW = 16, PS_RMAX = PS_W * PS_KPW;
struct PersistArgs {
    long n, r, rb; // rows, rank, rows per workgroup (<= 64 * RCH)
    const double *Bm;
    long ldb;
    double *ut, *v, *w, *x, *Pt, *Sb, *hist;
    long hist_cap;
    LsqrState *st;
    GridSync bar;
};

// Sum over the 64 lanes of each of 16 values; lane l returns the sum of value
// (l >> 2) & 15.  Halving exchanges: a lane keeps the half of its values that
// its bit selects and adds the partner's copy of that half; the partners are
// lane ^ 32 and lane ^ 16 (LDS permutes), then lane mirrored within its 16
// and within its 8 (DPP; both differ in the bit of the step and agree in the
// bits already used), then two butterfly steps over the quad.  The order is
// fixed, so every call sums the same way.
template <int H, int BIT, int CTRL>
__device__ inline void halve_step(double (&p)[PS_KPW], int lane) {
    const bool up = (lane & BIT) != 0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
        const double keep = up ? p[i + H] : p[i];
        const double send = up ? p[i] : p[i + H];
        p[i] = keep + (CTRL < 0 ? __shfl_xor(send, BIT, 64) : dpp_move<(CTRL < 0 ? 0 : CTRL)>(send));
    }
}
__device__ inline double wave_sum16(double (&p)[PS_KPW], int lane) {
    static_assert(PS_KPW == 16, "wave_sum16 reduces 16 values");
    halve_step<8, 32, -1>(p, lane);
    halve_step<4, 16, -1>(p, lane);
    halve_step<2, 8, DPP_MIRROR>(p, lane);
    halve_step<1, 4, DPP_HALF_MIRROR>(p, lane);
    double s = p[0];
    s += dpp_move<DPP_XOR2>(s);
    s += dpp_move<DPP_XOR1>(s);
    return s;
}

template <int RCH>
__global__ __launch_bounds__(PS_T) void k_lsqr_persist(PersistArgs a) {
    __shared__ double vs[PS_RMAX], wl[PS_RMAX], xl[PS_RMAX], tl[4][PS_RMAX];
    __shared__ double red[RCH][PS_W][64], wred[2][PS_W];
    const int P = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
    const int wv = tid >> 6, lane = tid & 63;
    const long r = a.r;
    const LsqrState s0 = *a.st;
    if (s0.done) return;
    // this workgroup's rows of B and ut
    double bb[RCH][PS_KPW], u[RCH];
    bool ok[RCH];
#pragma unroll
    for (int ch = 0; ch < RCH; ++ch) {
        const long lr = (long)ch * 64 + lane, i = (long)g * a.rb + lr;
        ok[ch] = lr < a.rb && i < a.n;
#pragma unroll
        for (int c = 0; c < PS_KPW; ++c) {
            const long k = wv + (long)PS_W * c;
            bb[ch][c] = (ok[ch] && k < r) ? a.Bm[i + k * a.ldb] : 0.0;
        }
        u[ch] = ok[ch] ? a.ut[i] : 0.0;
    }
    for (int k = tid; k < PS_RMAX; k += PS_T) {
        vs[k] = k < r ? a.v[k] : 0.0;
        wl[k] = k < r ? a.w[k] : 0.0;
        xl[k] = k < r ? a.x[k] : 0.0;
    }
    __syncthreads();
    double alpha = s0.alpha, beta = s0.beta, rhobar = s0.rhobar, phibar = s0.phibar;
    double nrm_a = s0.nrm_a, nrm_r = s0.nrm_r, nrm_ar = s0.nrm_ar, coef = s0.coef;
    long it = s0.it;
    int flag = 1;
    unsigned target = 0;
    for (int par = 0;; par ^= 1) {
        double *Pt = a.Pt + (size_t)par * P * r;
        double *Sb = a.Sb + (size_t)par * P;
        // ut = B v - coef ut over this workgroup's rows
#pragma unroll
        for (int ch = 0; ch < RCH; ++ch) {
            double f = 0.0;
#pragma unroll
            for (int c = 0; c < PS_KPW; ++c) f += bb[ch][c] * vs[wv + PS_W * c];
            red[ch][wv][lane] = f;
        }
        __syncthreads();
        double sq = 0.0;
#pragma unroll
        for (int ch = 0; ch < RCH; ++ch) {
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < PS_W; ++q) s += red[ch][q][lane];
            u[ch] = ok[ch] ? s - coef * u[ch] : 0.0;
            sq += u[ch] * u[ch];
        }
        // partial B^T ut over the rows, one column per four lanes
        double p[PS_KPW];
#pragma unroll
        for (int c = 0; c < PS_KPW; ++c) {
            double s = 0.0;
#pragma unroll
            for (int ch = 0; ch < RCH; ++ch) s += bb[ch][c] * u[ch];
            p[c] = s;
        }
        const double pc = wave_sum16(p, lane);
        const long kc = wv + (long)PS_W * ((lane >> 2) & 15);
        if ((lane & 3) == 0 && kc < r) Pt[(long)g * r + kc] = pc;
        if (wv == 0) {
            sq = wave_sum_bcast(sq);
            if (lane == 0) Sb[g] = sq;
        }
        if (!grid_barrier(a.bar, target)) return;
        // every workgroup: beta, t = B^T u, v', alpha (same order everywhere);
        // the partial sums of squares load in the same round as those of B^T ut
        const double sbv = tid < P ? Sb[tid] : 0.0;
        {
            const int k = tid & (PS_RMAX - 1), h = tid / PS_RMAX;
            double s = 0.0;
            if (k < r) {
                const double *pt = Pt + k;
                const int rr = (int)r;
#pragma unroll 8
                for (int q = h; q < P; q += PS_T / PS_RMAX) s += pt[q * rr];
            }
            tl[h][k] = s;
        }
        {
            const double ws = wave_sum_bcast(sbv);
            if (lane == 0) wred[0][wv] = ws;
        }
        __syncthreads();
        double sb = 0.0;
#pragma unroll
        for (int q = 0; q < PS_W; ++q) sb += wred[0][q];
        beta = sqrt(sb);
        const bool upd = beta > 0;
        double vp = 0.0, sv = 0.0;
        if (tid < r) {
            const double t = ((tl[0][tid] + tl[1][tid]) + (tl[2][tid] + tl[3][tid])) / beta;
            vp = upd ? t - beta * vs[tid] : vs[tid];
            sv = upd ? vp * vp : 0.0;
        }
        {
            const double ws = wave_sum_bcast(sv);
            if (lane == 0) wred[1][wv] = ws;
        }
        __syncthreads();
        sv = 0.0;
#pragma unroll
        for (int q = 0; q < PS_W; ++q) sv += wred[1][q];
        const double alpha_old = alpha;
        if (upd) {
            alpha = sqrt(sv);
            nrm_a = sqrt(nrm_a * nrm_a + alpha_old * alpha_old + beta * beta);
        }
        const double sc = (upd && alpha > 0) ? 1.0 / alpha : 1.0;
        const double rho = sqrt(rhobar * rhobar + beta * beta);
        const double cs = rhobar / rho;
        const double sn = beta / rho;
        const double theta = sn * alpha;
        rhobar = -cs * alpha;
        const double phi = cs * phibar;
        phibar = sn * phibar;
        const double c1 = phi / rho, c2 = -theta / rho;
        if (tid < r) {
            const double vn = sc * vp, wj = wl[tid];
            vs[tid] = vn;
            xl[tid] = xl[tid] + c1 * wj;
            wl[tid] = c2 * wj + 1.0 * vn;
        }
        ++it;
        nrm_r = phibar;
        nrm_ar = phibar * alpha * fabs(cs);
        coef = upd ? alpha / beta : 0.0;
        if (g == 0 && tid == 0 && a.hist && it <= a.hist_cap) {
            double *hh = a.hist + 5 * (it - 1);
            hh[0] = alpha;
            hh[1] = beta;
            hh[2] = rhobar;
            hh[3] = phibar;
            hh[4] = nrm_a;
        }
        bool done = false;
        if (nrm_ar < s0.tol * nrm_a * nrm_r) {
            flag = 0;
            done = true;
        } else if (it >= s0.max_it) {
            flag = 1;
            done = true;
        }
        __syncthreads(); // v, x, w complete before the next product reads them
        if (done) break;
    }
    if (g != 0) return;
    for (long k = tid; k < r; k += PS_T) {
        a.x[k] = xl[k];
        a.v[k] = vs[k];
        a.w[k] = wl[k];
    }
    if (tid == 0) {
        LsqrState *st = a.st;
        st->alpha = alpha;
        st->beta = beta;
        st->rhobar = rhobar;
        st->phibar = phibar;
        st->nrm_a = nrm_a;
        st->nrm_r = nrm_r;
        st->nrm_ar = nrm_ar;
        st->coef = coef;
        st->it = it;
        st->flag = flag;
        st->done = 1;
    }
}

// Bm = A(:, E) M (n x r, column-major, ld n) with M upper triangular (its
// entries below the diagonal are not read); E null is the identity.  64 x 64
// output tiles, 16-deep steps of A's columns and M's rows through LDS.
__global__ __launch_bounds__(256) void k_precond_op(long n, long r, const double *__restrict__ A,
                                                    long lda, const int32_t *__restrict__ E,
                                                    const double *__restrict__ M, long ldm,
                                                    double *__restrict__ Bm, long ldb) {
    __shared__ double As[16][64], Ms[16][65];
    const long i0 = (long)blockIdx.x * 64, k0 = (long)blockIdx.y * 64;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[4][4] = {};
    const long jend = min(r, k0 + 64); // M[j, k] = 0 for j > k
    for (long j0 = 0; j0 < jend; j0 += 16) {
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            const int jj = e >> 6, ii = e & 63;
            const long j = j0 + jj, i = i0 + ii, k = k0 + ii;
            As[jj][ii] = (j < r && i < n) ? A[i + (E ? (long)E[j] : j) * lda] : 0.0;
            Ms[jj][ii] = (j < r && k < r && j <= k) ? M[j + k * ldm] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            double av[4], mv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = As[jj][tx + 16 * q];
                mv[q] = Ms[jj][ty + 16 * q];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int s = 0; s < 4; ++s) acc[q][s] = fma(av[q], mv[s], acc[q][s]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const long i = i0 + tx + 16 * q, k = k0 + ty + 16 * s;
            if (i < n && k < r) Bm[i + k * ldb] = acc[q][s];
        }
}

// D (n x r, zeroed by the caller) = the columns E of CSR A, compacted:
// D[i, Einv[j]] = A[i, j] for the columns j that E selects (Einv null: all).
// Duplicate entries add up in CSR order.  One thread per row.
__global__ void k_csr_to_dense(long n, const int32_t *__restrict__ rp,
                               const int32_t *__restrict__ cc, const double *__restrict__ cv,
                               const int32_t *__restrict__ Einv, double *__restrict__ D) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int q = rp[i]; q < rp[i + 1]; ++q) {
        const long k = Einv ? (long)Einv[cc[q]] : (long)cc[q];
        if (k >= 0) D[i + k * n] += cv[q];
    }
}

// ---- sparse operator kernels ----------------------------------------------------
// ut[i] = sum_p val[p] z[col[p]] - coef ut[i] over CSR rows, LPR lanes per
// row; Sb per block of SP_ROWS rows (enough blocks to cover every CU several
// times over at the sparse shapes).
constexpr int SP_ROWS = 128;
template <int LPR>
__global__ __launch_bounds__(256) void k_spmv_n(long n, const int32_t *__restrict__ ptr,
                                                const int32_t *__restrict__ col,
                                                const double *__restrict__ val,
                                                const double *__restrict__ z,
                                                double *__restrict__ ut, double *__restrict__ Sb,
                                                LsqrState *__restrict__ st, int init) {
    if (st->done) return;
    const double coef = init ? 0.0 : st->coef;
    const int sub = threadIdx.x % LPR, grp = threadIdx.x / LPR;
    constexpr int RPI = 256 / LPR; // rows per pass
    double sq = 0.0;
    for (int q = 0; q < SP_ROWS / RPI; ++q) {
        long i = (long)blockIdx.x * SP_ROWS + q * RPI + grp;
        double s = 0.0;
        if (i < n) {
            int p0 = ptr[i], p1 = ptr[i + 1];
            for (int p = p0 + sub; p < p1; p += LPR) s += val[p] * z[col[p]];
        }
#pragma unroll
        for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (sub == 0 && i < n) {
            double vv = s - coef * ut[i];
            ut[i] = vv;
            sq += vv * vv;
        }
    }
    sq = block_sum<256>(sq);
    if (threadIdx.x == 0) Sb[blockIdx.x] = sq;
}

// The same product with z staged in LDS: a persistent grid of workgroups,
// each copies z (d values) once and then serves its rows' gathers from LDS
// instead of the cache hierarchy; Sb keeps one partial per SP_ROWS rows.
template <int LPR>
__global__ __launch_bounds__(256) void k_spmv_n_lds(long n, long d, int nrb,
                                                    const int32_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ col,
                                                    const double *__restrict__ val,
                                                    const double *__restrict__ z,
                                                    double *__restrict__ ut,
                                                    double *__restrict__ Sb,
                                                    LsqrState *__restrict__ st, int init) {
    if (st->done) return;
    extern __shared__ double zs[];
    for (long j = threadIdx.x; j < d; j += 256) zs[j] = z[j];
    __syncthreads();
    const double coef = init ? 0.0 : st->coef;
    const int sub = threadIdx.x % LPR, grp = threadIdx.x / LPR;
    constexpr int RPI = 256 / LPR; // rows per pass
    for (int rb = blockIdx.x; rb < nrb; rb += gridDim.x) {
        double sq = 0.0;
        for (int q = 0; q < SP_ROWS / RPI; ++q) {
            long i = (long)rb * SP_ROWS + q * RPI + grp;
            double s = 0.0;
            if (i < n) {
                // the lane's entries k = 0, 1, ... (p0 + sub + k LPR) eight at
                // a time, all loads issued first: even k into s, odd k into
                // s1, each in increasing k; past the row's end the entry
                // repeats and is multiplied by 0
                const int p0 = ptr[i], p1 = ptr[i + 1];
                double s1 = 0.0;
                for (int p = p0 + sub; p < p1; p += 8 * LPR) {
                    int cc[8];
                    double vv[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int pk = p + k * LPR, pc = pk < p1 ? pk : p;
                        cc[k] = col[pc];
                        vv[k] = val[pc];
                    }
#pragma unroll
                    for (int k = 0; k < 8; k += 2) {
                        const double m0 = p + k * LPR < p1 ? 1.0 : 0.0;
                        const double m1 = p + (k + 1) * LPR < p1 ? 1.0 : 0.0;
                        s = fma(vv[k] * m0, zs[cc[k]], s);
                        s1 = fma(vv[k + 1] * m1, zs[cc[k + 1]], s1);
                    }
                }
                s += s1;
            }
#pragma unroll
            for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
            if (sub == 0 && i < n) {
                double vv = s - coef * ut[i];
                ut[i] = vv;
                sq += vv * vv;
            }
        }
        sq = block_sum<256>(sq);
        if (threadIdx.x == 0) Sb[rb] = sq;
    }
}

// Pt[j] = sum_p val[p] ut[row[p]] over CSC column j (one wave per column)
__global__ __launch_bounds__(256) void k_csc_t(long d, const int32_t *__restrict__ ptr,
                                               const int32_t *__restrict__ row,
                                               const double *__restrict__ val,
                                               const double *__restrict__ ut,
                                               double *__restrict__ Pt,
                                               const LsqrState *__restrict__ st) {
    if (st && st->done) return;
    const int lane = threadIdx.x & 63;
    long j = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= d) return;
    const int p0 = ptr[j], p1 = ptr[j + 1];
    // one running sum (four independent ones measured slower: 32 vs 27 us at
    // 120000 x 5000, density 1%; so was this loop with four entries' loads
    // issued together, 33.5 us)
    double s = 0.0;
    for (int p = p0 + lane; p < p1; p += 64) s += val[p] * ut[row[p]];
    s = wave_sum(s);
    if (lane == 0) Pt[j] = s;
}

// ---- host driver ----------------------------------------------------------------
struct LsqrPlan {
    long n, d, r;
    int cw_a, nck_a;   // A z chunks
    int cw_m, nck_m;   // M v chunks
    int nrb_a, cpw_a;  // A^T u row blocks / columns per workgroup
    int nrb_m, cpw_m;  // M^T t
    int nsb;           // squares partials
    bool sparse;
    bool fused = false; // one pass over a row-major copy of A per iteration (k_fused_rows)
    int spg = 0;        // sparse A z with z in LDS: workgroups (0: z from memory)
    int G = 0, nc = 0;  // fused: workgroups, column pairs per thread (over 512 columns)
    long ldt = 0;       // fused: row stride of the copy
};

// the one-pass kernel holds up to 16 column pairs of a row per thread in
// registers (d <= 8192); small problems keep the two-pass kernels
bool lsqr_rows_wanted(long n, long d) {
    if (d > 16 * FR_COLS || n < 4096) return false;
    const char *e = getenv("SKL_LSQR_FUSED");
    return !(e && e[0] == '0');
}
static bool fused_wanted(const LsqrProblem &p) {
    return !p.csr_ptr && !ctx().no_alloc && lsqr_rows_wanted(p.n, p.d);
}

long lsqr_rows_ld(long d) { return d + (d & 1); }

void lsqr_make_rows(long n, long d, const double *A, long lda, double *AT, hipStream_t s) {
    const long ldt = lsqr_rows_ld(d);
    dim3 g((unsigned)((n + 63) / 64), (unsigned)((ldt + 63) / 64));
    k_to_rows<<<g, 256, 0, s>>>(n, d, A, lda, AT, ldt);
    SKL_CHECK_LAUNCH();
}

// the row-major copy doubles A's footprint: only when the device has room for
// it (free memory plus what the arena holds unused).  Under a memory
// allowance (a server slot) the admitted footprint, which counts the copy,
// decides instead of the device's free memory, which other slots change.
static bool rows_fit(const LsqrProblem &p, Arena &ar) {
    if (ctx().mem_allowance > 0)
        return 8.0 * (double)p.n * (double)lsqr_rows_ld(p.d) <= (double)ctx().mem_allowance;
    size_t fr = 0, tot = 0;
    SKL_HIP(hipMemGetInfo(&fr, &tot));
    const double need = 8.0 * (double)p.n * (double)(p.d + (p.d & 1));
    const double avail = (double)fr + (double)(ar.capacity() - ar.used());
    return need * 1.25 + 256.0 * (1 << 20) <= avail;
}

// ---- launch shapes (autotuned per problem-shape bucket) ----------------------
// target_mult: the workgroup count the two-pass products aim at, in multiples
// of the CU count; wpc: workgroups per CU of the one-pass kernel; spmv_lds:
// the sparse A z with z staged in LDS; persist: the workgroups of the
// persistent iteration (k_lsqr_persist), 0 for the launch-per-kernel one.
// Defaults: 8, 1, 1, and the persistent iteration with its smallest grid
// (ceil(n / 128) workgroups) for every problem that it takes.
struct LsqrShape {
    int target_mult = 8, wpc = 1, spmv_lds = 1;
    int persist = -1; // workgroups of the persistent iteration, 0 off, -1 by size
};
using ShapeKey = std::array<int, 4>;
static std::mutex g_shape_mu;
static std::map<ShapeKey, LsqrShape> g_shapes;
static thread_local const LsqrShape *t_shape_override = nullptr;

// buckets of half an octave in n, d and (sparse) the nonzeros per row
static ShapeKey shape_key(bool sparse, long n, long d, long nnz) {
    auto lg = [](double v) { return (int)std::floor(2.0 * std::log2(std::max(v, 1.0))); };
    return {sparse ? 1 : 0, lg((double)n), lg((double)d),
            sparse ? lg((double)nnz / (double)std::max(n, 1L)) : 0};
}

static LsqrShape shape_for(const LsqrProblem &p) {
    if (t_shape_override) return *t_shape_override;
    std::lock_guard<std::mutex> lk(g_shape_mu);
    auto it = g_shapes.find(shape_key(p.csr_ptr != nullptr, p.n, p.d, p.nnz));
    return it == g_shapes.end() ? LsqrShape() : it->second;
}

// The persistent iteration takes ranks up to 256 and at most 128 rows per
// workgroup on at most half the CUs; it runs with the smallest such grid
// unless the launch shape (or SKL_LSQR_PERSIST) asks for more workgroups.
static long persist_min_groups(const LsqrProblem &p) { return (p.n + 127) / 128; }
static bool persist_eligible(const LsqrProblem &p) {
    return p.r >= 1 && p.r <= PS_RMAX && p.n >= 1 &&
           persist_min_groups(p) <= std::max(1, ctx().num_cu / 2);
}
static int persist_groups(const LsqrProblem &p, const LsqrShape &sh) {
    if (!persist_eligible(p) || ctx().no_alloc) return 0;
    // SKL_LSQR_PERSIST=0 turns the persistent iteration off, =P (> 0) asks
    // for P workgroups on every problem that it takes
    const char *e = getenv("SKL_LSQR_PERSIST");
    const long env_p = e ? strtol(e, nullptr, 10) : -1;
    if (e && env_p <= 0) return 0;
    long P = env_p > 0 ? env_p : (long)sh.persist;
    if (P == 0) return 0;
    P = std::max(P, persist_min_groups(p));
    return (int)std::min<long>(P, std::max(1, ctx().num_cu / 2));
}

bool lsqr_shape_tuned(bool sparse, long n, long d, long nnz) {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    return g_shapes.count(shape_key(sparse, n, d, nnz)) != 0;
}

long lsqr_shape_entries() {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    return (long)g_shapes.size();
}

void lsqr_shape_clear() {
    std::lock_guard<std::mutex> lk(g_shape_mu);
    g_shapes.clear();
}

long long lsqr_shape_bucket(bool sparse, long n, long d, long nnz) {
    const ShapeKey k = shape_key(sparse, n, d, nnz);
    return ((long long)k[0] << 48) | ((long long)(k[1] & 0xffff) << 32) |
           ((long long)(k[2] & 0xffff) << 16) | (long long)(k[3] & 0xffff);
}

static LsqrPlan make_plan(const LsqrProblem &p, int num_cu) {
    LsqrPlan q;
    q.n = p.n;
    q.d = p.d;
    q.r = p.r;
    q.sparse = (p.csr_ptr != nullptr);
    const LsqrShape shape = shape_for(p);
    long rowblocks = (p.n + GN_ROWS - 1) / GN_ROWS;
    // enough workgroups to fill the chip several times over
    long target = (long)shape.target_mult * num_cu;
    long nck = std::max<long>(1, std::min<long>((target + rowblocks - 1) / rowblocks,
                                                (p.d + 63) / 64));
    q.cw_a = (int)std::min<long>(GN_MAXCW, (p.d + nck - 1) / nck);
    q.cw_a = std::max(q.cw_a, 1);
    q.nck_a = (int)((p.d + q.cw_a - 1) / q.cw_a);
    long rbm = (p.r + GN_ROWS - 1) / GN_ROWS;
    long nckm = std::max<long>(1, std::min<long>((target + rbm - 1) / rbm, (p.r + 63) / 64));
    q.cw_m = (int)std::min<long>(GN_MAXCW, (p.r + nckm - 1) / nckm);
    q.cw_m = std::max(q.cw_m, 1);
    q.nck_m = (int)((p.r + q.cw_m - 1) / q.cw_m);
    q.nrb_a = (int)((p.n + GT_RB - 1) / GT_RB);
    long cchunks = std::max<long>(1, (target + q.nrb_a - 1) / q.nrb_a);
    q.cpw_a = (int)std::max<long>(4, (p.d + cchunks - 1) / cchunks);
    q.nrb_m = (int)((p.r + GT_RB - 1) / GT_RB);
    long cchm = std::max<long>(1, (target + q.nrb_m - 1) / q.nrb_m);
    q.cpw_m = (int)std::max<long>(4, (p.r + cchm - 1) / cchm);
    q.nsb = (int)((p.n + (q.sparse ? SP_ROWS : RN_PER) - 1) / (q.sparse ? SP_ROWS : RN_PER));
    if (q.sparse) {
        q.nrb_a = 1;
        // z in LDS when it fits four workgroups per CU (or at least one)
        const size_t zb = sizeof(double) * (size_t)p.d;
        const int per_cu = zb <= 40960 ? 4 : zb <= 81920 ? 2 : zb <= 163840 ? 1 : 0;
        const char *e = getenv("SKL_SPMV_LDS");
        if (per_cu > 0 && shape.spmv_lds && !(e && e[0] == '0'))
            q.spg = std::min(q.nsb, per_cu * num_cu);
    }
    if (fused_wanted(p)) {
        q.fused = true;
        const long need = (p.d + FR_COLS - 1) / FR_COLS;
        q.nc = need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : need <= 8 ? 8 : 16;
        // one workgroup per CU (the registers allow four up to nc = 8, but one
        // keeps as many bytes in flight with the fewest partials: 0.210, 0.222
        // and 0.236 ms per pass at 40000 x 4000 for one, two and four)
        int wpc = shape.wpc;
        const char *e = getenv("SKL_FUSED_WPC");
        if (e && atoi(e) > 0) wpc = atoi(e);
        q.G = (int)std::min<long>((long)wpc * num_cu, p.n);
        q.ldt = p.d + (p.d & 1);
    }
    return q;
}

struct LsqrBufs {
    double *ut, *v, *w, *x, *z, *Pz, *Pn, *Sb, *Pt, *t, *Pm, *hist;
    double *AT;      // fused: row-major copy of A
    double *Sv;      // squares of v, one per k_update_v workgroup
    int *cnt;        // k_update_wx's finished-workgroup counter (zero between launches)
    long hist_cap;
    int32_t *Einv;
    LsqrState *st;
};

template <class T> struct Ident {
    typedef T type;
};

// Graph under construction on this thread: while `on`, launch_ev appends a
// kernel node (after the previous one) instead of launching.  The LSQR chunk
// graphs are built this way rather than by stream capture, which another
// thread's runtime calls can invalidate mid-recording (and the stream stays
// in the invalidated capture state afterwards).
struct GraphRec {
    bool on = false;
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;
    hipError_t err = hipSuccess;
    // argument values of the nodes, kept until the graph is instantiated
    std::vector<std::shared_ptr<void>> keep;
};
static thread_local GraphRec t_rec;

template <class Tup, size_t... I>
static void arg_ptrs(Tup &t, void **ptr, std::index_sequence<I...>) {
    ((ptr[I] = (void *)&std::get<I>(t)), ...);
}

// Launches k on s (or appends it to the graph being built).  With timing
// events (e0 non-null) they are recorded by the dispatch itself
// (hipExtLaunchKernelGGL), so their interval is the kernel's own execution
// and not the stream's event commands around it.
template <class... P>
static void launch_ev(void (*k)(P...), dim3 g, dim3 b, uint32_t sh, hipStream_t s, hipEvent_t e0,
                      hipEvent_t e1, typename Ident<P>::type... args) {
    if (t_rec.on) {
        if (t_rec.err != hipSuccess) return;
        using Tup = std::tuple<typename std::decay<P>::type...>;
        auto vals = std::make_shared<Tup>(args...);
        t_rec.keep.push_back(vals);
        void *ptr[sizeof...(P) > 0 ? sizeof...(P) : 1];
        arg_ptrs(*vals, ptr, std::index_sequence_for<P...>{});
        hipKernelNodeParams np{};
        np.func = (void *)k;
        np.gridDim = g;
        np.blockDim = b;
        np.sharedMemBytes = sh;
        np.kernelParams = ptr;
        np.extra = nullptr;
        hipGraphNode_t node = nullptr;
        t_rec.err = hipGraphAddKernelNode(&node, t_rec.graph, t_rec.last ? &t_rec.last : nullptr,
                                          t_rec.last ? 1 : 0, &np);
        if (t_rec.err == hipSuccess) t_rec.last = node;
        return;
    }
    if (e0)
        hipExtLaunchKernelGGL(k, g, b, sh, s, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(k, g, b, sh, s, args...);
}

// after_fused: A^T u already sits in Pt as per-workgroup partials and the
// squares in Sb (q.G of each); otherwise A^T u is computed here.
static void launch_op_t(const LsqrProblem &p, const LsqrPlan &q, const LsqrBufs &B,
                        hipStream_t s, int init, hipEvent_t *ev = nullptr,
                        bool after_fused = false) {
    hipEvent_t e2 = ev ? ev[2] : nullptr, e3 = ev ? ev[3] : nullptr;
    hipEvent_t e6 = ev ? ev[6] : nullptr, e7 = ev ? ev[7] : nullptr;
    int nrb = q.nrb_a, nsb = q.nsb;
    if (after_fused) {
        // t = Pt(E) / beta over the workgroup partials
        nsb = q.G;
        launch_ev(k_reduce_t_wide, dim3((unsigned)((q.r + RW_OUT - 1) / RW_OUT)), dim3(256), 0, s,
                  nullptr, nullptr, q.d, q.r, q.G, B.Pt, p.E, B.Sb, nsb, B.t, B.st);
    } else {
        // Pt = A^T ut
        if (!q.sparse) {
            dim3 g((unsigned)q.nrb_a, (unsigned)((q.d + q.cpw_a - 1) / q.cpw_a));
            launch_ev(k_gemv_t_part<0>, g, dim3(GT_THREADS), 0, s, e2, e3, q.n, q.d, p.A, p.lda,
                      B.ut, q.cpw_a, B.Pt, 0, B.st);
        } else {
            launch_ev(k_csc_t, dim3((unsigned)((q.d + 3) / 4)), dim3(256), 0, s, e2, e3, q.d,
                      p.csc_ptr, p.csc_row, p.csc_val, B.ut, B.Pt, B.st);
        }
        // t = Pt(E) / beta
        launch_ev(k_reduce_t, dim3((unsigned)((q.r + RED_OUT - 1) / RED_OUT)), dim3(256), 0, s,
                  nullptr, nullptr, q.d, q.r, nrb, B.Pt, p.E, B.Sb, nsb, B.t, B.st);
    }
    const unsigned nv = (unsigned)((q.r + UV_T - 1) / UV_T);
    if (init) {
        // Pm = M^T t, then the first step on one workgroup
        dim3 g((unsigned)q.nrb_m, (unsigned)((q.r + q.cpw_m - 1) / q.cpw_m));
        k_gemv_t_part<1><<<g, GT_THREADS, 0, s>>>(q.r, q.r, p.M, p.ldm, B.t, q.cpw_m, B.Pm, 1,
                                                   B.st);
        k_update<<<1, UPD_THREADS, 0, s>>>(q.r, q.nrb_m, B.Pm, B.Sb, nsb, B.v, B.w, B.x, B.st,
                                           B.hist, B.hist_cap, init);
    } else {
        // Pm = M^T t (row-block partials; those below the diagonal blocks stay zero)
        const unsigned g = (unsigned)trmv_t_chunks(q.r);
        launch_ev(k_trmv_t_chunks, dim3(g), dim3(256), 0, s, e6, e7, q.r, p.M, p.ldm, B.t, B.Pm,
                  q.nrb_m, B.st);
        launch_ev(k_update_v, dim3(nv), dim3(UV_T), 0, s, nullptr, nullptr, q.r, q.nrb_m, B.Pm,
                  B.Sb, nsb, B.v, B.Sv, B.st);
        launch_ev(k_update_wx, dim3(nv), dim3(UV_T), 0, s, nullptr, nullptr, q.r, (int)nv, B.Sv, B.Sb, nsb, B.v, B.w, B.x, B.st, B.cnt,
                                        B.hist, B.hist_cap);
    }
}

template <int NC> static void launch_fused_nc(const LsqrProblem &p, const LsqrPlan &q,
                                              const LsqrBufs &B, hipStream_t s, hipEvent_t e0,
                                              hipEvent_t e1) {
    const size_t sh = sizeof(double) * ((size_t)NC * FR_COLS + 2 * FR_W);
    launch_ev(k_fused_rows<NC>, dim3(q.G), dim3(FR_T), (uint32_t)sh, s, e0, e1, q.n, q.d, B.AT,
              q.ldt, B.z, B.ut, B.Sb, B.Pt, B.st);
}

// The one-pass kernel's dynamic LDS exceeds the default 64 KB at 16 column
// pairs per thread; the attribute is set before the first launch (never
// inside a graph capture).
template <int NC> static void fused_attr_nc() {
    const size_t sh = sizeof(double) * ((size_t)NC * FR_COLS + 2 * FR_W);
    if (sh > 65536)
        SKL_HIP(hipFuncSetAttribute((const void *)k_fused_rows<NC>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
}
static void fused_attr(int nc) {
    switch (nc) {
    case 1: fused_attr_nc<1>(); break;
    case 2: fused_attr_nc<2>(); break;
    case 4: fused_attr_nc<4>(); break;
    case 8: fused_attr_nc<8>(); break;
    default: fused_attr_nc<16>(); break;
    }
}

static void launch_fused(const LsqrProblem &p, const LsqrPlan &q, const LsqrBufs &B,
                         hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    switch (q.nc) {
    case 1: launch_fused_nc<1>(p, q, B, s, e0, e1); break;
    case 2: launch_fused_nc<2>(p, q, B, s, e0, e1); break;
    case 4: launch_fused_nc<4>(p, q, B, s, e0, e1); break;
    case 8: launch_fused_nc<8>(p, q, B, s, e0, e1); break;
    default: launch_fused_nc<16>(p, q, B, s, e0, e1); break;
    }
}

// Optional kernel timing: events around the four operator products of one
// iteration (slots: 0 A z, 1 A^T u, 2 M v, 3 M^T t), recorded on the stream
// that runs them.
struct ProfStats {
    double ms[4] = {0, 0, 0, 0};
    long cnt[4] = {0, 0, 0, 0};
    double bytes[4] = {0, 0, 0, 0};
};
static thread_local bool g_prof_on = false;
static thread_local ProfStats g_prof;
static thread_local hipEvent_t g_ev[8] = {};

static void prof_events() {
    if (g_ev[0]) return;
    AllocLock lk(capture_mutex());
    for (auto &e : g_ev) SKL_HIP(hipEventCreate(&e));
}

static inline hipEvent_t evi(hipEvent_t *ev, int i) { return ev ? ev[i] : nullptr; }

static void launch_iteration(const LsqrProblem &p, const LsqrPlan &q, const LsqrBufs &B,
                             hipStream_t s, hipEvent_t *ev = nullptr) {
    // z = scatter_E(M v)
    {
        dim3 g((unsigned)((q.r + GN_ROWS - 1) / GN_ROWS), (unsigned)q.nck_m);
        launch_ev(k_gemv_n_part<1>, g, dim3(GN_THREADS), 0, s, evi(ev, 4), evi(ev, 5), q.r, q.r,
                  p.M, p.ldm, B.v, q.cw_m, B.Pz, 1, B.st);
        launch_ev(k_z_reduce, dim3((unsigned)((q.d + RED_OUT - 1) / RED_OUT)), dim3(256), 0, s, nullptr, nullptr, q.d, q.r, q.nck_m, B.Pz, B.Einv,
                                                                 B.z, B.st);
    }
    // ut = A z - coef ut ; Sb
    if (q.fused) {
        launch_fused(p, q, B, s, evi(ev, 0), evi(ev, 1));
        launch_op_t(p, q, B, s, 0, ev, true);
        return;
    }
    if (!q.sparse) {
        dim3 g((unsigned)((q.n + GN_ROWS - 1) / GN_ROWS), (unsigned)q.nck_a);
        launch_ev(k_gemv_n_part<0>, g, dim3(GN_THREADS), 0, s, evi(ev, 0), evi(ev, 1), q.n, q.d,
                  p.A, p.lda, B.z, q.cw_a, B.Pn, 0, B.st);
        launch_ev(k_reduce_n, dim3((unsigned)q.nsb), dim3(256), 0, s, nullptr, nullptr, q.n,
                  q.nck_a, B.Pn, B.ut, B.Sb, B.st, 0);
    } else {
        if (q.spg > 0)
            launch_ev(k_spmv_n_lds<8>, dim3((unsigned)q.spg), dim3(256),
                      (uint32_t)(sizeof(double) * q.d), s, evi(ev, 0), evi(ev, 1), q.n, q.d, q.nsb,
                      p.csr_ptr, p.csr_col, p.csr_val, B.z, B.ut, B.Sb, B.st, 0);
        else
            launch_ev(k_spmv_n<8>, dim3((unsigned)q.nsb), dim3(256), 0, s, evi(ev, 0), evi(ev, 1),
                      q.n, p.csr_ptr, p.csr_col, p.csr_val, B.z, B.ut, B.Sb, B.st, 0);
    }
    launch_op_t(p, q, B, s, 0, ev);
}

static void prof_accumulate(const LsqrPlan &q, long nnz) {
    const int pairs[4][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}};
    double r = (double)q.r, n = (double)q.n, d = (double)q.d;
    double b[4];
    if (q.fused) {
        // one read of A, u read and written, z
        b[0] = 8.0 * (n * d + 2 * n + d);
        b[1] = 0.0;
    } else if (!q.sparse) {
        b[0] = 8.0 * (n * d + n + d);
        b[1] = 8.0 * (n * d + n + d);
    } else {
        b[0] = 12.0 * (double)nnz + 8.0 * (n + 1) + 8.0 * n + 8.0 * d;
        b[1] = 12.0 * (double)nnz + 8.0 * (d + 1) + 8.0 * n + 8.0 * d;
    }
    b[2] = 8.0 * (r * (r + 1) / 2 + 2 * r);
    b[3] = b[2];
    for (int k = 0; k < 4; ++k) {
        if (q.fused && k == 1) continue; // A^T u runs inside the fused pass
        float ms = 0;
        if (hipEventElapsedTime(&ms, g_ev[pairs[k][0]], g_ev[pairs[k][1]]) == hipSuccess) {
            g_prof.ms[k] += ms;
            g_prof.cnt[k] += 1;
            g_prof.bytes[k] = b[k];
        } else {
            (void)hipGetLastError(); // do not leave a sticky error for later checks
        }
    }
}

__global__ void k_build_einv(long d, long r, const int32_t *E, int32_t *Einv) {
    long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < d) Einv[j] = -1;
}
__global__ void k_build_einv2(long r, const int32_t *E, int32_t *Einv) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < r) Einv[E[i]] = (int32_t)i;
}
__global__ void k_set_state(LsqrState *st, double tol, long max_it, int *cnt) {
    *cnt = 0;
    st->alpha = st->beta = st->rhobar = st->phibar = 0.0;
    st->nrm_a = st->nrm_r = st->nrm_ar = st->coef = 0.0;
    st->tol = tol;
    st->it = 0;
    st->max_it = max_it;
    st->flag = 1;
    st->done = 0;
}

// Cached executable graphs, keyed by every pointer and size they bake in.
struct GraphKey {
    std::vector<long> k;
    bool operator<(const GraphKey &o) const { return k < o.k; }
};
static thread_local std::map<GraphKey, hipGraphExec_t> g_graphs;

void lsqr_release_graphs() {
    if (!g_graphs.empty()) {
        AllocLock lk(capture_mutex());
        for (auto &kv : g_graphs) (void)hipGraphExecDestroy(kv.second);
    }
    g_graphs.clear();
}

bool lsqr_start_rows(LsqrProblem &p, Arena &ar) {
    if (!fused_wanted(p) || !rows_fit(p, ar)) return false;
    Context &C = ctx();
    C.make_aux();
    const long ldt = p.d + (p.d & 1);
    double *AT = ar.alloc_n<double>((size_t)p.n * ldt);
    SKL_HIP(hipEventRecord(C.aux_ev[0], C.stream));
    SKL_HIP(hipStreamWaitEvent(C.aux, C.aux_ev[0], 0));
    dim3 g((unsigned)((p.n + 63) / 64), (unsigned)((ldt + 63) / 64));
    k_to_rows<<<g, 256, 0, C.aux>>>(p.n, p.d, p.A, p.lda, AT, ldt);
    SKL_CHECK_LAUNCH();
    SKL_HIP(hipEventRecord(C.aux_ev[1], C.aux));
    p.AT = AT;
    p.ldt = ldt;
    p.AT_ready = C.aux_ev[1];
    return true;
}

LsqrResult lsqr_solve(const LsqrProblem &p, const double *b, double tol, long max_it,
                      double *y, bool use_graph, int chunk, skl_lsqr_step *hist, long hist_len,
                      Arena &ar) {
    Context &C = ctx();
    hipStream_t s = C.stream;
    LsqrPlan q = make_plan(p, C.num_cu);
    if (q.fused && !p.AT && !rows_fit(p, ar)) q.fused = false;
    if (q.fused) fused_attr(q.nc);
    if (q.spg > 0 && sizeof(double) * q.d > 65536)
        SKL_HIP(hipFuncSetAttribute((const void *)k_spmv_n_lds<8>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(sizeof(double) * q.d)));
    LsqrBufs B;
    B.ut = ar.alloc_n<double>((size_t)q.n);
    B.v = ar.alloc_n<double>((size_t)q.r);
    B.w = ar.alloc_n<double>((size_t)q.r);
    B.x = y;
    B.z = ar.alloc_n<double>((size_t)q.d);
    B.Pz = ar.alloc_n<double>((size_t)q.nck_m * q.r);
    B.Pn = q.sparse ? nullptr : ar.alloc_n<double>((size_t)q.nck_a * q.n);
    B.Sb = ar.alloc_n<double>((size_t)std::max(q.nsb, q.G));
    B.Pt = ar.alloc_n<double>((size_t)std::max(q.nrb_a, q.G) * q.d);
    B.t = ar.alloc_n<double>((size_t)q.r);
    B.Pm = ar.alloc_n<double>((size_t)q.nrb_m * q.r);
    k_zero_d<<<256, 256, 0, s>>>((long)q.nrb_m * q.r, B.Pm);
    B.AT = nullptr;
    if (q.fused) {
        if (p.AT) {
            // made on the aux stream (lsqr_start_rows) with the same row stride
            SKL_REQUIRE(p.ldt == q.ldt, SKL_E_ARG, "row-major copy of A has the wrong stride");
            B.AT = const_cast<double *>(p.AT);
            if (p.AT_ready) SKL_HIP(hipStreamWaitEvent(s, p.AT_ready, 0));
        } else {
            B.AT = ar.alloc_n<double>((size_t)q.n * q.ldt);
            dim3 g((unsigned)((q.n + 63) / 64), (unsigned)((q.ldt + 63) / 64));
            k_to_rows<<<g, 256, 0, s>>>(q.n, q.d, p.A, p.lda, B.AT, q.ldt);
        }
    }
    B.st = ar.alloc_n<LsqrState>(1);
    B.Sv = ar.alloc_n<double>((size_t)(q.r + UV_T - 1) / UV_T);
    B.cnt = ar.alloc_n<int>(1);
    long hl = hist ? std::min(hist_len, max_it) : 0;
    B.hist = hl > 0 ? ar.alloc_n<double>((size_t)5 * hl) : nullptr;
    B.hist_cap = hl;
    B.Einv = nullptr;
    if (p.E) {
        B.Einv = ar.alloc_n<int32_t>((size_t)q.d);
        k_build_einv<<<(unsigned)((q.d + 255) / 256), 256, 0, s>>>(q.d, q.r, p.E, B.Einv);
        k_build_einv2<<<(unsigned)((q.r + 255) / 256), 256, 0, s>>>(q.r, p.E, B.Einv);
    }
    k_set_state<<<1, 1, 0, s>>>(B.st, tol, max_it, B.cnt);
    SKL_CHECK_LAUNCH();

    // init: ut = b, beta = ||b||, v = M^T A^T u / alpha ...
    k_init_u<<<(unsigned)q.nsb, 256, 0, s>>>(q.n, q.sparse ? SP_ROWS : RN_PER, b, B.ut, B.Sb);
    launch_op_t(p, q, B, s, 1);
    SKL_CHECK_LAUNCH();

    struct HostFlag {
        int done;
        long it;
    };
    HostFlag *hf = (HostFlag *)C.pinned;
    if (chunk <= 0) chunk = 16;
    // no graphs and no timing events in no-allocation mode
    bool graph_ok = use_graph && !C.no_alloc;
    const bool prof = g_prof_on && !C.no_alloc;
    hipEvent_t *ev = nullptr;
    if (prof) {
        prof_events();
        ev = g_ev;
    }
    hipGraphExec_t exec = nullptr;
    // with kernel timing on, the iterations run eagerly (events cannot be
    // timed inside a replayed graph) and the second iteration of every chunk
    // is timed, on a device already busy with the first
    if (prof) graph_ok = false;
    // small problems: every iteration in one persistent launch
    int pg = (prof || max_it <= 0) ? 0 : persist_groups(p, shape_for(p));
    // the barrier count of the launch stays within the 32-bit counter
    if ((double)max_it * pg >= 2.0e9) pg = 0;
    if (pg > 0) {
        graph_ok = false;
        // the preconditioned operator B = A(:, E) M, formed once
        double *Bm = ar.alloc_n<double>((size_t)q.n * q.r);
        const dim3 gb((unsigned)((q.n + 63) / 64), (unsigned)((q.r + 63) / 64));
        if (q.sparse) {
            double *D = ar.alloc_n<double>((size_t)q.n * q.r);
            k_zero_d<<<256, 256, 0, s>>>(q.n * q.r, D);
            k_csr_to_dense<<<(unsigned)((q.n + 255) / 256), 256, 0, s>>>(
                q.n, p.csr_ptr, p.csr_col, p.csr_val, B.Einv, D);
            k_precond_op<<<gb, 256, 0, s>>>(q.n, q.r, D, q.n, nullptr, p.M, p.ldm, Bm, q.n);
        } else {
            k_precond_op<<<gb, 256, 0, s>>>(q.n, q.r, p.A, p.lda, p.E, p.M, p.ldm, Bm, q.n);
        }
        SKL_CHECK_LAUNCH();
        PersistArgs pa;
        pa.n = q.n;
        pa.r = q.r;
        pa.rb = (q.n + pg - 1) / pg;
        pa.Bm = Bm;
        pa.ldb = q.n;
        pa.ut = B.ut;
        pa.v = B.v;
        pa.w = B.w;
        pa.x = B.x;
        pa.Pt = ar.alloc_n<double>((size_t)2 * pg * q.r);
        pa.Sb = ar.alloc_n<double>((size_t)2 * pg);
        pa.hist = B.hist;
        pa.hist_cap = B.hist_cap;
        pa.st = B.st;
        pa.bar = grid_sync_fresh(ar);
        SKL_REQUIRE(pa.rb <= 128, SKL_E_ARG, "persistent LSQR: more than 128 rows per workgroup");
        {
            BarrierLaunch guard(pg);
            if (pa.rb <= 64)
                k_lsqr_persist<1><<<pg, PS_T, 0, s>>>(pa);
            else
                k_lsqr_persist<2><<<pg, PS_T, 0, s>>>(pa);
            SKL_CHECK_LAUNCH();
        }
        SKL_HIP(hipStreamSynchronize(s));
        check_grid_fault(); // a barrier that gave up leaves the state unfinished
    }
    if (graph_ok) {
        GraphKey key;
        // every pointer and size baked into the captured kernels' arguments:
        // a replay is only valid when all of them are unchanged (inputs such as
        // the CSC values may live in a fresh allocation at each call)
        key.k = {(long)q.n, (long)q.d, (long)q.r, (long)(intptr_t)p.A, p.lda,
                 (long)(intptr_t)p.csr_ptr, (long)(intptr_t)p.csr_col,
                 (long)(intptr_t)p.csr_val, (long)(intptr_t)p.csc_ptr,
                 (long)(intptr_t)p.csc_row, (long)(intptr_t)p.csc_val, p.nnz,
                 (long)(intptr_t)p.E, (long)(intptr_t)p.M, p.ldm, (long)(intptr_t)B.ut, (long)(intptr_t)B.v,
                 (long)(intptr_t)B.x, (long)(intptr_t)B.st, (long)(intptr_t)B.hist,
                 (long)(intptr_t)B.Pn, (long)(intptr_t)B.Pt, (long)(intptr_t)B.Pm,
                 (long)(intptr_t)B.Einv, (long)(intptr_t)B.z, (long)(intptr_t)B.w,
                 (long)(intptr_t)B.t, (long)(intptr_t)B.Sb, (long)(intptr_t)B.Pz, B.hist_cap,
                 (long)(intptr_t)B.Sv, (long)(intptr_t)B.cnt,
                 q.fused ? 1L : 0L, chunk, (long)(intptr_t)s, prof ? 1L : 0L,
                 (long)q.spg, (long)(intptr_t)B.AT, (long)q.G,
                 // and the launch shape
                 (long)q.cw_a, (long)q.nck_a, (long)q.cw_m, (long)q.nck_m, (long)q.nrb_a,
                 (long)q.cpw_a, (long)q.nrb_m, (long)q.cpw_m, (long)q.nsb, (long)q.nc};
        auto it = g_graphs.find(key);
        if (it != g_graphs.end()) {
            exec = it->second;
        } else {
            // the chunk as an explicit graph: `chunk` iterations, then the
            // device-side done flag copied to the pinned host word
            hipGraph_t graph = nullptr;
            hipError_t e1 = hipGraphCreate(&graph, 0);
            if (e1 == hipSuccess) {
                t_rec.on = true;
                t_rec.graph = graph;
                t_rec.last = nullptr;
                t_rec.err = hipSuccess;
                for (int c = 0; c < chunk; ++c) launch_iteration(p, q, B, s);
                t_rec.on = false;
                e1 = t_rec.err;
                if (e1 == hipSuccess) {
                    hipGraphNode_t cp = nullptr;
                    e1 = hipGraphAddMemcpyNode1D(&cp, graph, t_rec.last ? &t_rec.last : nullptr,
                                                 t_rec.last ? 1 : 0, hf, &B.st->done, sizeof(int),
                                                 hipMemcpyDeviceToHost);
                }
            }
            if (e1 == hipSuccess) {
                // bounded cache: callers whose buffers move every call (host
                // inputs staged afresh) would otherwise grow it without end
                if (g_graphs.size() >= 32) lsqr_release_graphs();
                AllocLock lk(capture_mutex());
                e1 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            }
            if (graph) (void)hipGraphDestroy(graph);
            t_rec.keep.clear();
            if (e1 == hipSuccess) {
                g_graphs[key] = exec;
            } else {
                // run this solve's iterations without a graph
                graph_ok = false;
                exec = nullptr;
                (void)hipGetLastError();
            }
        }
    }
    // device-side done flag read back once per chunk
    SKL_HIP(hipMemcpyAsync(hf, &B.st->done, sizeof(int), hipMemcpyDeviceToHost, s));
    SKL_HIP(hipStreamSynchronize(s));
    long launched = 0;
    while (!hf->done) {
        if (graph_ok) {
            SKL_HIP(hipGraphLaunch(exec, s));
        } else {
            for (int c = 0; c < chunk; ++c)
                launch_iteration(p, q, B, s, (prof && c == 1) ? ev : nullptr);
            SKL_CHECK_LAUNCH();
            SKL_HIP(hipMemcpyAsync(hf, &B.st->done, sizeof(int), hipMemcpyDeviceToHost, s));
        }
        launched += chunk;
        SKL_HIP(hipStreamSynchronize(s));
        // a chunk that ended the solve may have timed kernels that returned
        // at once; only chunks that ran every iteration count
        if (prof && chunk > 1 && !hf->done) prof_accumulate(q, p.nnz);
        if (launched > max_it + chunk) break; // defensive: state always sets done by max_it
    }
    LsqrState hs;
    SKL_HIP(hipMemcpyAsync(&hs, B.st, sizeof hs, hipMemcpyDeviceToHost, s));
    SKL_HIP(hipStreamSynchronize(s));
    LsqrResult res;
    res.flag = hs.flag;
    res.it = hs.it;
    res.nrm_ar = hs.nrm_ar;
    res.nrm_a = hs.nrm_a;
    res.nrm_r = hs.nrm_r;
    if (hist && B.hist && hs.it > 0) {
        long cnt = std::min(hs.it, hl);
        std::vector<double> h((size_t)(5 * cnt));
        SKL_HIP(hipMemcpy(h.data(), B.hist, sizeof(double) * 5 * cnt, hipMemcpyDeviceToHost));
        for (long i = 0; i < cnt; ++i) {
            hist[i].alpha = h[5 * i];
            hist[i].beta = h[5 * i + 1];
            hist[i].rhobar = h[5 * i + 2];
            hist[i].phibar = h[5 * i + 3];
            hist[i].nrm_a = h[5 * i + 4];
        }
    }
    return res;
}

// ---- dense / sparse y = alpha A x + beta y (used for x0 residuals) --------------
__global__ __launch_bounds__(256) void k_axpby_part(long n, int nck, const double *__restrict__ P,
                                                    double alpha, double beta,
                                                    double *__restrict__ y) {
    long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < nck; ++c) s += P[(long)c * n + i];
    y[i] = (beta == 0.0 ? 0.0 : beta * y[i]) + alpha * s;
}

void dev_gemv_n_dense(long n, long d, const double *A, long lda, const double *x, double alpha,
                      double beta, double *y, Arena &ar) {
    hipStream_t s = ctx().stream;
    long rowblocks = (n + GN_ROWS - 1) / GN_ROWS;
    long target = 8L * ctx().num_cu;
    long nck = std::max<long>(1, std::min<long>((target + rowblocks - 1) / rowblocks, (d + 63) / 64));
    int cw = (int)std::max<long>(1, std::min<long>(GN_MAXCW, (d + nck - 1) / nck));
    int nc = (int)((d + cw - 1) / cw);
    double *P = ar.alloc_n<double>((size_t)nc * n);
    dim3 g((unsigned)rowblocks, (unsigned)nc);
    k_gemv_n_part<2><<<g, GN_THREADS, 0, s>>>(n, d, A, lda, x, cw, P, 0, nullptr);
    k_axpby_part<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, nc, P, alpha, beta, y);
    SKL_CHECK_LAUNCH();
}

// y = alpha A^T u + beta y (A n x d); the row-block partials are summed in a
// fixed order
void dev_gemv_t_dense(long n, long d, const double *A, long lda, const double *u, double alpha,
                      double beta, double *y, Arena &ar) {
    hipStream_t s = ctx().stream;
    long nrb = (n + GT_RB - 1) / GT_RB;
    long target = 8L * ctx().num_cu;
    long cchunks = std::max<long>(1, (target + nrb - 1) / nrb);
    int cpw = (int)std::max<long>(4, (d + cchunks - 1) / cchunks);
    double *P = ar.alloc_n<double>((size_t)nrb * d);
    dim3 g((unsigned)nrb, (unsigned)((d + cpw - 1) / cpw));
    k_gemv_t_part<2><<<g, GT_THREADS, 0, s>>>(n, d, A, lda, u, cpw, P, 0, nullptr);
    k_axpby_part<<<(unsigned)((d + 255) / 256), 256, 0, s>>>(d, (int)nrb, P, alpha, beta, y);
    SKL_CHECK_LAUNCH();
}

// ---- the autotuner -------------------------------------------------------------
// deterministic fill in [-1, 1) (splitmix64 of seed + index)
__global__ void k_fill_hash(long count, double *x, unsigned long long seed) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    x[i] = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
// M: upper triangular r x r with unit diagonal and small entries above it
__global__ void k_tune_upper(long r, double *M) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long j = blockIdx.y;
    for (; j < r; j += gridDim.y)
        if (i < r) M[i + j * r] = i < j ? M[i + j * r] * (0.1 / (double)r) : (i == j ? 1.0 : 0.0);
}

// Times every candidate launch shape on a synthetic problem of the given shape
// (the LSQR iteration as a solve runs it: `iters` iterations replayed as
// graphs, the operator A M with a dense upper triangular M) and stores the
// fastest for the shape's bucket.  Returns the chosen shape's time per
// iteration (ms); *default_ms: the default shape's.
double lsqr_autotune(bool sparse, long n, long d, long nnz, int iters, double *default_ms) {
    Context &C = ctx();
    SKL_REQUIRE(!C.no_alloc, SKL_E_UNSUPPORTED, "autotuning allocates device memory");
    SKL_REQUIRE(n >= d && d > 0 && (!sparse || (nnz > 0 && nnz <= n * d && nnz < 2147483647L)),
                SKL_E_ARG, "invalid autotuning shape");
    Arena &ar = C.arena;
    ar.reset();
    hipStream_t s = C.stream;
    LsqrProblem p;
    p.n = n;
    p.d = d;
    p.r = d;
    double *M = ar.alloc_n<double>((size_t)d * d);
    k_fill_hash<<<(unsigned)(((long)d * d + 255) / 256), 256, 0, s>>>((long)d * d, M, 11);
    k_tune_upper<<<dim3((unsigned)((d + 255) / 256), (unsigned)std::min<long>(d, 4096)), 256, 0, s>>>(
        d, M);
    p.M = M;
    p.ldm = d;
    if (!sparse) {
        double *A = ar.alloc_n<double>((size_t)n * d);
        k_fill_hash<<<(unsigned)(((long)n * d + 255) / 256), 256, 0, s>>>((long)n * d, A, 7);
        p.A = A;
        p.lda = n;
    } else {
        // nnz / d distinct rows per column, by a multiplicative hash of the
        // column, sorted
        const long per = std::max<long>(1, nnz / d);
        std::vector<int32_t> cp((size_t)d + 1), ri;
        ri.reserve((size_t)(per * d));
        std::vector<int32_t> col;
        for (long j = 0; j < d; ++j) {
            cp[(size_t)j] = (int32_t)ri.size();
            col.clear();
            unsigned long long h = 0x2545F4914F6CDD1Dull * (unsigned long long)(j + 1);
            for (long t = 0; t < per; ++t) {
                h ^= h << 13;
                h ^= h >> 7;
                h ^= h << 17;
                col.push_back((int32_t)(h % (unsigned long long)n));
            }
            std::sort(col.begin(), col.end());
            col.erase(std::unique(col.begin(), col.end()), col.end());
            ri.insert(ri.end(), col.begin(), col.end());
        }
        cp[(size_t)d] = (int32_t)ri.size();
        const long nz = (long)ri.size();
        int32_t *dcp = ar.alloc_n<int32_t>((size_t)d + 1);
        int32_t *dri = ar.alloc_n<int32_t>((size_t)nz);
        double *dv = ar.alloc_n<double>((size_t)nz);
        SKL_HIP(hipMemcpyAsync(dcp, cp.data(), sizeof(int32_t) * (d + 1), hipMemcpyHostToDevice, s));
        SKL_HIP(hipMemcpyAsync(dri, ri.data(), sizeof(int32_t) * nz, hipMemcpyHostToDevice, s));
        k_fill_hash<<<(unsigned)((nz + 255) / 256), 256, 0, s>>>(nz, dv, 5);
        int32_t *rp = ar.alloc_n<int32_t>((size_t)n + 1);
        int32_t *cc = ar.alloc_n<int32_t>((size_t)nz);
        double *cv = ar.alloc_n<double>((size_t)nz);
        csc_to_csr(n, d, nz, dcp, dri, dv, rp, cc, cv, ar);
        p.csr_ptr = rp;
        p.csr_col = cc;
        p.csr_val = cv;
        p.csc_ptr = dcp;
        p.csc_row = dri;
        p.csc_val = dv;
        p.nnz = nz;
        SKL_HIP(hipStreamSynchronize(s)); // the host arrays go out of scope
    }
    double *b = ar.alloc_n<double>((size_t)n);
    k_fill_hash<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, b, 3);
    double *y = ar.alloc_n<double>((size_t)d);
    SKL_CHECK_LAUNCH();
    // the default shape first, then the launch-per-kernel iteration's
    // shapes and, for the problems it takes, the persistent iteration's
    // workgroup counts
    std::vector<LsqrShape> cand(1);
    const bool fused = !sparse && lsqr_rows_wanted(n, d);
    for (int tm : {4, 8, 16, 2}) {
        if (fused) {
            if (tm == 2) continue;
            for (int w : {1, 2}) cand.push_back({tm, w, 1, 0});
        } else if (sparse) {
            if (tm == 2) continue;
            for (int l : {1, 0}) cand.push_back({tm, 1, l, 0});
        } else {
            cand.push_back({tm, 1, 1, 0});
        }
    }
    if (persist_eligible(p)) {
        const long cap = std::max(1, C.num_cu / 2), lo = persist_min_groups(p);
        std::vector<long> pw;
        for (long P : {lo, (n + 63) / 64, 2 * ((n + 63) / 64)})
            if (P <= cap && std::find(pw.begin(), pw.end(), P) == pw.end()) pw.push_back(P);
        for (long P : pw) cand.push_back({8, 1, 1, (int)P});
    }
    const size_t mark = ar.used();
    double best = INFINITY, def = INFINITY;
    LsqrShape chosen;
    for (const LsqrShape &c : cand) {
        t_shape_override = &c;
        double ms = INFINITY;
        try {
            for (int rep = 0; rep < 2; ++rep) {
                ar.rewind(mark);
                SKL_HIP(hipStreamSynchronize(s));
                const double t0 = wall();
                lsqr_solve(p, b, 0.0, iters, y, true, 16, nullptr, 0, ar);
                SKL_HIP(hipStreamSynchronize(s));
                ms = (wall() - t0) * 1e3 / iters; // the second repetition is kept
            }
        } catch (...) {
            t_shape_override = nullptr;
            throw;
        }
        t_shape_override = nullptr;
        if (&c == &cand[0]) def = ms;
        if (ms < best) {
            best = ms;
            chosen = c;
        }
    }
    ar.rewind(mark);
    lsqr_release_graphs(); // the synthetic problem's graphs
    {
        std::lock_guard<std::mutex> lk(g_shape_mu);
        g_shapes[shape_key(sparse, n, d, sparse ? nnz : 0)] = chosen;
    }
    if (default_ms) *default_ms = def;
    return best;
}

} // namespace skl

// ---- profiling hooks of the C ABI ------------------------------------------------
extern "C" {
void skl_profile_enable(int on) { skl::g_prof_on = on != 0; }
void skl_profile_reset(void) { skl::g_prof = skl::ProfStats(); }
int skl_profile_read(int slot, double *avg_ms, long *launches, double *bytes_per_launch) {
    if (slot < 0 || slot > 3) return SKL_E_ARG;
    const skl::ProfStats &P = skl::g_prof;
    if (launches) *launches = P.cnt[slot];
    if (avg_ms) *avg_ms = P.cnt[slot] ? P.ms[slot] / (double)P.cnt[slot] : 0.0;
    if (bytes_per_launch) *bytes_per_launch = P.bytes[slot];
    return SKL_OK;
}
}
