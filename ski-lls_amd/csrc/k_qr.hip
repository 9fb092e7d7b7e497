// k_qr.hip -- Householder QR family of the small sketched matrix SA on gfx950.
//
//  * qr_householder: blocked right-looking QR (dgeqrf semantics,
//    src/dense_qr.cpp:8-33).  Panels are factored by a multi-workgroup kernel
//    in which each workgroup holds a slice of rows in LDS; the column reductions
//    of one Householder step (||x||^2 and the products with every other panel
//    column) travel in a single grid-wide all-reduce, so a column costs one
//    grid barrier.  Trailing updates are GEMMs on the matrix cores.
//  * cpqr_greedy: dgeqp3 pivoting (LAPACK 3.9 dlaqp2 downdating) with the
//    same kernel over the whole matrix; every workgroup keeps a replica of the
//    partial column norms and makes the identical pivot decision.
//  * cpqr_hqrrp: HQRRP (hqrrp/lapack_compatible_sources/
//    NoFLA_HQRRP_WY_blk_var4_skint.c:313-581): block pivots from a Gaussian
//    sketch Y = G A (column-sliced multi-workgroup kernel), panel QR with pivoting,
//    GEMM updates and the Y downdate.
#include <algorithm>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

#include "k_common.h"

// Phase marks for an external timing harness; they compile to nothing here.
#ifndef SKL_PHASE
#define SKL_PHASE(kernel, mark)
#endif

namespace skl {

// ---------------------------------------------------------------------------
// Row-sliced multi-workgroup Householder kernel (one grid barrier per column)
// ---------------------------------------------------------------------------
template <class T> struct RsArgs {
    long m;        // rows
    int nc;        // columns held (all receive the reflectors)
    int npiv;      // leading columns that are pivot candidates (0 = no pivoting)
    int nsteps;    // reflectors to compute
    T *W;
    long ldw;
    T *tau;
    int32_t *jpvt;  // swapped along with pivoted columns (may be null)
    int32_t *swaps; // pivot column of each step (may be null)
    int hq_norms;   // 1: HQRRP downdate formula (1+r)(1-r); 0: LAPACK 1 - r^2
    int rpw;        // rows per workgroup
    int ldl;        // leading dimension of the LDS slice
    int two_phase;  // reduce-scatter + all-gather instead of all-to-all
    T *part;        // [2][P][nc]
    T *red;         // [2][nc]
    T *rowv;        // [2][nc]
    T *rcp;         // [P][nc]
    GridSync bar;   // barrier counter (zero at launch), error word, poll limit
    const int *only_if; // when set, the kernel runs only if *only_if != 0
};

template <class T> __device__ inline T dlapy2_d(T x, T y) {
    T xa = fabs(x), ya = fabs(y);
    T w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
    if (z == T(0)) return w;
    T q = z / w;
    return w * sqrt(T(1) + q * q);
}

template <class T> __device__ inline T eps_half();
template <> __device__ inline double eps_half<double>() { return 1.1102230246251565e-16; }
template <> __device__ inline float eps_half<float>() { return 5.9604644775390625e-08f; }

constexpr int RS_T = 512;

template <class T>
__global__ __launch_bounds__(RS_T) void k_qr_rowslice(RsArgs<T> a) {
    if (a.only_if && *a.only_if == 0) return;
    unsigned bar_target = 0;
    extern __shared__ unsigned char smem_raw[];
    const int P = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
    const int nc = a.nc, ldl = a.ldl;
    T *S = reinterpret_cast<T *>(smem_raw);      // ldl x nc
    T *Gs = S + (size_t)ldl * nc;                 // nc
    T *Rs = Gs + nc;                              // nc
    T *vn1 = Rs + nc;                             // npiv
    T *vn2 = vn1 + (a.npiv > 0 ? a.npiv : 1);     // npiv
    T *redbuf = vn2 + (a.npiv > 0 ? a.npiv : 1);  // RS_T
    int *flags = reinterpret_cast<int *>(redbuf + RS_T); // npiv + 2
    int *jps = flags + a.npiv + 2;                       // npiv: jpvt of the columns (workgroup 0)
    const long r0 = (long)w * a.rpw;
    const long r1 = min(a.m, r0 + a.rpw);
    const int nr = (int)max(0L, r1 - r0);
    const int npiv = a.npiv;
    const T tol3z = sqrt(eps_half<T>());

    const int lane = tid & 63, wave = tid >> 6;
    for (int q = wave; q < nc; q += RS_T / 64)
        for (int i = lane; i < nr; i += 64) S[i + q * ldl] = a.W[(r0 + i) + (long)q * a.ldw];
    // column indices follow their columns in LDS; stored back at the end
    const bool keep_jp = (w == 0 && a.jpvt != nullptr);
    if (keep_jp)
        for (int q = tid; q < npiv; q += RS_T) jps[q] = a.jpvt[q];
    __syncthreads();

    // all-reduce of vals[q] for q in [q0, nc) (q0 relative), result into Gs;
    // with_row also copies the published pivot row (rowv[buf]) into Rs, its
    // loads issued beside the partial sums
    auto allreduce = [&](int buf, int q0, bool with_row) -> bool {
        T *pw = a.part + ((size_t)buf * P + w) * nc;
        // values already in pw[q] for q >= q0
        if (!grid_barrier(a.bar, bar_target)) return false;
        if (!a.two_phase) {
            // eight adjacent lanes per column, each summing every eighth workgroup
            const int sub = tid & 7;
            for (int qb = q0; qb < nc; qb += RS_T / 8) {
                const int q = qb + (tid >> 3);
                T s = 0, rv = 0;
                if (q < nc) {
                    if (with_row && sub == 0) rv = a.rowv[(size_t)buf * nc + q];
                    for (int u = sub; u < P; u += 8) s += a.part[((size_t)buf * P + u) * nc + q];
                }
                s += __shfl_xor(s, 1, 64);
                s += __shfl_xor(s, 2, 64);
                s += __shfl_xor(s, 4, 64);
                if (q < nc && sub == 0) {
                    Gs[q] = s;
                    if (with_row) Rs[q] = rv;
                }
            }
        } else {
            int ncol = nc - q0;
            int per = (ncol + P - 1) / P;
            int qb = q0 + w * per, qe = min(nc, qb + per);
            for (int q = qb + tid; q < qe; q += RS_T) {
                T s = 0;
                for (int u = 0; u < P; ++u) s += a.part[((size_t)buf * P + u) * nc + q];
                a.red[(size_t)buf * nc + q] = s;
            }
            if (!grid_barrier(a.bar, bar_target)) return false;
            for (int q = q0 + tid; q < nc; q += RS_T) {
                Gs[q] = a.red[(size_t)buf * nc + q];
                if (with_row) Rs[q] = a.rowv[(size_t)buf * nc + q];
            }
        }
        (void)pw;
        return true;
    };

    // initial column norms for pivoting (over all rows)
    if (npiv > 0) {
        T *pw = a.part + ((size_t)1 * P + w) * nc;
        for (int q = tid; q < nc; q += RS_T) {
            T s = 0;
            for (int i = 0; i < nr; ++i) {
                T v = S[i + (size_t)q * ldl];
                s += v * v;
            }
            pw[q] = s;
        }
        if (!allreduce(1, 0, false)) return;
        __syncthreads();
        for (int q = tid; q < npiv; q += RS_T) {
            vn1[q] = sqrt(Gs[q]);
            vn2[q] = vn1[q];
        }
        __syncthreads();
    }

    for (int jj = 0; jj < a.nsteps; ++jj) {
        const int buf = jj & 1;
        SKL_PHASE(0, 0);
        if (npiv > 0) {
            // first index of max |vn1[jj..npiv)| (idamax), identical in all workgroups
            if (wave == 0) {
                int best = 0x7fffffff;
                T bv = T(-1);
                for (int q = jj + lane; q < npiv; q += 64)
                    if (fabs(vn1[q]) > bv) {
                        bv = fabs(vn1[q]);
                        best = q;
                    }
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    T ov = __shfl_xor(bv, o, 64);
                    int ob = __shfl_xor(best, o, 64);
                    if (ov > bv || (ov == bv && ob < best)) {
                        bv = ov;
                        best = ob;
                    }
                }
                if (lane == 0) flags[npiv] = best < npiv ? best : jj; // all NaN: stay
            }
            __syncthreads();
            const int pvt = flags[npiv];
            if (pvt != jj) {
                for (int i = tid; i < nr; i += RS_T) {
                    T t0 = S[i + (size_t)jj * ldl];
                    S[i + (size_t)jj * ldl] = S[i + (size_t)pvt * ldl];
                    S[i + (size_t)pvt * ldl] = t0;
                }
                if (tid == 0) {
                    vn1[pvt] = vn1[jj];
                    vn2[pvt] = vn2[jj];
                    if (w == 0 && a.jpvt) {
                        int32_t t1 = jps[pvt];
                        jps[pvt] = jps[jj];
                        jps[jj] = t1;
                    }
                }
            }
            if (w == 0 && tid == 0 && a.swaps) a.swaps[jj] = pvt;
            __syncthreads();
        }
        SKL_PHASE(0, 1);
        // partial Gram row: g[q] = sum_{rows > jj} S(:, jj) S(:, q), q >= jj
        T *pw = a.part + ((size_t)buf * P + w) * nc;
        const int ib = (int)max(0L, (long)jj + 1 - r0); // first local row below the diagonal
        {
            // eight adjacent lanes per column, each summing every eighth row
            const int sub = tid & 7;
            const T *sj = S + (size_t)jj * ldl;
            for (int qb = jj; qb < nc; qb += RS_T / 8) {
                const int q = qb + (tid >> 3);
                T s = 0;
                if (q < nc) {
                    const T *sq = S + (size_t)q * ldl;
                    for (int i = ib + sub; i < nr; i += 8) s += sj[i] * sq[i];
                }
                s += __shfl_xor(s, 1, 64);
                s += __shfl_xor(s, 2, 64);
                s += __shfl_xor(s, 4, 64);
                if (q < nc && sub == 0) pw[q] = s;
            }
        }
        const bool owner = (jj >= r0 && jj < r1);
        if (owner)
            for (int q = jj + tid; q < nc; q += RS_T)
                a.rowv[(size_t)buf * nc + q] = S[(jj - r0) + (size_t)q * ldl];
        SKL_PHASE(0, 2);
        if (!allreduce(buf, jj, true)) return;
        __syncthreads();
        SKL_PHASE(0, 3);
        // Householder scalars (dlarfg), computed redundantly
        const T alpha = Rs[jj];
        const T xnorm = sqrt(Gs[jj]);
        T tau = 0, beta = alpha, sc = 0;
        if (xnorm != T(0)) {
            beta = -copysign(dlapy2_d(alpha, xnorm), alpha);
            tau = (beta - alpha) / beta;
            sc = T(1) / (alpha - beta);
        }
        if (tau != T(0)) {
            // w_q = v^T a_q = a(jj, q) + sc * g_q ; C -= tau v w^T
            for (int q = jj + 1 + tid; q < nc; q += RS_T) Gs[q] = -tau * (Rs[q] + sc * Gs[q]);
            __syncthreads();
            // one wave per column, lanes over rows
            for (int i = ib + lane; i < nr; i += 64) {
                const T vi = sc * S[i + (size_t)jj * ldl];
                for (int q = jj + 1 + wave; q < nc; q += RS_T / 64)
                    S[i + (size_t)q * ldl] = S[i + (size_t)q * ldl] + vi * Gs[q];
            }
            if (owner)
                for (int q = jj + 1 + tid; q < nc; q += RS_T)
                    S[(jj - r0) + (size_t)q * ldl] = Rs[q] + Gs[q];
            __syncthreads();
            for (int i = ib + tid; i < nr; i += RS_T) S[i + (size_t)jj * ldl] *= sc;
            if (owner && tid == 0) S[(jj - r0) + (size_t)jj * ldl] = beta;
        } else {
            for (int q = jj + 1 + tid; q < nc; q += RS_T) Gs[q] = T(0);
        }
        if (w == 0 && tid == 0) a.tau[jj] = tau;
        __syncthreads();
        SKL_PHASE(0, 4);
        if (npiv > 0) {
            // partial norm downdating with the new row jj of R
            int any = 0;
            for (int q = jj + 1 + tid; q < npiv; q += RS_T) {
                int fl = 0;
                if (vn1[q] != T(0)) {
                    T rn = Rs[q] + Gs[q];
                    T temp = fabs(rn) / vn1[q];
                    if (a.hq_norms)
                        temp = (T(1) + temp) * (T(1) - temp);
                    else
                        temp = T(1) - temp * temp;
                    if (temp < T(0)) temp = T(0);
                    T q5 = vn1[q] / vn2[q];
                    T temp2 = temp * q5 * q5;
                    if (temp2 <= tol3z)
                        fl = 1;
                    else
                        vn1[q] = vn1[q] * sqrt(temp);
                }
                flags[q] = fl;
                any |= fl;
            }
            if (tid == 0) flags[npiv + 1] = 0;
            __syncthreads();
            if (any) atomicOr(&flags[npiv + 1], 1);
            __syncthreads();
            if (flags[npiv + 1]) {
                // exact norms of the flagged columns over the rows below jj
                T *rw = a.rcp + (size_t)w * nc;
                const int ib2 = (int)max(0L, (long)jj + 1 - r0);
                for (int q = jj + 1 + tid; q < npiv; q += RS_T) {
                    T s = 0;
                    if (flags[q])
                        for (int i = ib2; i < nr; ++i) {
                            T v = S[i + (size_t)q * ldl];
                            s += v * v;
                        }
                    rw[q] = s;
                }
                if (!grid_barrier(a.bar, bar_target)) return;
                for (int q = jj + 1 + tid; q < npiv; q += RS_T) {
                    if (!flags[q]) continue;
                    if (jj < a.m - 1) {
                        T s = 0;
                        for (int u = 0; u < P; ++u) s += a.rcp[(size_t)u * nc + q];
                        vn1[q] = sqrt(s);
                        vn2[q] = vn1[q];
                    } else {
                        vn1[q] = T(0);
                        vn2[q] = T(0);
                    }
                }
                __syncthreads();
                // rcp is rewritten only after another grid barrier
            }
        }
        __syncthreads();
    }
    for (int q = wave; q < nc; q += RS_T / 64)
        for (int i = lane; i < nr; i += 64) a.W[(r0 + i) + (long)q * a.ldw] = S[i + q * ldl];
    if (keep_jp)
        for (int q = tid; q < npiv; q += RS_T) a.jpvt[q] = jps[q];
}

// Workgroups of one grid-barrier launch: one per CU at most, and no more than
// half the device, so every workgroup is resident even beside a second such
// launch from another stream.
static int barrier_grid_cap() { return std::max(1, ctx().num_cu / 2); }

// ---------------------------------------------------------------------------
// Register-resident panel kernel (at most 64 columns)
// ---------------------------------------------------------------------------
// Lane q of every wave holds column q of the panel, wave w the rows w + 8 k
// (k < K) of its workgroup's slice of 8 K rows, so the column sums of a
// Householder step (||x||^2 below the diagonal and the products of the
// reflector column with every column) are plain per-lane sums: no shuffles,
// no LDS traffic for the matrix itself.  The reflector column's rows are
// broadcast through a small per-wave LDS vector (measured faster here than
// v_readlane, whose scalar results spill at these register counts).  Pivoting (dlaqp2 /
// HQRRP panel rules, as in k_qr_rowslice) moves no data: each lane keeps the
// logical position of its column, and the partial norms live in the lane of
// their column; the pivot is found by a wave-wide argmax that every wave of
// every workgroup computes identically.  The matrix is read and written in
// 64-row chunks staged through LDS so that the global accesses run along
// columns.  One grid barrier per column, none with one workgroup.
template <class T> struct RpArgs {
    long m;       // rows
    int nc;       // columns (<= 64)
    int npiv;     // leading columns that are pivot candidates (0 = no pivoting)
    int nsteps;   // reflectors to compute
    T *W;
    long ldw;
    T *tau;
    int32_t *jpvt;  // (may be null)
    int32_t *swaps; // pivot column of each step (may be null)
    int hq_norms;   // 1: HQRRP downdate formula (1+r)(1-r); 0: LAPACK 1 - r^2
    T *part;        // [2][P][64] workgroup sums
    T *rowv;        // [2][64] row jj of the panel (workgroup 0)
    T *rcp;         // [P][64] norm recomputation sums
    GridSync bar;   // barrier counter (zero at launch), error word, poll limit
    const int *only_if; // when set, the kernel runs only if *only_if != 0
};

constexpr int RP_W = 8, RP_T = 64 * RP_W;

template <class T, int K>
__global__ __launch_bounds__(RP_T) void k_qr_regpanel(RpArgs<T> a) {
    static_assert(K % 8 == 0, "rows come in 64-row chunks");
    if (a.only_if && *a.only_if == 0) return;
    __shared__ T tile[64 * 65];
    __shared__ T s_g[2][RP_W][64];
    __shared__ T s_row[2][64];
    __shared__ T s_r[RP_W][64];
    __shared__ T s_gs[RP_W][64];
    __shared__ T s_col[RP_W][K];
    const int P = gridDim.x, blk = blockIdx.x, tid = threadIdx.x;
    const int lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const long r0 = (long)blk * (8 * K);
    const int nc = a.nc, npiv = a.npiv;
    const int BIG = 0x7fffffff;
    const T tol3z = sqrt(eps_half<T>());
    unsigned bar_target = 0;

    T x[K];
#pragma unroll
    for (int c = 0; c < K / 8; ++c) {
        const long rb = r0 + 64 * c;
        for (int e = tid; e < 64 * 64; e += RP_T) {
            const int r = e & 63, q = e >> 6;
            const long row = rb + r;
            tile[r + q * 65] = (q < nc && row < a.m) ? a.W[row + (long)q * a.ldw] : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) x[8 * c + kk] = tile[(w + 8 * kk) + lane * 65];
        __syncthreads();
    }
    int pos = lane < nc ? lane : BIG; // logical position of this lane's column
    int jp = 0;
    T vn1 = 0, vn2 = 0;
    if (npiv > 0) {
        if (a.jpvt && lane < npiv) jp = a.jpvt[lane];
        T sq = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) sq += x[k] * x[k];
        s_r[w][lane] = sq;
        __syncthreads();
        T tot = 0;
#pragma unroll
        for (int u = 0; u < RP_W; ++u) tot += s_r[u][lane];
        if (P > 1) {
            if (w == 0) a.rcp[(size_t)blk * 64 + lane] = tot;
            if (!grid_barrier(a.bar, bar_target)) return;
            tot = 0;
            for (int u = 0; u < P; ++u) tot += a.rcp[(size_t)u * 64 + lane];
        }
        vn1 = sqrt(tot);
        vn2 = vn1;
        __syncthreads();
    }

    for (int jj = 0; jj < a.nsteps; ++jj) {
        const int buf = jj & 1;
        SKL_PHASE(2, 0);
        // ---- pivot: largest partial norm, first logical position on ties ----
        int L = jj;
        if (npiv > 0) {
            const T av = fabs(vn1);
            T bv = T(-1);
            int bp = BIG, bl = -1;
            if (pos >= jj && pos < npiv && av == av) {
                bv = av;
                bp = pos;
                bl = lane;
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const T ov = __shfl_xor(bv, o, 64);
                const int op = __shfl_xor(bp, o, 64), ol = __shfl_xor(bl, o, 64);
                if (ov > bv || (ov == bv && op < bp)) {
                    bv = ov;
                    bp = op;
                    bl = ol;
                }
            }
            const unsigned long long atj = __ballot(pos == jj);
            const int lj = __builtin_ctzll(atj); // positions 0..nc-1 are all held
            if (bl < 0) bl = lj; // no finite candidate: the column stays
            const int pvt = __shfl(pos, bl, 64);
            if (lane == lj) pos = pvt;
            if (lane == bl) pos = jj;
            if (blk == 0 && tid == 0 && a.swaps) a.swaps[jj] = pvt;
            L = __builtin_amdgcn_readfirstlane(bl); // identical in every lane
        }
        SKL_PHASE(2, 1);
        // ---- column sums over this workgroup's rows below jj ----
        // rows w + 8 k with k >= kfirst lie below row jj; row jj itself is
        // register kj of wave jj % 8 of workgroup 0.  Only the first 64 rows
        // (k < 8) can lie at or above jj (< 64).
        const int jl = jj - (int)r0;
        const int kfirst = jl < w ? 0 : (jl - w) / 8 + 1;
        const int kj = (blk == 0 && w == (jj & 7)) ? (jj >> 3) : -1;
        if (lane == L) {
#pragma unroll
            for (int k = 0; k < K; ++k) s_col[w][k] = x[k];
        }
        T g = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) g += (k >= kfirst ? s_col[w][k] : T(0)) * x[k];
#pragma unroll
        for (int k = 8; k < K; ++k) g += s_col[w][k] * x[k];
        s_g[buf][w][lane] = g;
        if (kj >= 0) {
            T rv = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) rv = k == kj ? x[k] : rv;
            s_row[buf][lane] = rv;
        }
        __syncthreads();
        SKL_PHASE(2, 2);
        T G = 0, Rs;
        if (P == 1) {
#pragma unroll
            for (int u = 0; u < RP_W; ++u) G += s_g[buf][u][lane];
            Rs = s_row[buf][lane];
        } else {
            if (w == 0) {
                T tot = 0;
#pragma unroll
                for (int u = 0; u < RP_W; ++u) tot += s_g[buf][u][lane];
                a.part[((size_t)buf * P + blk) * 64 + lane] = tot;
                if (blk == 0) a.rowv[(size_t)buf * 64 + lane] = s_row[buf][lane];
            }
            if (!grid_barrier(a.bar, bar_target)) return;
            // wave w sums the workgroups u = w mod 8, then every wave adds
            // the eight partial sums in the same order
            T pt = 0;
            for (int u = w; u < P; u += RP_W) pt += a.part[((size_t)buf * P + u) * 64 + lane];
            Rs = a.rowv[(size_t)buf * 64 + lane];
            s_gs[w][lane] = pt;
            __syncthreads();
#pragma unroll
            for (int u = 0; u < RP_W; ++u) G += s_gs[u][lane];
        }
        SKL_PHASE(2, 3);
        // ---- Householder scalars (dlarfg), replicated ----
        const T alpha = readlane(Rs, L);
        const T xnorm = sqrt(readlane(G, L));
        T tau = 0, beta = alpha, sc = 0;
        if (xnorm != T(0)) {
            beta = -copysign(dlapy2_d(alpha, xnorm), alpha);
            tau = (beta - alpha) / beta;
            sc = T(1) / (alpha - beta);
        }
        // w_q = v^T a_q = a(jj, q) + sc g_q ; C -= tau v w^T
        const bool act = pos > jj && pos != BIG;
        const T f = act ? -tau * (Rs + sc * G) : T(0);
        const bool isL = lane == L;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const T v = sc * s_col[w][k];
            const T upd = isL ? v : (act ? x[k] + v * f : x[k]);
            const T rowj = isL ? beta : (act ? Rs + f : x[k]);
            x[k] = k >= kfirst ? upd : (k == kj ? rowj : x[k]);
        }
#pragma unroll
        for (int k = 8; k < K; ++k) {
            const T v = sc * s_col[w][k];
            x[k] = isL ? v : (act ? x[k] + v * f : x[k]);
        }
        if (blk == 0 && tid == 0) a.tau[jj] = tau;
        SKL_PHASE(2, 4);
        // ---- partial norm downdating (dlaqp2 / HQRRP) ----
        if (npiv > 0) {
            bool fl = false;
            if (act && pos < npiv && vn1 != T(0)) {
                const T rn = Rs + f;
                T temp = fabs(rn) / vn1;
                if (a.hq_norms)
                    temp = (T(1) + temp) * (T(1) - temp);
                else
                    temp = T(1) - temp * temp;
                if (temp < T(0)) temp = T(0);
                const T q5 = vn1 / vn2;
                const T temp2 = temp * q5 * q5;
                if (temp2 <= tol3z)
                    fl = true;
                else
                    vn1 = vn1 * sqrt(temp);
            }
            if (__ballot(fl)) { // same decision in every wave and workgroup
                T sq = 0;
#pragma unroll
                for (int k = 0; k < 8; ++k) sq += k >= kfirst ? x[k] * x[k] : T(0);
#pragma unroll
                for (int k = 8; k < K; ++k) sq += x[k] * x[k];
                s_r[w][lane] = sq;
                __syncthreads();
                T tot = 0;
#pragma unroll
                for (int u = 0; u < RP_W; ++u) tot += s_r[u][lane];
                if (P > 1) {
                    if (w == 0) a.rcp[(size_t)blk * 64 + lane] = tot;
                    if (!grid_barrier(a.bar, bar_target)) return;
                    tot = 0;
                    for (int u = 0; u < P; ++u) tot += a.rcp[(size_t)u * 64 + lane];
                }
                if (fl) {
                    vn1 = (jj < a.m - 1) ? sqrt(tot) : T(0);
                    vn2 = vn1;
                }
            }
        }
    }
    // ---- store, columns at their logical positions ----
#pragma unroll
    for (int c = 0; c < K / 8; ++c) {
        const long rb = r0 + 64 * c;
        __syncthreads();
        if (pos < nc) {
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) tile[(w + 8 * kk) + pos * 65] = x[8 * c + kk];
        }
        __syncthreads();
        for (int e = tid; e < 64 * nc; e += RP_T) {
            const int r = e & 63, q = e >> 6;
            const long row = rb + r;
            if (row < a.m) a.W[row + (long)q * a.ldw] = tile[r + q * 65];
        }
    }
    if (npiv > 0 && a.jpvt && blk == 0 && w == 0 && pos < npiv) a.jpvt[pos] = jp;
}

template <class T, int K>
static void launch_regpanel(int P, const RpArgs<T> &a, hipStream_t s) {
    k_qr_regpanel<T, K><<<P, RP_T, 0, s>>>(a);
}

// Panel factorization with the register-resident kernel; false when the panel
// does not fit it (more than 64 columns, or more rows than one grid holds).
// Slices of 8 K rows: one workgroup when the panel fits one (K as small as
// that allows), otherwise the largest K that stays in registers (32 rows
// per lane in FP64, 64 in FP32), which keeps the workgroup count -- and with
// it the barrier latency -- low.  SKL_QR_PANEL=lds selects the LDS kernel instead.
template <class T>
static bool regpanel_qr(long m, int nc, int npiv, int nsteps, T *W, long ldw, T *tau,
                        int32_t *jpvt, int32_t *swaps, int hq_norms, const int *only_if,
                        Arena &ar) {
    static const bool off = [] {
        const char *e = getenv("SKL_QR_PANEL");
        return e && std::string(e) == "lds";
    }();
    if (off || nc > 64 || nsteps > 64) return false;
    constexpr int KMAX = sizeof(T) == 8 ? 32 : 64;
    int K = 8;
    while (K < KMAX && 8L * K < m) K *= 2;
    const long P = (m + 8L * K - 1) / (8L * K);
    if (P > barrier_grid_cap()) return false;
    RpArgs<T> a;
    a.m = m;
    a.nc = nc;
    a.npiv = npiv;
    a.nsteps = nsteps;
    a.W = W;
    a.ldw = ldw;
    a.tau = tau;
    a.jpvt = jpvt;
    a.swaps = swaps;
    a.hq_norms = hq_norms;
    a.only_if = only_if;
    a.part = ar.alloc_n<T>((size_t)2 * P * 64);
    a.rowv = ar.alloc_n<T>((size_t)2 * 64);
    a.rcp = ar.alloc_n<T>((size_t)P * 64);
    a.bar = grid_sync_fresh(ar);
    hipStream_t s = ctx().stream;
    BarrierLaunch guard((int)P);
    switch (K) {
    case 8: launch_regpanel<T, 8>((int)P, a, s); break;
    case 16: launch_regpanel<T, 16>((int)P, a, s); break;
    case 32: launch_regpanel<T, 32>((int)P, a, s); break;
    default: launch_regpanel<T, KMAX>((int)P, a, s); break;
    }
    SKL_CHECK_LAUNCH();
    return true;
}

static constexpr size_t RS_LDS_BUDGET = 150 * 1024;

// Factor (or pivot-factor) the m x nc block at W: nsteps reflectors, pivoting
// over the first npiv columns when npiv > 0.
// The CholeskyQR2 + Householder-reconstruction panel (k_cqr.hip) for FP64
// panels of at most 64 columns that are factored whole (all pivoting
// columns), tall enough that the column-by-column kernel needs several
// workgroups; SKL_QR_CQR=0 turns it off.
int *cqr_panel(long mp, int b, int pivot, double *Wp, long ldw, double *tau, int32_t *jpvt,
               int32_t *swaps, int hq_norms, double *Tout, Arena &ar);
static bool cqr_wanted(long m, int nc, int npiv, int nsteps) {
    static const bool on = [] {
        const char *e = getenv("SKL_QR_CQR");
        return !(e && e[0] == '0');
    }();
    return on && !ctx().no_alloc && nc == nsteps && nc <= 64 && (npiv == 0 || npiv == nc) &&
           m >= 512;
}

// Returns the CholeskyQR2 panel's device flag when that panel was queued (it
// then wrote the reflectors' T factor to Tout, when given, unless the flag
// ends nonzero), else null.
template <class T>
static const int *rowslice_qr(long m, int nc, int npiv, int nsteps, T *W, long ldw, T *tau,
                              int32_t *jpvt, int32_t *swaps, int hq_norms, Arena &ar,
                              T *Tout = nullptr) {
    if (nsteps <= 0) return nullptr;
    // with the fast panel queued first, the kernels below run only when it
    // gave up (its flag nonzero)
    const int *only_if = nullptr;
    if constexpr (std::is_same<T, double>::value) {
        if (cqr_wanted(m, nc, npiv, nsteps))
            only_if = cqr_panel(m, nc, npiv > 0 ? 1 : 0, W, ldw, tau, jpvt, swaps, hq_norms,
                                Tout, ar);
    }
    if (regpanel_qr<T>(m, nc, npiv, nsteps, W, ldw, tau, jpvt, swaps, hq_norms, only_if, ar))
        return only_if;
    const int P_half = barrier_grid_cap();
    const int P_max = std::max(1, ctx().num_cu);
    size_t fixed = sizeof(T) * (2 * (size_t)nc + 2 * (size_t)std::max(npiv, 1) + RS_T) +
                   sizeof(int) * (2 * (size_t)npiv + 4) + 64;
    // fewest workgroups whose slices fit in LDS (each step costs one barrier,
    // whose latency grows with the workgroup count) ...
    long rpw = -1;
    int P = 0;
    for (int p = 1; p <= P_max; ++p) {
        long r = (m + p - 1) / p;
        long ld = r | 1;
        size_t need = sizeof(T) * (size_t)ld * nc + fixed;
        if (need <= RS_LDS_BUDGET) {
            rpw = r;
            P = p;
            break;
        }
    }
    SKL_REQUIRE(rpw > 0, SKL_E_UNSUPPORTED, "QR panel too large for the LDS-resident kernel");
    // ... but at least ~16 workgroups so the per-step row work stays short
    long target_rows = 256;
    if (rpw > target_rows) {
        int p2 = (int)std::min<long>(std::min(P_half, 16), (m + target_rows - 1) / target_rows);
        if (p2 > P) {
            P = p2;
            rpw = (m + P - 1) / P;
        }
    }
    P = (int)((m + rpw - 1) / rpw);
    RsArgs<T> a;
    a.m = m;
    a.nc = nc;
    a.npiv = npiv;
    a.nsteps = nsteps;
    a.W = W;
    a.ldw = ldw;
    a.tau = tau;
    a.jpvt = jpvt;
    a.swaps = swaps;
    a.hq_norms = hq_norms;
    a.rpw = (int)rpw;
    a.only_if = only_if;
    a.ldl = (int)(rpw | 1);
    a.two_phase = ((long)P * nc > 32768) ? 1 : 0;
    a.part = ar.alloc_n<T>((size_t)2 * P * nc);
    a.red = ar.alloc_n<T>((size_t)2 * nc);
    a.rowv = ar.alloc_n<T>((size_t)2 * nc);
    a.rcp = ar.alloc_n<T>((size_t)P * nc);
    a.bar = grid_sync_fresh(ar);
    size_t sh = sizeof(T) * (size_t)a.ldl * nc + fixed;
    hipStream_t s = ctx().stream;
    if (sh > 65536)
        SKL_HIP(hipFuncSetAttribute((const void *)k_qr_rowslice<T>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    // grids over more than half the device (wide matrices in the greedy path)
    // always run alone
    BarrierLaunch guard(P, P > P_half);
    k_qr_rowslice<T><<<P, RS_T, sh, s>>>(a);
    SKL_CHECK_LAUNCH();
    return only_if;
}

// ---------------------------------------------------------------------------
// Block reflector helpers
// ---------------------------------------------------------------------------
// Vf (mp x b) = unit lower trapezoid of the factored panel
template <class T>
__global__ void k_vfull(long mp, int b, const T *W, long ldw, T *Vf) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int q = blockIdx.y;
    if (i >= mp) return;
    T v = (i > q) ? W[i + (long)q * ldw] : (i == q ? T(1) : T(0));
    Vf[i + (long)q * mp] = v;
}

// T (b x b, ld b, b <= 64) of the compact WY form I - V T V^T of the
// reflectors H_0 .. H_{b-1} (dlarft forward columnwise), from VtV = V^T V and
// tau, by recursive doubling: splitting the reflectors into halves V = [V1
// V2], Q = (I - V1 T1 V1^T)(I - V2 T2 V2^T) gives T = [T1, -T1 (V1^T V2) T2;
// 0, T2].  Eight 8 x 8 diagonal blocks come from the column recurrence
// T(0:i, i) = -tau_i T(0:i, 0:i) VtV(0:i, i), then three merge levels
// (8 -> 16 -> 32 -> 64) form the off-diagonal blocks with two small products
// each.  Reflectors past b have tau = 0 (zero rows and columns).
template <class T>
// only_if: run only when *only_if is nonzero, with VtV null: V^T V is then
// formed here from Vf (mp x b), 64 rows at a time through LDS.
__global__ __launch_bounds__(256) void k_build_t(int b, const T *VtV, const T *tau, T *Tm,
                                                 const int *only_if, long mp, const T *Vf) {
    constexpr int LD = 65;
    __shared__ T sT[64 * LD], sV[64 * LD], sY[64 * LD];
    __shared__ T st[64];
    if (only_if && *only_if == 0) return;
    const int tid = threadIdx.x;
    if (VtV) {
        for (int e = tid; e < 64 * 64; e += 256) {
            const int r = e & 63, q = e >> 6;
            sV[r + q * LD] = (r < b && q < b) ? VtV[r + q * b] : T(0);
        }
    } else {
        // thread (r = t & 63, g = t >> 6) sums entries (r, g + 4 m) over the
        // rows; chunks of 64 rows of Vf staged in sY
        T acc[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) acc[m] = T(0);
        const int r = tid & 63, g = tid >> 6;
        for (long i0 = 0; i0 < mp; i0 += 64) {
            __syncthreads();
            for (int e = tid; e < 64 * 64; e += 256) {
                const int ii = e & 63, q = e >> 6;
                sY[ii + q * LD] = (i0 + ii < mp && q < b) ? Vf[(i0 + ii) + (long)q * mp] : T(0);
            }
            __syncthreads();
            for (int ii = 0; ii < 64; ++ii) {
                const T vr = sY[ii + r * LD];
#pragma unroll
                for (int m = 0; m < 16; ++m) acc[m] += vr * sY[ii + (g + 4 * m) * LD];
            }
        }
#pragma unroll
        for (int m = 0; m < 16; ++m) sV[r + (g + 4 * m) * LD] = acc[m];
    }
    for (int e = tid; e < 64 * 64; e += 256) {
        const int r = e & 63, q = e >> 6;
        sT[r + q * LD] = T(0);
    }
    if (tid < 64) st[tid] = tid < b ? tau[tid] : T(0);
    __syncthreads();
    // 8 x 8 diagonal blocks: one thread per block runs the column recurrence
    if (tid < 8) {
        const int o = tid * 8;
        for (int i = 0; i < 8; ++i) {
            const int c = o + i;
            const T ti = st[c];
            T tmp[8];
            for (int k = 0; k < i; ++k) tmp[k] = -ti * sV[(o + k) + c * LD];
            for (int r = 0; r < i; ++r) {
                T acc = T(0);
                for (int k = r; k < i; ++k) acc += sT[(o + r) + (o + k) * LD] * tmp[k];
                sT[(o + r) + c * LD] = acc;
            }
            sT[c + c * LD] = ti;
        }
    }
    __syncthreads();
    for (int sz = 8; sz < 64; sz *= 2) {
        // pairs (A, C) of sz x sz diagonal blocks at rows a0 and a0 + sz:
        // Y = VtV(A, C) T_C, then T(A, C) = -T_A Y
        const int npair = 64 / (2 * sz), per = sz * sz;
        for (int e = tid; e < npair * per; e += 256) {
            const int pr = e / per, w = e % per;
            const int i = w % sz, j = w / sz;
            const int a0 = pr * 2 * sz, c0 = a0 + sz;
            T acc = T(0);
            for (int k = 0; k <= j; ++k) acc += sV[(a0 + i) + (c0 + k) * LD] * sT[(c0 + k) + (c0 + j) * LD];
            sY[(a0 + i) + (c0 + j) * LD] = acc;
        }
        __syncthreads();
        for (int e = tid; e < npair * per; e += 256) {
            const int pr = e / per, w = e % per;
            const int i = w % sz, j = w / sz;
            const int a0 = pr * 2 * sz, c0 = a0 + sz;
            T acc = T(0);
            for (int k = i; k < sz; ++k) acc += sT[(a0 + i) + (a0 + k) * LD] * sY[(a0 + k) + (c0 + j) * LD];
            sT[(a0 + i) + (c0 + j) * LD] = -acc;
        }
        __syncthreads();
    }
    for (int e = tid; e < b * b; e += 256) Tm[e] = sT[(e % b) + (e / b) * LD];
}

// C (mp x nc2) := (I - V T V^T)^T C
template <class T>
static void apply_block_left_t(long mp, int b, const T *Vf, const T *Tm, long nc2, T *Cm, long ldc,
                               Arena &ar) {
    if (nc2 <= 0 || b <= 0) return;
    T *Wt = ar.alloc_n<T>((size_t)b * nc2);
    T *Wt2 = ar.alloc_n<T>((size_t)b * nc2);
    gemm<T>('T', 'N', b, nc2, mp, T(1), Vf, mp, Cm, ldc, T(0), Wt, b);
    gemm<T>('T', 'N', b, nc2, b, T(1), Tm, b, Wt, b, T(0), Wt2, b);
    gemm<T>('N', 'N', mp, nc2, b, T(-1), Vf, mp, Wt2, b, T(1), Cm, ldc);
}

// Vf and the T factor of a factored panel.  cqr_flag: the CholeskyQR2 panel's
// flag (rowslice_qr), whose kernels wrote T already unless the flag is
// nonzero; k_build_t then forms V^T V from Vf itself (the rare case).
template <class T>
static void panel_t(long mp, int b, const T *Wp, long ldw, const T *tau, T *Vf, T *Tm, Arena &ar,
                    const int *cqr_flag = nullptr) {
    dim3 g((unsigned)((mp + 255) / 256), (unsigned)b);
    k_vfull<T><<<g, 256, 0, ctx().stream>>>(mp, b, Wp, ldw, Vf);
    if (cqr_flag) {
        k_build_t<T><<<1, 256, 0, ctx().stream>>>(b, nullptr, tau, Tm, cqr_flag, mp, Vf);
    } else {
        T *VtV = ar.alloc_n<T>((size_t)b * b);
        gemm<T>('T', 'N', b, b, mp, T(1), Vf, mp, Vf, mp, T(0), VtV, b);
        k_build_t<T><<<1, 256, 0, ctx().stream>>>(b, VtV, tau, Tm, nullptr, 0, nullptr);
    }
    SKL_CHECK_LAUNCH();
}

// ---------------------------------------------------------------------------
// Blocked Householder QR: factor the first nfac columns, apply Q^T to the rest
// ---------------------------------------------------------------------------
// Look-ahead: after panel p, only the next panel's columns are updated on the
// main stream; the rest of the trailing matrix is updated on the side stream
// while the next panel is factored (the panel kernel holds few CUs, the GEMMs
// take the others).  Hand-offs: the side update of panel p starts after the
// main stream has built V_p, T_p; the main stream updates the columns of
// panel p+2 with V_{p+1} only after the side update of panel p (which also
// reached them) is done.  V and T are double-buffered for that distance.
// SKL_QR_LOOKAHEAD=0 turns it off.
static bool lookahead_on() {
    static const bool on = [] {
        const char *e = getenv("SKL_QR_LOOKAHEAD");
        return !(e && e[0] == '0');
    }();
    return on;
}

template <class T>
void qr_householder(long m, long ncol, long nfac, T *W, long ldw, T *tau, Arena &ar) {
    long kmax = std::min(m, nfac);
    const int nb = (nfac >= 256) ? 64 : 32;
    const long npan = (kmax + nb - 1) / nb;
    const bool ahead = npan >= 3 && lookahead_on() && !ctx().no_alloc;
    T *Vf[2], *Tm[2];
    Vf[0] = ar.alloc_n<T>((size_t)m * nb);
    Tm[0] = ar.alloc_n<T>((size_t)nb * nb);
    Vf[1] = ahead ? ar.alloc_n<T>((size_t)m * nb) : Vf[0];
    Tm[1] = ahead ? ar.alloc_n<T>((size_t)nb * nb) : Tm[0];
    Context &c = ctx();
    if (ahead) c.make_side();
    bool side_busy = false;
    for (long j = 0, p = 0; j < kmax; j += nb, ++p) {
        int b = (int)std::min<long>(nb, kmax - j);
        long mp = m - j;
        T *V = Vf[p & 1], *Tp = Tm[p & 1];
        const int *cf = rowslice_qr<T>(mp, b, 0, b, W + j + j * ldw, ldw, tau + j, nullptr,
                                       nullptr, 0, ar, Tp);
        long nc2 = ncol - j - b;
        if (nc2 <= 0) continue;
        panel_t<T>(mp, b, W + j + j * ldw, ldw, tau + j, V, Tp, ar, cf);
        T *C2 = W + j + (j + b) * ldw;
        if (!ahead) {
            apply_block_left_t<T>(mp, b, V, Tp, nc2, C2, ldw, ar);
            continue;
        }
        const long nn = std::min<long>(nb, nc2), nr = nc2 - nn;
        if (side_busy) SKL_HIP(hipStreamWaitEvent(c.stream, c.ev[1], 0));
        apply_block_left_t<T>(mp, b, V, Tp, nn, C2, ldw, ar);
        side_busy = false;
        if (nr > 0) {
            SKL_HIP(hipEventRecord(c.ev[0], c.stream));
            OnSideStream on_side;
            SKL_HIP(hipStreamWaitEvent(c.stream, c.ev[0], 0));
            apply_block_left_t<T>(mp, b, V, Tp, nr, C2 + nn * ldw, ldw, ar);
            SKL_HIP(hipEventRecord(c.ev[1], c.stream));
            side_busy = true;
        }
    }
    if (side_busy) SKL_HIP(hipStreamWaitEvent(c.stream, c.ev[1], 0));
}
template void qr_householder<double>(long, long, long, double *, long, double *, Arena &);
template void qr_householder<float>(long, long, long, float *, long, float *, Arena &);

// ---------------------------------------------------------------------------
// Greedy CPQR (dgeqp3) over the first nfac columns; columns past nfac are
// carried along (they receive Q^T, never pivot).
// ---------------------------------------------------------------------------
template <class T>
void cpqr_greedy(long m, long ncol, long nfac, T *W, long ldw, int32_t *jpvt, T *tau, Arena &ar) {
    long kmax = std::min(m, nfac);
    rowslice_qr<T>(m, (int)ncol, (int)nfac, (int)kmax, W, ldw, tau, jpvt, nullptr, 0, ar);
}
template void cpqr_greedy<double>(long, long, long, double *, long, int32_t *, double *, Arena &);
template void cpqr_greedy<float>(long, long, long, float *, long, int32_t *, float *, Arena &);

// ---------------------------------------------------------------------------
// HQRRP pieces
// ---------------------------------------------------------------------------
constexpr int YQ_MAXROWS = 128; // two rows per lane
constexpr int YQ_MAXP = 128;    // workgroups of one launch

// Column-sliced multi-workgroup QRP of Y (mY x nv, mY <= 128) for nsteps pivots
// with the HQRRP norm downdating (skint.c:920-990).  Every wave keeps K
// columns in registers (lane l holds rows l and l + 64), together with their
// partial norms and current logical positions; Y itself is only read.
// One grid barrier per pivot: before it each workgroup publishes its best
// candidate (norm, logical position) and that column's values; after it every
// wave picks the global winner (largest norm, first position on ties -- the
// idamax rule), forms the Householder vector itself and updates its columns.
// perm_out[t] is the column at logical position t after nsteps pivots.
template <class T> struct YqArgs {
    int mY, nv, nsteps;
    const T *Y;
    long ldy;
    int32_t *perm_out;
    T *cand;        // [2][P] candidate norm
    int32_t *candp; // [2][P] candidate logical position
    T *ctau;        // [2][P] tau of the candidate's reflector
    T *cvec;        // [2][P][YQ_MAXROWS] candidate's reflector v (v_jj = 1, zero above)
    GridSync bar;   // barrier counter (zero at launch), error word, poll limit
    const int32_t *skip = nullptr; // k_yqrp_lanes: nonzero *skip = nothing to do
};

template <class T>
__device__ inline bool yq_better(T v, int p, T bv, int bp) {
    return v > bv || (v == bv && p < bp);
}

template <int K> struct Log2 {
    static constexpr int v = K == 1 ? 0 : K == 2 ? 1 : K == 4 ? 2 : K == 8 ? 3 : 4;
};

// Sums each of v[0..K) over the wave.  log2 K halving exchanges (on lane bits
// 5, 4, ...) leave one value per lane, then plain butterflies finish; every
// lane ends with the total of value c(lane), whose holder split_lane(c) names.
template <int K, class T> __device__ inline T wave_sum_split(T (&v)[K], int lane) {
    constexpr int L = Log2<K>::v;
#pragma unroll
    for (int s = 0; s < L; ++s) {
        const int m = 32 >> s, half = K >> (s + 1);
        const bool up = (lane & m) != 0;
#pragma unroll
        for (int j = 0; j < half; ++j) {
            T send = up ? v[j] : v[j + half];
            T keep = up ? v[j + half] : v[j];
            v[j] = keep + __shfl_xor(send, m, 64);
        }
    }
    T r = v[0];
#pragma unroll
    for (int m = 32 >> L; m > 0; m >>= 1) r += __shfl_xor(r, m, 64);
    return r;
}
template <int K> __device__ inline int split_lane(int c) {
    constexpr int L = Log2<K>::v;
    int l = 0;
#pragma unroll
    for (int b = 0; b < L; ++b)
        if ((c >> b) & 1) l |= 32 >> (L - 1 - b);
    return l;
}

// K columns per wave, W waves per workgroup
template <class T, int K, int W>
__global__ __launch_bounds__(64 * W) void k_yqrp(YqArgs<T> a) {
    __shared__ T s_wv[W];
    __shared__ int s_wp[W];
    __shared__ T s_norm[2 * W * K]; // partial norms (dn, en) of each wave's columns
    extern __shared__ unsigned char yq_smem[];
    const int P = gridDim.x, w = blockIdx.x;
    T *s_vec = reinterpret_cast<T *>(yq_smem);      // P x YQ_MAXROWS candidate reflectors
    T *s_cv = s_vec + (size_t)P * YQ_MAXROWS;       // P candidate norms
    T *s_tau = s_cv + P;                            // P candidate taus
    int *s_cp = reinterpret_cast<int *>(s_tau + P); // P candidate positions
    unsigned bar_target = 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = lane, r1 = lane + 64;
    const int mY = a.mY, nv = a.nv;
    const int NONE = 0x7fffffff;
    const T tol3z = sqrt(eps_half<T>());
    const int cbase = (w * W + wave) * K;
    T x0[K], x1[K];
    // dn, en of skint.c:920-990 are wave-uniform: kept in LDS, and the scalar
    // arithmetic on them runs lane-parallel (lane k works on column k)
    T *dn = s_norm + 2 * K * wave, *en = dn + K;
    int pos[K];
    {
        T sq[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int c = cbase + k;
            const bool ok = c < nv;
            const T *col = a.Y + (long)(ok ? c : 0) * a.ldy;
            x0[k] = (ok && r0 < mY) ? col[r0] : T(0);
            x1[k] = (ok && r1 < mY) ? col[r1] : T(0);
            sq[k] = x0[k] * x0[k] + x1[k] * x1[k];
            pos[k] = ok ? c : NONE;
        }
        T tot = wave_sum_split<K>(sq, lane);
        T mine = __shfl(tot, split_lane<K>(lane & (K - 1)), 64);
        if (lane < K) {
            T nrm = sqrt(mine);
            dn[lane] = nrm;
            en[lane] = nrm;
        }
    }
    for (int jj = 0; jj < a.nsteps; ++jj) {
        const int buf = jj & 1;
        SKL_PHASE(1, 0);
        // best remaining column of this wave, then of this workgroup
        T bv = T(-1);
        int bp = NONE, bk = 0;
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (pos[k] >= jj && pos[k] != NONE && yq_better(fabs(dn[k]), pos[k], bv, bp)) {
                bv = fabs(dn[k]);
                bp = pos[k];
                bk = k;
            }
        if (lane == 0) {
            s_wv[wave] = bv;
            s_wp[wave] = bp;
        }
        __syncthreads();
        T gv = T(-1);
        int gp = NONE, gw = 0;
        for (int u = 0; u < W; ++u)
            if (yq_better(s_wv[u], s_wp[u], gv, gp)) {
                gv = s_wv[u];
                gp = s_wp[u];
                gw = u;
            }
        if (wave == gw) {
            // publish the workgroup's candidate as its reflector over rows
            // jj.. (dlarfg), so the other workgroups need no scalar work
            T c0 = T(0), c1 = T(0);
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (k == bk) {
                    c0 = x0[k];
                    c1 = x1[k];
                }
            const T alpha = (jj < 64) ? readlane(c0, jj) : readlane(c1, jj - 64);
            const T xnorm = sqrt(wave_sum((r0 > jj ? c0 * c0 : T(0)) + (r1 > jj ? c1 * c1 : T(0))));
            T tau = 0, sc = 0;
            if (xnorm != T(0)) {
                T beta = -copysign(dlapy2_d(alpha, xnorm), alpha);
                tau = (beta - alpha) / beta;
                sc = T(1) / (alpha - beta);
            }
            T *dst = a.cvec + ((size_t)buf * P + w) * YQ_MAXROWS;
            dst[r0] = r0 > jj ? sc * c0 : (r0 == jj ? T(1) : T(0));
            dst[r1] = r1 > jj ? sc * c1 : (r1 == jj ? T(1) : T(0));
            if (lane == 0) {
                a.cand[buf * P + w] = gv;
                a.candp[buf * P + w] = gp;
                a.ctau[buf * P + w] = tau;
            }
        }
        SKL_PHASE(1, 1);
        if (!grid_barrier(a.bar, bar_target)) return;
        SKL_PHASE(1, 2);
        // every workgroup's candidate, staged in LDS in one round trip
        for (int e = threadIdx.x; e < P * YQ_MAXROWS; e += 64 * W)
            s_vec[e] = a.cvec[(size_t)buf * P * YQ_MAXROWS + e];
        for (int t = threadIdx.x; t < P; t += 64 * W) {
            s_cv[t] = a.cand[buf * P + t];
            s_cp[t] = a.candp[buf * P + t];
            s_tau[t] = a.ctau[buf * P + t];
        }
        __syncthreads();
        SKL_PHASE(1, 3);
        // global winner, computed identically by every wave
        T v = T(-1);
        int p = NONE, u = 0;
        for (int t = lane; t < P; t += 64) {
            T cv = s_cv[t];
            int cp = s_cp[t];
            if (yq_better(cv, cp, v, p)) {
                v = cv;
                p = cp;
                u = t;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            T ov = __shfl_xor(v, o, 64);
            int op = __shfl_xor(p, o, 64), ou = __shfl_xor(u, o, 64);
            if (yq_better(ov, op, v, p)) {
                v = ov;
                p = op;
                u = ou;
            }
        }
        const int pl = p;
        const T vv0 = s_vec[(size_t)u * YQ_MAXROWS + r0];
        const T vv1 = s_vec[(size_t)u * YQ_MAXROWS + r1];
        const T tau = s_tau[u];
        SKL_PHASE(1, 4);
        // own columns: logical swap of positions jj and pl, then C -= tau v (v^T C)
        bool act[K];
        T sd[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (pos[k] == pl)
                pos[k] = jj;
            else if (pos[k] == jj)
                pos[k] = pl;
            act[k] = pos[k] > jj && pos[k] != NONE;
            sd[k] = vv0 * x0[k] + vv1 * x1[k];
        }
        if (tau != T(0)) {
            T tot = wave_sum_split<K>(sd, lane);
            const T f = -tau * __shfl(tot, split_lane<K>(lane & (K - 1)), 64);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const T fk = readlane(f, k);
                if (act[k]) {
                    if (r0 >= jj) x0[k] = x0[k] + vv0 * fk;
                    if (r1 >= jj) x1[k] = x1[k] + vv1 * fk;
                }
            }
        }
        // norm downdating, lane k for column k
        T xjv = T(0);
        bool av = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const T xj = (jj < 64) ? readlane(x0[k], jj) : readlane(x1[k], jj - 64);
            if (lane == k) {
                xjv = xj;
                av = act[k];
            }
        }
        bool rec = false;
        if (av) {
            const T d0 = dn[lane];
            if (d0 != T(0)) {
                T temp = fabs(xjv) / d0;
                temp = (T(1) + temp) * (T(1) - temp);
                if (temp < T(0)) temp = T(0);
                const T t5 = d0 / en[lane];
                const T temp2 = temp * t5 * t5;
                if (temp2 <= tol3z)
                    rec = true;
                else
                    dn[lane] = d0 * sqrt(temp);
            }
        }
        if (__ballot(rec)) { // rare: exact norms of the trailing rows
            T rs[K];
#pragma unroll
            for (int k = 0; k < K; ++k)
                rs[k] = (r0 > jj ? x0[k] * x0[k] : T(0)) + (r1 > jj ? x1[k] * x1[k] : T(0));
            T tot = wave_sum_split<K>(rs, lane);
            const T mine = __shfl(tot, split_lane<K>(lane & (K - 1)), 64);
            if (rec) {
                const T nrm = (mY - jj - 1 > 0) ? sqrt(mine) : T(0);
                dn[lane] = nrm;
                en[lane] = nrm;
            }
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (pos[k] != NONE) a.perm_out[pos[k]] = cbase + k;
    }
}

// Column-per-lane pivot search for the HQRRP sketch height (74 rows): lane l
// of a wave holds one column of Y entirely in registers, with its partial
// norms and logical position.  The product of a column with the reflector,
// its update and the norm downdating -- the exact recomputation included --
// are loops inside one lane; the only cross-lane work of a step is the
// argmax.  Two kernels run these steps with the same per-column arithmetic
// (so they choose the same pivots bit for bit):
//  * k_yqrp_top: one workgroup over the (at most 512) columns of largest
//    initial norm, no grid barrier;
//  * k_yqrp_lanes: every column, 512 per workgroup, one grid barrier per pivot.
constexpr int YQ_ROWS = 74; // nb + pp of HQRRP (64 + 10)

// Register layout: during group g of eight pivot steps, y[i] holds row 8 g + i
// of the column, so at step jj = 8 g + s the pivot row is y[s] (s static), the
// rows above it are retired and never read again, and the rows at or beyond
// live = 74 - 8 g are zero.  The column shifts up by eight rows between
// groups, so every access has a fixed register index.  The loops run over
// chunks of eight rows and skip the chunks that start at or beyond `live` (a
// wave-uniform branch; skipping terms that are exact zeros changes no sum).
constexpr int YQ_CH = 8, YQ_NCH = (YQ_ROWS + YQ_CH - 1) / YQ_CH;
#define YQ_ROWS_FROM(S, body)                                                                 \
    _Pragma("unroll") for (int c_ = 0; c_ < YQ_NCH; ++c_) {                                   \
        if (YQ_CH * c_ + YQ_CH > (S) && YQ_CH * c_ < live) {                                  \
            _Pragma("unroll") for (int i_ = 0; i_ < YQ_CH; ++i_) {                            \
                const int r = YQ_CH * c_ + i_;                                                \
                if (r >= (S) && r < YQ_ROWS) { body }                                         \
            }                                                                                 \
        }                                                                                     \
    }

template <int S> using yq_sub = std::integral_constant<int, S>;

// step(yq_sub<s>{}, jj, live) for jj = 8 g + s < nsteps, then regroup(live)
// after each group; false when a step says stop
template <class F, class G>
__device__ inline bool yq_for_steps(int nsteps, F &&step, G &&regroup) {
    for (int g = 0; 8 * g < nsteps; ++g) {
        const int live = YQ_ROWS - 8 * g;
#define YQ_SUB(S)                                                                             \
    if (8 * g + (S) < nsteps && !step(yq_sub<(S)>{}, 8 * g + (S), live)) return false;
        YQ_SUB(0) YQ_SUB(1) YQ_SUB(2) YQ_SUB(3) YQ_SUB(4) YQ_SUB(5) YQ_SUB(6) YQ_SUB(7)
#undef YQ_SUB
        regroup(live);
    }
    return true;
}

// acc += v[R] * x, with v[R] taken from lane R % 16 of each row of sixteen
// lanes of vb (lane l of vb[c] holds v[16 c + l % 16]) by a fused DPP
// row-broadcast FMA: one instruction per term, no scalar round trip
template <int R>
__device__ inline void yq_bcast_fma(double &acc, const double (&vb)[5], double x) {
    bcast16_fma<R>(acc, vb[R >> 4], x);
}
template <int R> __device__ inline void yq_bcast_fma(float &acc, const float (&vb)[5], float x) {
    const float v = __int_as_float(
        __builtin_amdgcn_update_dpp(0, __float_as_int(vb[R >> 4]), 0x150 + (R & 15), 0xF, 0xF, false));
    acc = fmaf(v, x, acc);
}
template <class T> __device__ inline void yq_fence(T (&vb)[5]) {
    asm volatile("s_nop 1" : "+v"(vb[0]), "+v"(vb[1]), "+v"(vb[2]), "+v"(vb[3]), "+v"(vb[4]));
}
// (float: the broadcast goes through v_mov_dpp, whose hazards the compiler
// tracks)

// rows [S, live) of the column, chunk by chunk (a chunk is skipped when it
// starts at or beyond live), with compile-time row numbers
template <int S, class F> __device__ inline void yq_rows_static(int live, F &&f) {
    static_for<0, YQ_NCH>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        if constexpr (YQ_CH * c + YQ_CH > S) {
            if (YQ_CH * c < live) {
                static_for<0, YQ_CH>([&](auto ii) {
                    constexpr int r = YQ_CH * c + decltype(ii)::value;
                    if constexpr (r >= S && r < YQ_ROWS) f(std::integral_constant<int, r>{});
                });
            }
        }
    });
}

// y -= tau v (v^T y) for an active column (f = 0 otherwise), v broadcast
// from vb (yq_v) by DPP: the dot in four interleaved partial sums.
template <int S, class T>
__device__ inline void yq_apply(T (&y)[YQ_ROWS], T (&vb)[5], T tau, bool act, int live) {
    yq_fence(vb);
    T g[4] = {0, 0, 0, 0};
    yq_rows_static<S>(live, [&](auto rc) {
        constexpr int r = decltype(rc)::value;
        yq_bcast_fma<r>(g[(r - S) & 3], vb, y[r]);
    });
    T f = (act && tau != T(0)) ? -tau * ((g[0] + g[2]) + (g[1] + g[3])) : T(0);
    yq_rows_static<S>(live, [&](auto rc) {
        constexpr int r = decltype(rc)::value;
        yq_bcast_fma<r>(y[r], vb, f);
    });
}

// norm downdating of skint.c:920-990 (NoFLA_QRP_downdate_partial_norms) with
// the updated pivot-row entry y[S], recomputed in the lane when needed (the
// quotients and square roots from rcp_nr / rsqrt_nr)
template <int S, class T>
__device__ inline void yq_downdate(const T (&y)[YQ_ROWS], T &dn, T &en, bool act, int mY, int jj,
                                   T tol3z, int live) {
    if (act && dn != T(0)) {
        const T rd = rcp_nr(dn);
        T temp = fabs(y[S]) * rd;
        temp = (T(1) + temp) * (T(1) - temp);
        if (temp < T(0)) temp = T(0);
        const T t5 = dn * rcp_nr(en);
        const T temp2 = temp * t5 * t5;
        if (temp2 <= tol3z) {
            T sq = 0;
            YQ_ROWS_FROM(S + 1, sq = fma(y[r], y[r], sq);)
            dn = (mY - jj - 1 > 0) ? sqrt_nr(sq) : T(0);
            en = dn;
        } else {
            dn = dn * sqrt_nr(temp);
        }
    }
}

// next group: drop the eight retired rows (rows of skipped chunks are zero
// and stay so)
template <class T> __device__ inline void yq_shift8(T (&y)[YQ_ROWS], int live) {
    YQ_ROWS_FROM(0, y[r] = (r + YQ_CH < YQ_ROWS) ? y[r + YQ_CH] : T(0);)
}

// a column's registers and initial norms (the same arithmetic in both kernels)
template <class T>
__device__ inline void yq_load(T (&y)[YQ_ROWS], T &dn, T &en, const T *Y, long ldy, int col,
                               bool valid, int mY) {
    const T *src = Y + (long)(valid ? col : 0) * ldy;
    T sq = 0;
#pragma unroll
    for (int r = 0; r < YQ_ROWS; ++r) {
        y[r] = (valid && r < mY) ? src[r] : T(0);
        sq += y[r] * y[r];
    }
    dn = sqrt(sq);
    en = dn;
}

// The reflector (dlarfg) of the wave's column in lane wl over its live rows
// jj.. (pivot row y[S]): yq_store_rows has the lane store its raw rows S.. in
// pairs to `rows` (LDS, 16-byte aligned); yq_reflector has a wave read them
// back spread over its lanes (after the stores: the same wave, whose LDS
// operations complete in order, or any wave after a barrier) for alpha = y[S]
// and the sum of squares of the rows below it, summed by DPP lane moves; then
// tau and the scale sc of the rows below the pivot (v[S] = 1, v[r] = sc y[r];
// a reader forms v with yq_v).  Wave-uniform results, the same bits in every
// wave that reads the same rows.
template <class T> struct alignas(2 * sizeof(T)) yq_pair { T a, b; };
template <int S, class T>
__device__ inline void yq_store_rows(const T (&y)[YQ_ROWS], int wl, int lane, T *rows, int live) {
    if (lane == wl) {
        yq_pair<T> *d2 = reinterpret_cast<yq_pair<T> *>(rows);
#pragma unroll
        for (int r0 = S & ~1; r0 < YQ_ROWS; r0 += 2)
            if (YQ_CH * (r0 / YQ_CH) < live) d2[r0 / 2] = {y[r0], r0 + 1 < YQ_ROWS ? y[r0 + 1] : T(0)};
    }
}
template <int S, class T>
__device__ inline void yq_reflector(const T *rows, int lane, int live, T &tau, T &sc) {
    const T x0 = (lane > S && lane < live) ? rows[lane] : T(0);
    const T x1 = (lane + 64 < live) ? rows[64 + lane] : T(0);
    const T xnorm2 = wave_sum_bcast(fma(x1, x1, x0 * x0));
    const T alpha = rows[S];
    tau = 0;
    sc = 0;
    if (xnorm2 != T(0)) {
        const T beta = -copysign(sqrt_nr(fma(alpha, alpha, xnorm2)), alpha);
        tau = (beta - alpha) * rcp_nr(beta);
        sc = rcp_nr(alpha - beta);
    }
}
template <int S, class T>
__device__ inline void yq_publish(const T (&y)[YQ_ROWS], int wl, int lane, T *rows, int live,
                                  T &tau, T &sc) {
    yq_store_rows<S>(y, wl, lane, rows, live);
    yq_reflector<S>(rows, lane, live, tau, sc);
}

// the reflector for yq_apply from the published raw rows: lane l of vb[c]
// holds v[16 c + l % 16] (v[S] = 1, v[r] = sc y[r] below; rows outside
// [S, live) are never read)
template <int S, class T> __device__ inline void yq_v(const T *src, T sc, int lane, T (&vb)[5]) {
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const int r = 16 * c + (lane & 15);
        vb[c] = r == S ? T(1) : sc * src[r];
    }
}

template <class T, int W>
__global__ __launch_bounds__(64 * W) void k_yqrp_lanes(YqArgs<T> a) {
    __shared__ T s_wv[W];
    __shared__ int s_wp[W], s_wl[W];
    __shared__ __align__(16) T s_rows[80]; // the workgroup's candidate's raw rows
    extern __shared__ unsigned char yq_smem[];
    // nothing to do when the single-workgroup search already chose the pivots
    if (a.skip && *a.skip != 0) return;
    const int P = gridDim.x, blk = blockIdx.x;
    T *s_vec = reinterpret_cast<T *>(yq_smem);      // P x YQ_MAXROWS candidate reflectors
    T *s_cv = s_vec + (size_t)P * YQ_MAXROWS;       // P candidate norms
    T *s_tau = s_cv + P;                            // P candidate taus
    int *s_cp = reinterpret_cast<int *>(s_tau + P); // P candidate positions
    unsigned bar_target = 0;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int mY = a.mY, nv = a.nv;
    const int NONE = 0x7fffffff;
    const T tol3z = sqrt(eps_half<T>());
    const int col = (blk * W + wave) * 64 + lane;
    const bool valid = col < nv;
    T y[YQ_ROWS];
    T dn, en;
    yq_load(y, dn, en, a.Y, a.ldy, col, valid, mY);
    int pos = valid ? col : NONE;
    auto step = [&](auto Sc, int jj, int live) -> bool {
        constexpr int S = decltype(Sc)::value;
        const int buf = jj & 1;
        SKL_PHASE(1, 0);
        // best column of this wave (every lane ends with the same answer) ...
        T bv = T(-1);
        int bp = NONE, bl = 0;
        {
            const T av = fabs(dn);
            const T cand = (pos >= jj && pos != NONE && av == av) ? av : T(-1);
            const int wl = wave_argmax(cand, pos);
            if (wl >= 0) {
                bv = readlane(cand, wl);
                bp = __builtin_amdgcn_readlane(pos, wl);
                bl = wl;
            }
        }
        if (lane == 0) {
            s_wv[wave] = bv;
            s_wp[wave] = bp;
            s_wl[wave] = bl;
        }
        __syncthreads();
        // ... and of this workgroup, published as its reflector over rows jj..
        T gv = T(-1);
        int gp = NONE, gw = 0;
        for (int u = 0; u < W; ++u)
            if (yq_better(s_wv[u], s_wp[u], gv, gp)) {
                gv = s_wv[u];
                gp = s_wp[u];
                gw = u;
            }
        if (wave == gw) {
            // raw rows to cvec through LDS, sc after them (entry YQ_MAXROWS - 1)
            T *dst = a.cvec + ((size_t)buf * P + blk) * YQ_MAXROWS;
            T tau, sc;
            yq_publish<S>(y, s_wl[gw], lane, s_rows, live, tau, sc);
            dst[lane] = s_rows[lane];
            if (lane < 80 - 64) dst[64 + lane] = s_rows[64 + lane];
            if (lane == 0) {
                dst[YQ_MAXROWS - 1] = sc;
                a.cand[buf * P + blk] = gv;
                a.candp[buf * P + blk] = gp;
                a.ctau[buf * P + blk] = tau;
            }
        }
        SKL_PHASE(1, 1);
        if (!grid_barrier(a.bar, bar_target)) return false;
        SKL_PHASE(1, 2);
        // every workgroup's candidate, staged in LDS in one round trip (rows
        // at or beyond R are never read)
        for (int e = threadIdx.x; e < P * YQ_MAXROWS; e += 64 * W)
            s_vec[e] = a.cvec[(size_t)buf * P * YQ_MAXROWS + e];
        for (int t = threadIdx.x; t < P; t += 64 * W) {
            s_cv[t] = a.cand[buf * P + t];
            s_cp[t] = a.candp[buf * P + t];
            s_tau[t] = a.ctau[buf * P + t];
        }
        __syncthreads();
        SKL_PHASE(1, 3);
        // global winner, computed identically by every wave
        T v = T(-1);
        int p = NONE, u = 0;
        for (int t = lane; t < P; t += 64)
            if (yq_better(s_cv[t], s_cp[t], v, p)) {
                v = s_cv[t];
                p = s_cp[t];
                u = t;
            }
        {
            const int wl = wave_argmax(v, p);
            if (wl >= 0) {
                p = __builtin_amdgcn_readlane(p, wl);
                u = __builtin_amdgcn_readlane(u, wl);
            } else {
                p = NONE;
                u = 0;
            }
        }
        const int pl = p;
        const T tau = s_tau[u];
        T vb[5];
        yq_v<S>(s_vec + (size_t)u * YQ_MAXROWS, s_vec[(size_t)u * YQ_MAXROWS + YQ_MAXROWS - 1],
                lane, vb);
        SKL_PHASE(1, 4);
        // logical swap of positions jj and pl, then y -= tau v (v^T y)
        if (pos == pl)
            pos = jj;
        else if (pos == jj)
            pos = pl;
        const bool act = pos > jj && pos != NONE;
        yq_apply<S>(y, vb, tau, act, live);
        yq_downdate<S>(y, dn, en, act, mY, jj, tol3z, live);
        // s_v, s_wv and the staged candidates are rewritten only after the
        // next step's barriers
        return true;
    };
    if (!yq_for_steps(a.nsteps, step, [&](int live) { yq_shift8(y, live); })) return;
    if (pos != NONE) a.perm_out[pos] = col;
}

// Column norms of Y (mY rows), eight lanes per column (each row read by one
// lane of a group of eight, so a column's rows come in as whole cache lines)
template <class T> __global__ __launch_bounds__(256) void k_ynorms(int nv, int mY, const T *Y,
                                                                   long ldy, T *nrm) {
    const int c = blockIdx.x * 32 + (threadIdx.x >> 3), part = threadIdx.x & 7;
    T sq = 0;
    if (c < nv) {
        const T *src = Y + (long)c * ldy;
        for (int r = part; r < mY; r += 8) sq = fma(src[r], src[r], sq);
    }
    sq += __shfl_xor(sq, 1, 64);
    sq += __shfl_xor(sq, 2, 64);
    sq += __shfl_xor(sq, 4, 64);
    if (c < nv && part == 0) nrm[c] = sqrt(sq);
}

// Pivot search of a block on the columns that can win it, in one workgroup.
// Partial norms never grow: a downdate multiplies by sqrt((1+t)(1-t)) <= 1,
// and a recomputation returns the norm of the trailing rows of the column,
// which the orthogonal reflectors keep at its initial norm up to rounding
// (relative 64 x 74 ulps).  So a column of initial norm N never has a partial
// norm above N (1 + 1e-9) in this search.  The kernel keeps the (at most 512)
// columns of largest initial norm -- those above a threshold chosen from a
// histogram of log2 of the norms in 1/16-octave bins -- one per lane, runs the
// pivot steps of k_yqrp_lanes on them, and checks at every step that the
// winner's partial norm exceeds X (1 + margin), where X is the largest
// initial norm left out (from k_ynorms, whose summation order moves a norm
// by a few ulps, far inside the margin of 1e-6): then no left-out column
// could have won, and the
// pivots are those of the search over all columns.  Otherwise it stops, and
// *done = 0 tells k_yqrp_lanes (launched after it) to run the search over all
// columns; on success *done = 1 and perm_out is written.  Left-out columns
// move only when a pivot displaces them from position jj; their positions are
// tracked in LDS (column at each position).
constexpr int YT_W = 8, YT_MAXNV = 16384;
template <class T> struct YtArgs {
    int mY, nv, nsteps;
    const T *Y;
    long ldy;
    const T *nrm0; // initial column norms (k_ynorms)
    int32_t *perm_out;
    int32_t *done;
};

template <class T>
__global__ __launch_bounds__(64 * YT_W) void k_yqrp_top(YtArgs<T> a) {
    constexpr int W = YT_W, NT = 64 * W, VR = 80;
    __shared__ __align__(16) T s_v[2][VR]; // the winner's raw rows
    __shared__ T s_bv[2][W], s_tau[2], s_sc[2];
    __shared__ int s_bp[2][W], s_bl[2][W];
    __shared__ int s_hist[256];
    __shared__ int s_cols[NT];
    __shared__ int s_cnt, s_cut;
    __shared__ T s_red[W];
    extern __shared__ unsigned char yt_smem[];
    int *s_colat = reinterpret_cast<int *>(yt_smem); // column at each logical position
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int mY = a.mY, nv = a.nv;
    const int NONE = 0x7fffffff;
    const T tol3z = sqrt(eps_half<T>());
    const T margin = sizeof(T) == 8 ? T(1e-6) : T(1e-2);
    for (int t = tid; t < nv; t += NT) s_colat[t] = t;
    // choose the columns: all of them when they fit, else the histogram cut
    T xmax = T(-1); // largest initial norm left out (-1: none)
    int nsel;
    if (nv <= NT) {
        nsel = nv;
        s_cols[tid] = tid;
    } else {
        T mx = T(0);
        for (int c = tid; c < nv; c += NT) {
            const T v = a.nrm0[c];
            if (v == v) mx = fmax(mx, v);
        }
        mx = wave_max(mx);
        if (lane == 0) s_red[wave] = mx;
        for (int b = tid; b < 256; b += NT) s_hist[b] = 0;
        if (tid == 0) s_cnt = 0;
        __syncthreads();
        mx = s_red[0];
        for (int q = 1; q < W; ++q) mx = fmax(mx, s_red[q]);
        // bin 0..254: 16 log2(max / norm); 255: zero, NaN or far below
        auto bin_of = [&](T v) -> int {
            if (!(v > T(0))) return 255;
            const float l = 16.0f * log2f((float)(mx / v));
            return l < 254.0f ? (int)l : 254;
        };
        for (int c = tid; c < nv; c += NT) atomicAdd(&s_hist[bin_of(a.nrm0[c])], 1);
        __syncthreads();
        if (wave == 0) {
            // cut = largest B with sum of bins below B at most NT
            int h[4], run = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                h[q] = s_hist[4 * lane + q];
                run += h[q];
            }
            int incl = run; // inclusive prefix over lanes of the four-bin sums
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(incl, o, 64);
                if (lane >= o) incl += t;
            }
            int before = incl - run, cut = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (before + h[q] <= NT) cut = 4 * lane + q + 1;
                s_hist[4 * lane + q] = before; // first slot of the bin
                before += h[q];
            }
            // bins are consecutive, so the largest passing cut is the answer
            cut = wave_max_int(cut);
            if (lane == 0) s_cut = cut;
        }
        __syncthreads();
        const int cut = s_cut;
        T lo = T(-1);
        for (int c = tid; c < nv; c += NT) {
            const T v = a.nrm0[c];
            const int bn = bin_of(v);
            if (bn < cut) {
                // slots in bin order (largest norms in the first waves)
                s_cols[atomicAdd(&s_hist[bn], 1)] = c;
                atomicAdd(&s_cnt, 1);
            } else if (v == v)
                lo = fmax(lo, v);
        }
        lo = wave_max(lo);
        __syncthreads();
        if (lane == 0) s_red[wave] = lo;
        __syncthreads();
        xmax = s_red[0];
        for (int q = 1; q < W; ++q) xmax = fmax(xmax, s_red[q]);
        nsel = s_cnt;
    }
    __syncthreads();
    SKL_PHASE(3, 0);
    const T lim = xmax < T(0) ? T(-1) : xmax * (T(1) + margin);
    const bool valid = tid < nsel;
    const int col = valid ? s_cols[tid] : 0;
    T y[YQ_ROWS];
    T dn, en;
    yq_load(y, dn, en, a.Y, a.ldy, col, valid, mY);
    int pos = valid ? col : NONE;
    SKL_PHASE(3, 1);
    auto step = [&](auto Sc, int jj, int live) -> bool {
        constexpr int S = decltype(Sc)::value;
        const int buf = jj & 1;
        SKL_PHASE(3, 2);
        // every wave offers its best column (value, position, lane); the
        // winner over the waves alone stores its rows and forms its
        // reflector, which the others read after a second barrier (eight
        // waves storing and reducing their candidates at once, one barrier a
        // step, measured slower: their LDS stores queue in the CU's one LDS
        // pipeline)
        {
            const T av = fabs(dn);
            const T cand = (pos >= jj && pos != NONE && av == av) ? av : T(-1);
            const int l = wave_argmax(cand, pos);
            if (lane == 0) {
                s_bv[buf][wave] = l >= 0 ? readlane(cand, l) : T(-1);
                s_bp[buf][wave] = l >= 0 ? __builtin_amdgcn_readlane(pos, l) : NONE;
                s_bl[buf][wave] = l;
            }
        }
        SKL_PHASE(3, 3);
        __syncthreads();
        SKL_PHASE(3, 4);
        // the winner over the waves (largest norm, first position on ties)
        T gv = T(-1);
        int gp = NONE, u = 0;
        for (int q = 0; q < W; ++q)
            if (yq_better(s_bv[buf][q], s_bp[buf][q], gv, gp)) {
                gv = s_bv[buf][q];
                gp = s_bp[buf][q];
                u = q;
            }
        // a left-out column might have won: leave the block to k_yqrp_lanes
        if (gp == NONE || !(gv > lim)) return false;
        if (wave == u) {
            T tau, sc;
            yq_publish<S>(y, s_bl[buf][u], lane, &s_v[buf][0], live, tau, sc);
            if (lane == 0) {
                s_tau[buf] = tau;
                s_sc[buf] = sc;
            }
        }
        __syncthreads();
        const T tau = s_tau[buf];
        T vb[5];
        yq_v<S>(&s_v[buf][0], s_sc[buf], lane, vb);
        if (tid == 0) {
            const int t = s_colat[jj];
            s_colat[jj] = s_colat[gp];
            s_colat[gp] = t;
        }
        if (pos == gp)
            pos = jj;
        else if (pos == jj)
            pos = gp;
        const bool act = pos > jj && pos != NONE;
        SKL_PHASE(3, 5);
        yq_apply<S>(y, vb, tau, act, live);
        SKL_PHASE(3, 6);
        yq_downdate<S>(y, dn, en, act, mY, jj, tol3z, live);
        SKL_PHASE(3, 7);
        // the buffers of this step are rewritten two steps on, after the
        // next step's barrier
        return true;
    };
    const bool ok = yq_for_steps(a.nsteps, step, [&](int live) { yq_shift8(y, live); });
    if (!ok) {
        if (tid == 0) *a.done = 0;
        return;
    }
    __syncthreads();
    for (int t = tid; t < nv; t += NT) a.perm_out[t] = s_colat[t];
    if (tid == 0) *a.done = 1;
}

// changed[0..*nch) = ascending positions t with perm[t] != t (one workgroup)
// (none when a grid barrier gave up: perm may then hold stale entries)
__global__ __launch_bounds__(1024) void k_perm_changed(int nv, const int32_t *perm,
                                                       int32_t *changed, int32_t *nch,
                                                       const unsigned *err) {
    __shared__ int s_cnt[16];
    __shared__ int s_base;
    if (grid_failed(err)) {
        if (threadIdx.x == 0) *nch = 0;
        return;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) s_base = 0;
    __syncthreads();
    for (int t0 = 0; t0 < nv; t0 += 1024) {
        const int t = t0 + threadIdx.x;
        const bool f = t < nv && perm[t] != t;
        const unsigned long long bal = __ballot(f);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) s_cnt[wave] = __popcll(bal);
        __syncthreads();
        int off = s_base;
        for (int u = 0; u < wave; ++u) off += s_cnt[u];
        if (f) changed[off + before] = t;
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int u = 0; u < 16; ++u) tot += s_cnt[u];
            s_base += tot;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *nch = s_base;
}

// The block permutation of W (m rows), Y (mY rows) and jpvt in one gather
// and one scatter launch: blocks x < gw of row tile x take W, the others Y;
// block (0, k) also moves jpvt's entry k.
template <class T>
__global__ void k_perm_gather_all(long m, const T *Wc, long ldw, int mY, const T *Yc,
                                  const int32_t *jp, const int32_t *perm, const int32_t *changed,
                                  const int32_t *nch, unsigned gw, T *tmpW, T *tmpY,
                                  int32_t *tmpJ) {
    const int k = blockIdx.y;
    if (k >= *nch) return;
    const int src = perm[changed[k]];
    if (blockIdx.x < gw) {
        const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
        if (i < m) tmpW[i + (long)k * m] = Wc[i + (long)src * ldw];
        if (blockIdx.x == 0 && threadIdx.x == 0) tmpJ[k] = jp[src];
    } else {
        const int i = (int)(blockIdx.x - gw) * blockDim.x + threadIdx.x;
        if (i < mY) tmpY[i + (long)k * mY] = Yc[i + (long)src * mY];
    }
}
template <class T>
__global__ void k_perm_scatter_all(long m, T *Wc, long ldw, int mY, T *Yc, int32_t *jp,
                                   const int32_t *changed, const int32_t *nch, unsigned gw,
                                   const T *tmpW, const T *tmpY, const int32_t *tmpJ) {
    const int k = blockIdx.y;
    if (k >= *nch) return;
    const int dst = changed[k];
    if (blockIdx.x < gw) {
        const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
        if (i < m) Wc[i + (long)dst * ldw] = tmpW[i + (long)k * m];
        if (blockIdx.x == 0 && threadIdx.x == 0) jp[dst] = tmpJ[k];
    } else {
        const int i = (int)(blockIdx.x - gw) * blockDim.x + threadIdx.x;
        if (i < mY) Yc[i + (long)dst * mY] = tmpY[i + (long)k * mY];
    }
}


// Apply the panel's column swaps (s <-> swaps[s], in order, b <= 64) to the
// columns 0..b-1 of two row blocks.  The swaps are composed into one gather
// (new column t = old column src[t]); each workgroup moves a 64-row tile
// through LDS.  Workgroups [0, ceil(r1/64)) cover X1, the rest X2.  Nothing
// moves when a grid barrier gave up (the swaps may be stale).
template <class T>
__global__ __launch_bounds__(256) void k_apply_swaps(int b, const int32_t *swaps, long r1, T *X1,
                                                     long ld1, long r2, T *X2, long ld2,
                                                     const unsigned *err) {
    __shared__ int src[64];
    __shared__ T tile[64 * 65];
    if (grid_failed(err)) return;
    const long nb1 = (r1 + 63) / 64;
    long blk = blockIdx.x;
    T *X = X1;
    long ld = ld1, rows = r1;
    if (blk >= nb1) {
        blk -= nb1;
        X = X2;
        ld = ld2;
        rows = r2;
    }
    const long i0 = blk * 64;
    const int ni = (int)min(64L, rows - i0);
    // the block's rows of the b columns, every load in flight before the
    // first is used (clamped indices; rows past ni are never written back)
    constexpr int NQ = 64 * 64 / 256;
    T v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = threadIdx.x + q * 256, i = e & 63, t = e >> 6;
        v[q] = X[(i0 + min(i, ni - 1)) + (long)min(t, b - 1) * ld];
    }
    // wave 0: the column order from the sequential swaps, lane t holding
    // the source of column t (each swap exchanges two lanes' values)
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int sv = lane;
        const int sw = lane < b ? swaps[lane] : lane;
        for (int s = 0; s < b; ++s) {
            const int p = __builtin_amdgcn_readlane(sw, s);
            if (p != s) {
                const int a = __builtin_amdgcn_readlane(sv, s), c = __builtin_amdgcn_readlane(sv, p);
                sv = lane == s ? c : (lane == p ? a : sv);
            }
        }
        src[lane] = sv;
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int e = threadIdx.x + q * 256, i = e & 63, t = e >> 6;
        tile[i + t * 65] = v[q];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * b; e += 256) {
        int i = e & 63, t = e >> 6;
        if (i < ni && src[t] != t) X[(i0 + i) + (long)t * ld] = tile[i + src[t] * 65];
    }
}

// The small middle of the Y downdate, one workgroup per 16 rows of Y:
// BT = B0 Tm (B0 mY x b, Tm b x b upper, ld ldt) and B = G1(:, 0:b) - BT U11^T
// (U11 = the unit lower top b x b block of Vf, ld mp); both mY x b, ld mY.
// Tm and U11 pass through LDS 16 columns (rows) at a time, so the workgroup
// needs 24 KB of LDS in FP64 and fits beside the GEMM tiles of another stream.
// Thread (i = t & 15, c = t >> 4) forms entry (i, 16 s + c) of slice s.
template <class T>
__global__ __launch_bounds__(256) void k_yd_small(int mY, int b, const T *B0, const T *Tm, int ldt,
                                                  const T *G1, long ldg, const T *Vf, long mp,
                                                  T *BT, T *B) {
    // (row, col) of the 16 x 64 row blocks at row + 16 col; sS: the slice of
    // Tm (q + 64 c: Tm(q, 16 s + c)) or of U11 (c + 16 q: U11(16 s + c, q)).
    // Every global read is unconditional (clamped indices) and issued in a
    // batch before its first use: entries outside the triangles, the panel
    // or the rows are never read by the sums (q <= r < b) or stored (i <
    // rows), so they need no zeroing.
    __shared__ T sB0[16 * 64], sBT[16 * 64], sS[64 * 16];
    const int tid = threadIdx.x, i = tid & 15, c = tid >> 4;
    const int r0 = blockIdx.x * 16;
    const int rows = min(16, mY - r0);
    const int ic = min(i, rows - 1);
    T g1[4]; // G1(r0 + i, 16 s + c) for the four slices
    {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u, ii = e & 15, q = e >> 4;
            v[u] = B0[(r0 + min(ii, rows - 1)) + (long)min(q, b - 1) * mY];
            g1[u] = G1[(r0 + ic) + (long)min(16 * u + c, b - 1) * ldg];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u, ii = e & 15, q = e >> 4;
            sB0[ii + q * 16] = v[u];
        }
    }
    for (int s0 = 0; s0 < b; s0 += 16) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u, q = e & 63, cc = e >> 6;
            v[u] = Tm[q + (long)min(s0 + cc, b - 1) * ldt];
        }
        __syncthreads(); // sB0 stored / previous slice consumed
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u, q = e & 63, cc = e >> 6;
            sS[q + 64 * cc] = v[u];
        }
        __syncthreads();
        const int r = s0 + c;
        T acc = 0;
        for (int q = 0; q <= r && q < b; ++q) acc += sB0[i + q * 16] * sS[q + 64 * c];
        sBT[i + r * 16] = acc;
        if (i < rows && r < b) BT[(r0 + i) + (long)r * mY] = acc;
    }
    for (int s0 = 0; s0 < b; s0 += 16) {
        T v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u, cc = e & 15, q = e >> 4;
            v[u] = Vf[min(s0 + cc, b - 1) + (long)q * mp];
        }
        __syncthreads(); // sBT complete / previous slice consumed
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u, cc = e & 15, q = e >> 4;
            sS[cc + 16 * q] = v[u];
        }
        __syncthreads();
        // (BT U11^T)(i, r) = sum_{q <= r} BT(i, q) U11(r, q), U11(r, r) = 1
        const int r = s0 + c;
        T acc = 0;
        for (int q = 0; q < r && q < b; ++q) acc += sBT[i + q * 16] * sS[c + 16 * q];
        if (r < b) acc += sBT[i + r * 16];
        if (i < rows && r < b) B[(r0 + i) + (long)r * mY] = g1[s0 / 16] - acc;
    }
}

// the column-per-lane pivot kernel unless SKL_YQRP=rows selects the
// row-per-lane one
static bool yqrp_lanes_on() {
    static const bool on = [] {
        const char *e = getenv("SKL_YQRP");
        return !(e && std::string(e) == "rows");
    }();
    return on;
}

// the single-workgroup search over the leading columns first, unless
// SKL_YQRP_TOP=0 (read at every factorization)
static bool yqrp_top_on() {
    const char *e = getenv("SKL_YQRP_TOP");
    return !(e && std::string(e) == "0");
}

template <class T>
void cpqr_hqrrp(long m, long ncol, long nfac, T *W, long ldw, int32_t *jpvt, T *tau,
                const double *G_host, Arena &ar) {
    const int nb = 64, pp = 10, mY = nb + pp;
    hipStream_t s = ctx().stream;
    long n = nfac;
    long mn = std::min(m, n);
    // G (mY x m) and Y = G W(:, 0:n)
    T *G = ar.alloc_n<T>((size_t)mY * m);
    {
        std::vector<T> gh((size_t)mY * m);
        for (size_t i = 0; i < gh.size(); ++i) gh[i] = (T)G_host[i];
        SKL_HIP(hipMemcpyAsync(G, gh.data(), sizeof(T) * gh.size(), hipMemcpyHostToDevice, s));
        SKL_HIP(hipStreamSynchronize(s));
    }
    T *Y = ar.alloc_n<T>((size_t)mY * n);
    gemm<T>('N', 'N', mY, n, m, T(1), G, mY, W, ldw, T(0), Y, mY);
    T *Vf = ar.alloc_n<T>((size_t)m * nb);
    T *Tm = ar.alloc_n<T>((size_t)nb * nb);
    int32_t *perm = ar.alloc_n<int32_t>((size_t)n);
    int32_t *changed = ar.alloc_n<int32_t>((size_t)n);
    int32_t *nch = ar.alloc_n<int32_t>(4);
    int32_t *swaps = ar.alloc_n<int32_t>(nb);
    T *tmpc = ar.alloc_n<T>((size_t)std::max<long>(m, mY) * 2 * nb);
    T *tmpy = ar.alloc_n<T>((size_t)mY * 2 * nb);
    int32_t *tmpj = ar.alloc_n<int32_t>((size_t)2 * nb);
    T *B0 = ar.alloc_n<T>((size_t)mY * nb);
    T *BT = ar.alloc_n<T>((size_t)mY * nb);
    T *B = ar.alloc_n<T>((size_t)mY * nb);
    const int P_max = std::min(YQ_MAXP, barrier_grid_cap());
    static_assert(mY <= YQ_MAXROWS, "sketch rows exceed the register layout");
    SKL_REQUIRE(n <= (long)16 * 8 * P_max, SKL_E_UNSUPPORTED,
                "HQRRP pivot kernel: too many columns for one grid");
    T *cand = ar.alloc_n<T>((size_t)2 * P_max);
    int32_t *candp = ar.alloc_n<int32_t>((size_t)2 * P_max);
    T *ctau = ar.alloc_n<T>((size_t)2 * P_max);
    T *cvec = ar.alloc_n<T>((size_t)2 * P_max * YQ_MAXROWS);
    GridSync bar = grid_sync_fresh(ar); // (a fresh counter per pivot launch below)
    const bool top_on = yqrp_top_on();
    T *ynrm = ar.alloc_n<T>((size_t)n);
    int32_t *ydone = ar.alloc_n<int32_t>((size_t)(n / nb + 1));
    Context &c = ctx();
    const bool ahead = lookahead_on() && !ctx().no_alloc;
    if (ahead) c.make_side();
    bool side_busy = false;
    auto join_side = [&] {
        if (side_busy) SKL_HIP(hipStreamWaitEvent(c.stream, c.ev[1], 0));
        side_busy = false;
    };
    for (long j = 0; j < mn; j += nb) {
        int b = (int)std::min<long>(nb, std::min(n - j, m - j));
        bool last = (j + nb >= m) || (j + nb >= n);
        long nv = n - j;
        if (!last) {
            // pivot block from Y(:, j:n)
            YqArgs<T> ya;
            ya.mY = mY;
            ya.nv = (int)nv;
            ya.nsteps = b;
            ya.Y = Y + j * mY;
            ya.ldy = mY;
            ya.perm_out = perm;
            ya.cand = cand;
            ya.candp = candp;
            ya.ctau = ctau;
            ya.cvec = cvec;
            bar = grid_sync_fresh(ar);
            ya.bar = bar;
            if (mY == YQ_ROWS && yqrp_lanes_on()) {
                // the columns that can win the block in one workgroup first;
                // the grid-wide search runs only if that one stops
                if (top_on && nv <= YT_MAXNV) {
                    k_ynorms<T><<<(unsigned)((nv + 31) / 32), 256, 0, s>>>((int)nv, mY,
                                                                         Y + j * mY, mY, ynrm);
                    YtArgs<T> ta;
                    ta.mY = mY;
                    ta.nv = (int)nv;
                    ta.nsteps = b;
                    ta.Y = Y + j * mY;
                    ta.ldy = mY;
                    ta.nrm0 = ynrm;
                    ta.perm_out = perm;
                    ta.done = ydone + j / nb;
                    const size_t sht = (size_t)nv * sizeof(int32_t);
                    if (sht > 32768)
                        SKL_HIP(hipFuncSetAttribute((const void *)k_yqrp_top<T>,
                                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                                    (int)sht));
                    k_yqrp_top<T><<<1, 64 * YT_W, sht, s>>>(ta);
                    SKL_CHECK_LAUNCH();
                    ya.skip = ydone + j / nb;
                }
                // 512 columns per workgroup, one per lane
                constexpr int WL = 8;
                const int P = (int)((nv + 64 * WL - 1) / (64 * WL));
                size_t sh = (size_t)P * ((YQ_MAXROWS + 2) * sizeof(T) + sizeof(int));
                if (sh > 65536)
                    SKL_HIP(hipFuncSetAttribute((const void *)k_yqrp_lanes<T, WL>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                                (int)sh));
                BarrierLaunch guard(P);
                k_yqrp_lanes<T, WL><<<P, 64 * WL, sh, s>>>(ya);
                SKL_CHECK_LAUNCH();
            } else {
                // columns per workgroup: at most 32 workgroups (barrier latency
                // grows with the workgroup count) unless the columns need more
                // (128-column workgroups: 16 columns x 8 waves in FP64, whose
                // registers allow it, 8 x 16 in FP32)
                const bool dbl = sizeof(T) == 8;
                int K = nv <= 512 ? 1 : nv <= 1024 ? 2 : nv <= 2048 ? 4 : dbl ? 16 : 8;
                int Wv = (K == 16) ? 8 : 16;
                int P = (int)((nv + (long)Wv * K - 1) / ((long)Wv * K));
                constexpr int KW = sizeof(T) == 8 ? 16 : 8, WW = sizeof(T) == 8 ? 8 : 16;
                auto fn = K == 1 ? k_yqrp<T, 1, 16> : K == 2 ? k_yqrp<T, 2, 16>
                          : K == 4 ? k_yqrp<T, 4, 16> : k_yqrp<T, KW, WW>;
                size_t sh = (size_t)P * ((YQ_MAXROWS + 2) * sizeof(T) + sizeof(int));
                if (sh > 65536)
                    SKL_HIP(hipFuncSetAttribute((const void *)fn,
                                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                                (int)sh));
                BarrierLaunch guard(P);
                fn<<<P, 64 * Wv, sh, s>>>(ya);
                SKL_CHECK_LAUNCH();
            }
            // the previous block's update of the rows below its R12 ran on
            // the side stream beside this pivot search; the permutation
            // moves those rows
            join_side();
            k_perm_changed<<<1, 1024, 0, s>>>((int)nv, perm, changed, nch, bar.err);
            // permute W(:, j:n), Y(:, j:n), jpvt(j:n) by the new column order
            // (at most 2 b columns move: the b pivots and the ones they displace)
            const unsigned gwx = (unsigned)((m + 255) / 256), gyx = (unsigned)((mY + 255) / 256);
            dim3 gp(gwx + gyx, (unsigned)(2 * b));
            k_perm_gather_all<T><<<gp, 256, 0, s>>>(m, W + j * ldw, ldw, mY, Y + j * mY, jpvt + j,
                                                    perm, changed, nch, gwx, tmpc, tmpy, tmpj);
            k_perm_scatter_all<T><<<gp, 256, 0, s>>>(m, W + j * ldw, ldw, mY, Y + j * mY, jpvt + j,
                                                     changed, nch, gwx, tmpc, tmpy, tmpj);
            SKL_CHECK_LAUNCH();
        }
        // panel QRP of W(j:m, j:j+b), pivoting W(0:j, j:j+b) and Y(:, j:j+b)
        join_side();
        long mp = m - j;
        const int *cf = rowslice_qr<T>(mp, b, b, b, W + j + j * ldw, ldw, tau + j, jpvt + j,
                                       swaps, 1, ar, Tm);
        {
            unsigned nblk = (unsigned)((j + 63) / 64 + (mY + 63) / 64);
            k_apply_swaps<T><<<nblk, 256, 0, s>>>(b, swaps, j, W + j * ldw, ldw, mY, Y + j * mY,
                                                  mY, bar.err);
        }
        SKL_CHECK_LAUNCH();
        panel_t<T>(mp, b, W + j + j * ldw, ldw, tau + j, Vf, Tm, ar, cf);
        long nc2 = ncol - j - b;
        T *C2 = W + j + (j + b) * ldw;
        T *Wt2 = nullptr; // T^T V^T C2 when the rows below R12 are left to the side stream
        if (nc2 > 0 && ahead && !last && mp > b) {
            // C2 -= V (T^T (V^T C2)): the top b rows (R12, which the Y
            // downdate needs) here; the rows below on the side stream (below)
            T *Wt = ar.alloc_n<T>((size_t)b * nc2);
            Wt2 = ar.alloc_n<T>((size_t)b * nc2);
            gemm<T>('T', 'N', b, nc2, mp, T(1), Vf, mp, C2, ldw, T(0), Wt, b);
            gemm<T>('T', 'N', b, nc2, b, T(1), Tm, b, Wt, b, T(0), Wt2, b);
            gemm<T>('N', 'N', b, nc2, b, T(-1), Vf, mp, Wt2, b, T(1), C2, ldw);
        } else if (nc2 > 0) {
            apply_block_left_t<T>(mp, b, Vf, Tm, nc2, C2, ldw, ar);
        }
        if (!last) {
            // Y2 -= (G1 - (G V) T U11^T) R12 ; G(:, j:m) -= (G V) T V^T
            T *G1 = G + j * mY;
            gemm<T>('N', 'N', mY, b, mp, T(1), G1, mY, Vf, mp, T(0), B0, mY);
            k_yd_small<T><<<(unsigned)((mY + 15) / 16), 256, 0, s>>>(mY, b, B0, Tm, nb, G1, mY,
                                                                    Vf, mp, BT, B);
            SKL_CHECK_LAUNCH();
            long n2 = n - j - b;
            if (n2 > 0)
                gemm<T>('N', 'N', mY, n2, b, T(-1), B, mY, W + j + (j + b) * ldw, ldw, T(1),
                        Y + (j + b) * mY, mY);
            gemm<T>('N', 'T', mY, mp, b, T(-1), BT, mY, Vf, mp, T(1), G1, mY);
        }
        if (Wt2) {
            // the rows below R12 on the side stream, queued after the Y
            // downdate: the update (every CU's worth of GEMM tiles) then
            // overlaps the next pivot search, which holds a few CUs, instead
            // of keeping the downdate's small kernels from being scheduled
            SKL_HIP(hipEventRecord(c.ev[0], c.stream));
            OnSideStream on_side;
            SKL_HIP(hipStreamWaitEvent(c.stream, c.ev[0], 0));
            gemm<T>('N', 'N', mp - b, nc2, b, T(-1), Vf + b, mp, Wt2, b, T(1), C2 + b, ldw);
            SKL_HIP(hipEventRecord(c.ev[1], c.stream));
            side_busy = true;
        }
    }
    join_side();
}
template void cpqr_hqrrp<double>(long, long, long, double *, long, int32_t *, double *,
                                 const double *, Arena &);
template void cpqr_hqrrp<float>(long, long, long, float *, long, int32_t *, float *,
                                const double *, Arena &);

} // namespace skl
