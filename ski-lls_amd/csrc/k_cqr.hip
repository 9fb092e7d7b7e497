// k_cqr.hip -- Householder panel factorization by CholeskyQR2 and Householder
// reconstruction, with the HQRRP panel pivoting rule.
//
// The panel kernels of k_qr.hip factor a tall panel A_p (mp x b, b <= 64) one
// column at a time, one grid-wide exchange per column.  Here the panel's
// Householder factors come from b-column-wide steps instead:
//
//   1. G = A_p^T A_p (GEMM), then a pivoted Cholesky of G in one workgroup:
//      pivots by the largest partial norm with HQRRP's downdating
//      (hqrrp/lapack_compatible_sources/NoFLA_HQRRP_WY_blk_var4_skint.c:
//      920-990: vn1 *= sqrt((1 + t)(1 - t)), t = |r_kj| / vn1, recomputed
//      when the downdate loses accuracy), so the column order is the one the
//      column-by-column panel QRP picks, in exact arithmetic;
//   2. Q1 = A_p P R1^-1 (the triangle inverted in LDS, then a 64-column
//      product), then one more CholeskyQR pass, Q = Q1 R2^-1, R = R2 R1:
//      orthonormal to working precision while cond(A_p) stays far below
//      1/sqrt(u);
//   3. Householder reconstruction (Ballard, Demmel, Grigori, Jacquelin,
//      Knight, Nguyen, "Reconstructing Householder vectors from tall-skinny
//      QR", JPDC 85, 2015): the LU factorization without pivoting of
//      Q - [S; 0], with S_kk = -sign of the k-th pivot, gives the unit lower
//      trapezoidal V = L, U and tau_k = -U_kk S_kk, and the R of the
//      Householder QR is S R -- the factors the panel's dgeqrf / QRP would
//      store, in exact arithmetic -- and the compact WY factor of the
//      reflectors is T = -U S L1^-T (L1 the top block of V), so the
//      trailing update needs no V^T V product.
//
// A panel whose Cholesky meets a partial norm below 1e-4 of the column's own
// norm (near dependence: the Gram matrix no longer resolves it) or a pivot
// that is not positive sets a device flag; the fast path's later kernels then
// do nothing and the column-by-column kernel, launched after them with the
// flag as its condition, factors the untouched panel instead.
#include "k_common.h"
#include "k_tri64.h"

#ifndef SKL_PHASE
#define SKL_PHASE(kernel, mark)
#endif

namespace skl {

// Pivoted (PIV) or plain Cholesky of the b x b Gram matrix G (ld ldg) into
// R (upper, ld 64) and Z = P R^-1 (b x b, ld b; A_p Z = A_p P R^-1).  pass 1:
// perm (final position -> panel column) and the sequential swaps; flag = 0 /
// 1 (failed).  pass 2 (plain): runs only when flag is 0 and sets it on
// failure.
// Blocked, 16 steps at a time, with no barrier inside a block: wave 0 runs
// the steps, lane p for panel column p (rows and columns stay where they are;
// the lane keeps the column's logical position, partial norms, Schur
// diagonal and its entries of the block's rows of R).  A step reads column pk
// of the Schur complement as it was at the block's start (LDS) and subtracts
// the block's earlier rows of R (the pivot's entries by v_readlane); after
// the block the four waves update the Schur complement by those 16 rows on
// the matrix cores.
template <bool PIV>
__global__ __launch_bounds__(CQ_T) void k_cqr_chol(int b, long mp, const double *G, int ldg,
                                                  double *R, double *Z, int32_t *perm,
                                                  int32_t *swp, int *flag, int hq_norms,
                                                  int pass) {
    __shared__ double Rb[64 * CQ_LD]; // row k of R by panel column; later X = R^-1
    __shared__ double Rr[64 * CQ_LD]; // R in logical order
    __shared__ double Ss[64 * CQ_LD]; // Schur complement at the start of the block
    __shared__ int s_perm[64];
    __shared__ int s_k;               // steps done (the first failed step when short)
    const int tid = threadIdx.x, p = tid & 63, g = tid >> 6;
    if (pass == 2 && *flag != 0) return;
    SKL_PHASE(4, 0);
    tile64_fill(
        Ss, [&](int i, int j) { return G[min(i, b - 1) + (long)min(j, b - 1) * ldg]; },
        [&](int i, int j) { return i < b && j < b; });
    for (int e = tid; e < 64 * CQ_LD; e += CQ_T) Rb[e] = 0.0;
    // wave 0 runs the steps, lane p for panel column p: its Schur diagonal,
    // partial norms, logical position, whether it is still active, and its
    // entries of the block's rows of R
    double dg = 0.0, g0 = 0.0, vn1 = 0.0, vn2 = 0.0;
    int pos = p < b ? p : 0x7fffffff;
    bool mine = p < b;
    const double tol3z = sqrt(1.1102230246251565e-16);
    if (tid == 0) s_k = 0;
    __syncthreads();
    SKL_PHASE(4, 1);
    if (g == 0 && p < b) {
        dg = Ss[p + p * CQ_LD];
        g0 = dg;
        vn1 = sqrt(fmax(dg, 0.0));
        vn2 = vn1;
    }
    int k = 0;
    for (int k0 = 0; k0 < b; k0 += 16) {
        if (g == 0) {
            double rl[16], pre[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                rl[t] = 0.0;
                // unpivoted: the block's columns of the Schur complement, all
                // read before the first step (the pivoted steps read theirs
                // once the pivot is known)
                if (!PIV) pre[t] = Ss[p + min(k0 + t, 63) * CQ_LD];
            }
            bool ok = true;
            // the steps run without a branch on ok: after a failed step the
            // values are garbage, but the block ends and the panel is given up
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int kk = k0 + t;
                if (kk < b) {
                    int pk = kk;
                    if (PIV) {
                        // largest partial norm among the columns left, first
                        // logical position on ties
                        const double bv = (mine && vn1 == vn1) ? fabs(vn1) : -1.0;
                        int bl = wave_argmax(bv, pos);
                        if (bl < 0) bl = __builtin_ctzll(__ballot(pos == kk)); // no finite candidate
                        pk = bl;
                        // logical position of the pivot (pk is wave-uniform)
                        const int lp = __builtin_amdgcn_readlane(pos, pk);
                        const bool atk = pos == kk;
                        if (p == pk)
                            pos = kk;
                        else if (atk)
                            pos = lp;
                        if (p == 0 && swp) swp[kk] = lp;
                    }
                    const double g0k = readlane(g0, pk);
                    if (p == pk) mine = false;
                    // row pk of the current Schur complement: the block's start
                    // minus this block's earlier rows of R
                    double sc = PIV ? Ss[p + pk * CQ_LD] : pre[t];
#pragma unroll
                    for (int u = 0; u < t; ++u) sc = fma(-readlane(rl[u], pk), rl[u], sc);
                    const double d = readlane(sc, pk);
                    // not positive, or the column's partial norm below 1e-4 of its norm
                    ok = ok && d > 0.0 && d >= 1e-8 * g0k; // the same in every lane
                    const double inv = rsqrt_nr(d), rkk = d * inv;
                    const double rp = sc * inv; // R(kk, column p) for an active p
                    rl[t] = p == pk ? rkk : (mine ? rp : 0.0);
                    if (mine) {
                        dg -= rp * rp;
                        if (PIV && vn1 != 0.0) {
                            // partial norm downdating with the new row of R
                            // (quotients and square roots from rcp_nr and
                            // rsqrt_nr; both outcomes formed, one kept)
                            double temp = fabs(rp) * rcp_nr(vn1);
                            if (hq_norms)
                                temp = (1.0 + temp) * (1.0 - temp);
                            else
                                temp = 1.0 - temp * temp;
                            if (temp < 0.0) temp = 0.0;
                            const double q5 = vn1 * rcp_nr(vn2);
                            const double temp2 = temp * q5 * q5;
                            const double rec = (mp - kk - 1 > 0) ? sqrt_nr(fmax(dg, 0.0)) : 0.0;
                            const double dwn = vn1 * sqrt_nr(temp);
                            if (temp2 <= tol3z) {
                                vn1 = rec;
                                vn2 = rec;
                            } else {
                                vn1 = dwn;
                            }
                        }
                    }
                }
            }
            // the block's rows of R, by panel column
#pragma unroll
            for (int t = 0; t < 16; ++t)
                if (k0 + t < b) Rb[(k0 + t) + p * CQ_LD] = rl[t];
            if (p == 0) {
                int done = k0;
#pragma unroll
                for (int t = 0; t < 16; ++t) done += (k0 + t < b) ? 1 : 0;
                s_k = ok ? done : -1;
            }
            SKL_PHASE(4, 2);
        }
        __syncthreads();
        const int kd = s_k;
        if (kd < 0) break;
        k = kd;
        if (k >= b) break;
        // Schur complement of the next block: S -= R(k0:k0+16, :)^T R(k0:k0+16, :)
        // on the matrix cores, 16 x 16 tiles, four per wave
        {
            typedef double acc4 __attribute__((ext_vector_type(4)));
            const int lane = p, li = lane & 15, lk = lane >> 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int tile = g * 4 + q, ti = tile & 3, tj = tile >> 2;
                acc4 c;
#pragma unroll
                for (int r = 0; r < 4; ++r) c[r] = Ss[(16 * ti + lk + 4 * r) + (16 * tj + li) * CQ_LD];
#pragma unroll
                for (int u0 = 0; u0 < 16; u0 += 4)
                    c = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        -Rb[(k0 + u0 + lk) + (16 * ti + li) * CQ_LD],
                        Rb[(k0 + u0 + lk) + (16 * tj + li) * CQ_LD], c, 0, 0, 0);
                // each tile is read and written by its own wave only
#pragma unroll
                for (int r = 0; r < 4; ++r) Ss[(16 * ti + lk + 4 * r) + (16 * tj + li) * CQ_LD] = c[r];
            }
        }
        __syncthreads();
        SKL_PHASE(4, 3);
    }
    // k: the steps done, b unless a pivot failed (pos is wave 0's)
    const int fail = k < b ? 1 : 0;
    if (tid == 0) {
        if (pass == 1)
            *flag = fail;
        else if (fail)
            *flag = 1;
    }
    if (fail) return;
    if (g == 0 && p < b) s_perm[pos] = p; // logical position -> panel column
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += CQ_T) {
        const int i = e & 63, j = e >> 6;
        const double v = (i <= j && j < b) ? Rb[i + s_perm[j] * CQ_LD] : (i == j ? 1.0 : 0.0);
        Rr[i + j * CQ_LD] = v;
        R[i + j * 64] = (j < b) ? ((i <= j) ? v : 0.0) : 0.0;
    }
    if (PIV && perm && tid < b) perm[tid] = s_perm[tid];
    __syncthreads();
    SKL_PHASE(4, 4);
    // Z = P R^-1: row perm[t] of Z is row t of R^-1
    tri_inv_upper_lds(Rr, Rb, Ss);
    SKL_PHASE(4, 5);
    for (int e = tid; e < b * b; e += CQ_T) {
        const int t = e % b, j = e / b;
        Z[s_perm[t] + (long)j * b] = Rb[t + j * CQ_LD];
    }
    SKL_PHASE(4, 6);
}

// Out^T Out of a 64-row tile held in LDS (ld CQ_LD, zero past the rows and
// columns): the transpose goes to T (LDS scratch) and the 64 x 64 product to
// P (ld 64) on the matrix cores.  All CQ_T threads; T is free after return.
__device__ inline void gram_tile_store(const double *O, double *T, double *P) {
    for (int e = threadIdx.x; e < 64 * 64; e += CQ_T) {
        const int i = e & 63, k = e >> 6;
        T[i + k * CQ_LD] = O[k + i * CQ_LD]; // T(i, k) = O(k, i)
    }
    __syncthreads();
    mm64_lds(T, O, [&](int r, int c, double v) { P[r + 64 * c] = v; });
}

// Partial Gram matrices of a tall panel (rows x b, b <= 64), one per 64-row
// tile: P[t] = A(tile t, :)^T A(tile t, :) (64 x 64, zero past b), summed by
// k_gram_reduce in tile order.  Replaces the split-K GEMM for the CholeskyQR
// Gram matrices: a 64-row tile is one 64 x 64 x 64 product on MFMA.
__global__ __launch_bounds__(CQ_T) void k_gram_tiles(long rows, int b, const double *A, long lda,
                                                     double *P, const int *flag) {
    __shared__ double sA[64 * CQ_LD];
    __shared__ double sT[64 * CQ_LD];
    if (flag && *flag != 0) return;
    const long r0 = (long)blockIdx.x * 64;
    tile64_fill(
        sA,
        [&](int i, int j) {
            return A[(r0 + i < rows ? r0 + i : rows - 1) + (long)min(j, b - 1) * lda];
        },
        [&](int i, int j) { return j < b && r0 + i < rows; });
    __syncthreads();
    gram_tile_store(sA, sT, P + (size_t)blockIdx.x * 64 * 64);
}

// G (b x b, ld b) = sum over the nt tiles of P[t] in increasing t: eight
// threads per entry, each over every eighth tile (loads eight at a time),
// combined in a fixed order by lane moves.
__global__ __launch_bounds__(256) void k_gram_reduce(int nt, int b, const double *P, double *G,
                                                     const int *flag) {
    if (flag && *flag != 0) return;
    const int e = blockIdx.x * 32 + (threadIdx.x >> 3), sub = threadIdx.x & 7;
    double s = 0.0;
    if (e < 64 * 64) {
        int t = sub;
        for (; t + 56 < nt; t += 64) {
            double w[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) w[q] = P[(size_t)(t + 8 * q) * 4096 + e];
#pragma unroll
            for (int q = 0; q < 8; ++q) s += w[q];
        }
        for (; t < nt; t += 8) s += P[(size_t)t * 4096 + e];
    }
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    s += __shfl_xor(s, 4, 64);
    if (e < 64 * 64 && sub == 0) {
        const int i = e & 63, j = e >> 6;
        if (i < b && j < b) G[i + (long)j * b] = s;
    }
}

// Out(i, :) = In(i, :) X for rows i < rows (X b x b, ld b): 64-row tiles of
// In in LDS, thread t computes row t & 63 for every fourth column from
// t >> 6.  Runs only when *flag is 0.
__global__ __launch_bounds__(CQ_T) void k_cqr_rmul(long rows, int b, const double *In, long ldi,
                                                   const double *X, double *Out, long ldo,
                                                   const int *flag, double *Pg) {
    __shared__ double sI[64 * CQ_LD];
    __shared__ double sX[64 * CQ_LD];
    __shared__ double sO[64 * CQ_LD]; // the output tile, for its Gram
    if (*flag != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, grp = tid >> 6;
    const long r0 = (long)blockIdx.x * 64;
    tile64_fill2(
        sX, [&](int i, int j) { return X[min(i, b - 1) + (long)min(j, b - 1) * b]; },
        [&](int i, int j) { return i < b && j < b; }, sI,
        [&](int i, int j) {
            return In[(r0 + i < rows ? r0 + i : rows - 1) + (long)min(j, b - 1) * ldi];
        },
        [&](int i, int j) { return j < b && r0 + i < rows; });
    __syncthreads();
    const long row = r0 + lane;
    // the row in registers (zero past b), then 16 columns per thread
    double xi[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) xi[k] = sI[lane + k * CQ_LD];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int j = grp + 4 * m;
        double acc = 0.0;
#pragma unroll
        for (int k = 0; k < 64; ++k) acc += xi[k] * sX[k + j * CQ_LD];
        if (row < rows && j < b) Out[row + (long)j * ldo] = acc;
        sO[lane + j * CQ_LD] = (row < rows && j < b) ? acc : 0.0;
    }
    if (!Pg) return;
    // this tile's part of Out^T Out (64 x 64) for k_gram_reduce
    __syncthreads();
    gram_tile_store(sO, sI, Pg + (size_t)blockIdx.x * 64 * 64);
}

// Householder reconstruction of the panel from Q = Q1 Z2 (Q1 mp x b, ld ldq;
// Z2 = R2^-1, b x b upper, ld b) and R = R2 R1: only the top block of Q is
// formed here (the rows below never are: the final product applies Z2 U^-1
// to Q1).  The LU of the top block of Q - [S; 0] (S_kk = -sign of the k-th
// pivot) gives L1 (below the diagonal of the panel's top block), U and
// tau_k = -U_kk S_kk; the top block's upper triangle receives S R.  Outputs:
// Zc = Z2 U^-1 (b x b, ld b) for V2 = Q1(b:, :) Zc, and, when Tout is not
// null, the compact WY factor of the reflectors, T = -U S L1^-T (upper, b x b,
// ld b): (I - V T V^T)[S; 0] = Q means Q - [S; 0] = V (-T L1^T S), an LU
// factorization with unit lower V.  The panel's column order goes into jpvt
// (entries permuted) and swaps.  Runs only when *flag is 0.  Thread t owns
// row t & 63 and every fourth column from t >> 6 on: one barrier per step.
__global__ __launch_bounds__(CQ_T) void k_cqr_hr(int b, const double *Q1, long ldq,
                                                 const double *Z2, const double *R1,
                                                 const double *R2, const int32_t *perm,
                                                 const int32_t *swp, double *Wp, long ldw,
                                                 double *tau, double *Tout, double *Zc,
                                                 int32_t *jpvt, int32_t *swaps, const int *flag) {
    __shared__ double Ub[64 * CQ_LD];
    __shared__ double Lb[64 * CQ_LD];
    __shared__ double Ys[64 * CQ_LD];
    __shared__ double Xs[64 * CQ_LD];
    __shared__ double sg[64];
    __shared__ int jp[64];
    if (*flag != 0) return;
    SKL_PHASE(5, 0);
    const int tid = threadIdx.x, i = tid & 63, g = tid >> 6;
    // the top block of Q = Q1(0:64, :) Z2 (Q1's top block in Ys, Z2 in Lb,
    // the product into Xs)
    tile64_fill2(
        Ys, [&](int r, int c) { return Q1[min(r, b - 1) + (long)min(c, b - 1) * ldq]; },
        [&](int r, int c) { return r < b && c < b; }, Lb,
        [&](int r, int c) { return Z2[min(r, b - 1) + (long)min(c, b - 1) * b]; },
        [&](int r, int c) { return r < b && c < b && r <= c; });
    __syncthreads();
    mm64_lds(Ys, Lb, [&](int r, int c, double v) { Xs[r + c * CQ_LD] = v; });
    __syncthreads();
    for (int e = tid; e < 64 * CQ_LD; e += CQ_T) {
        Ub[e] = 0.0;
        Lb[e] = 0.0;
    }
    __syncthreads();
    SKL_PHASE(5, 1);
    // LU without pivoting of Xs - S, blocked by 16 steps: wave 0 runs a block
    // with no barrier, lane p keeping L(p, k) and U(k, p) of the block's steps;
    // step k reads column k and row k of the block-start matrix (LDS) and
    // subtracts the block's earlier terms (the other factor's entries by
    // v_readlane); after the block the four waves update the matrix by
    // L(:, block) U(block, :) on the matrix cores
    for (int k0 = 0; k0 < b; k0 += 16) {
        if (g == 0) {
            double lc[16], ur[16], pc[16], pr[16];
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                lc[t] = 0.0;
                ur[t] = 0.0;
                // the block's columns and rows of the block-start matrix,
                // all read before the first step
                const int k = min(k0 + t, 63);
                pc[t] = Xs[i + k * CQ_LD]; // column k, row i
                pr[t] = Xs[k + i * CQ_LD]; // row k, column i
            }
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int k = k0 + t;
                if (k < b) {
                    double cv = pc[t];
                    double rv = pr[t];
#pragma unroll
                    for (int u = 0; u < t; ++u) {
                        cv = fma(-lc[u], readlane(ur[u], k), cv);
                        rv = fma(-readlane(lc[u], k), ur[u], rv);
                    }
                    const double akk = readlane(rv, k);
                    const double sk = akk >= 0.0 ? -1.0 : 1.0;
                    const double piv = akk - sk;
                    const double l = cv * rcp_nr(piv);
                    lc[t] = (i > k && i < b) ? l : 0.0;
                    ur[t] = (i > k && i < b) ? rv : 0.0;
                    if (i > k && i < b) Lb[i + k * CQ_LD] = l;
                    if (i >= k && i < b) Ub[k + i * CQ_LD] = i == k ? piv : rv;
                    if (i == 0) sg[k] = sk;
                }
            }
            SKL_PHASE(5, 2);
        }
        __syncthreads();
        if (k0 + 16 >= b) break;
        // Xs -= L(:, k0:k0+16) U(k0:k0+16, :): 16 x 16 tiles, four per wave,
        // each read and written by its own wave only
        {
            typedef double acc4 __attribute__((ext_vector_type(4)));
            const int li = i & 15, lk = i >> 4;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int tile = g * 4 + q, ti = tile & 3, tj = tile >> 2;
                acc4 c;
#pragma unroll
                for (int r = 0; r < 4; ++r) c[r] = Xs[(16 * ti + lk + 4 * r) + (16 * tj + li) * CQ_LD];
#pragma unroll
                for (int u0 = 0; u0 < 16; u0 += 4)
                    c = __builtin_amdgcn_mfma_f64_16x16x4f64(
                        -Lb[(16 * ti + li) + (k0 + u0 + lk) * CQ_LD],
                        Ub[(k0 + u0 + lk) + (16 * tj + li) * CQ_LD], c, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) Xs[(16 * ti + lk + 4 * r) + (16 * tj + li) * CQ_LD] = c[r];
            }
        }
        __syncthreads();
        SKL_PHASE(5, 3);
    }
    if (tid < b) tau[tid] = -Ub[tid + tid * CQ_LD] * sg[tid];
    // top block: S R2 R1 on and above the diagonal, L1 below (R2 in Xs, R1
    // in Ys; both upper with zeros elsewhere, ld 64)
    tile64_fill2(
        Xs, [&](int r, int c) { return R2[r + 64 * c]; }, [](int, int) { return true; }, Ys,
        [&](int r, int c) { return R1[r + 64 * c]; }, [](int, int) { return true; });
    __syncthreads();
    mm64_lds(Xs, Ys, [&](int r, int c, double v) {
        if (r < b && c < b) Wp[r + (long)c * ldw] = r <= c ? sg[r] * v : Lb[r + c * CQ_LD];
    });
    if (jpvt) {
        if (tid < b) jp[tid] = jpvt[perm[tid]];
        __syncthreads();
        if (tid < b) jpvt[tid] = jp[tid];
    }
    if (swaps && tid < b) swaps[tid] = swp[tid];
    __syncthreads(); // L1 (Lb) and R1 (Ys) read
    SKL_PHASE(5, 4);
    if (Tout) {
        // T = -U S X with X = L1^-T: L1^T (unit upper) into Ys, its inverse
        // into Xs (Lb as the inversion's scratch)
        for (int e = tid; e < 64 * 64; e += CQ_T) {
            const int r = e & 63, c = e >> 6;
            Ys[r + c * CQ_LD] = r < c ? Lb[c + r * CQ_LD] : (r == c ? 1.0 : 0.0);
        }
        __syncthreads();
        tri_inv_upper_lds(Ys, Xs, Lb);
        // S X into Ys (rows scaled), then T = -U (S X)
        for (int e = tid; e < 64 * 64; e += CQ_T) {
            const int r = e & 63, c = e >> 6;
            Ys[r + c * CQ_LD] = (r < b ? sg[r] : 1.0) * Xs[r + c * CQ_LD];
        }
        __syncthreads();
        mm64_lds(Ub, Ys, [&](int r, int c, double v) {
            if (r < b && c < b) Tout[r + (long)c * b] = r <= c ? -v : 0.0;
        });
        __syncthreads();
    }
    SKL_PHASE(5, 5);
    // U^-1, then Zc = Z2 U^-1 (Z2 into Ys)
    for (int e = tid; e < 64 * 64; e += CQ_T) {
        const int r = e & 63, c = e >> 6;
        if (c >= b) Ub[r + c * CQ_LD] = (r == c) ? 1.0 : 0.0;
    }
    __syncthreads();
    tri_inv_upper_lds(Ub, Lb, Ys);
    SKL_PHASE(5, 6);
    tile64_fill(
        Ys, [&](int r, int c) { return Z2[min(r, b - 1) + (long)min(c, b - 1) * b]; },
        [&](int r, int c) { return r < b && c < b && r <= c; });
    __syncthreads();
    mm64_lds(Ys, Lb, [&](int r, int c, double v) {
        if (r < b && c < b) Zc[r + (long)c * b] = r <= c ? v : 0.0;
    });
    SKL_PHASE(5, 7);
}

// Queues the fast path for the panel Wp (mp x b, ld ldw): pivot != 0 uses the
// HQRRP panel pivoting (hq_norms) and fills jpvt (may be null) and swaps;
// Tout (b x b, ld b, may be null) receives the reflectors' T factor.
// Returns the device flag that is nonzero at run time when the fast path gave
// up and left the panel untouched.
int *cqr_panel(long mp, int b, int pivot, double *Wp, long ldw, double *tau, int32_t *jpvt,
               int32_t *swaps, int hq_norms, double *Tout, Arena &ar) {
    hipStream_t s = ctx().stream;
    double *G = ar.alloc_n<double>(64 * 64);
    double *R1 = ar.alloc_n<double>(64 * 64);
    double *R2 = ar.alloc_n<double>(64 * 64);
    double *Z1 = ar.alloc_n<double>(64 * 64);
    double *Z2 = ar.alloc_n<double>(64 * 64);
    double *Zc = ar.alloc_n<double>(64 * 64);
    double *Q1 = ar.alloc_n<double>((size_t)mp * b);
    int32_t *perm = ar.alloc_n<int32_t>(64);
    int32_t *swp = ar.alloc_n<int32_t>(64);
    int *flag = ar.alloc_n<int>(1);
    const unsigned gr = (unsigned)((mp + 63) / 64);
    // the Gram matrices from per-tile partials (one 64-row tile per workgroup)
    double *Pg = ar.alloc_n<double>((size_t)gr * 64 * 64);
    k_gram_tiles<<<gr, CQ_T, 0, s>>>(mp, b, Wp, ldw, Pg, nullptr);
    k_gram_reduce<<<128, 256, 0, s>>>((int)gr, b, Pg, G, nullptr);
    if (pivot)
        k_cqr_chol<true><<<1, CQ_T, 0, s>>>(b, mp, G, b, R1, Z1, perm, swp, flag, hq_norms, 1);
    else
        k_cqr_chol<false><<<1, CQ_T, 0, s>>>(b, mp, G, b, R1, Z1, nullptr, nullptr, flag, 0, 1);
    k_cqr_rmul<<<gr, CQ_T, 0, s>>>(mp, b, Wp, ldw, Z1, Q1, mp, flag, Pg);
    k_gram_reduce<<<128, 256, 0, s>>>((int)gr, b, Pg, G, flag);
    k_cqr_chol<false><<<1, CQ_T, 0, s>>>(b, mp, G, b, R2, Z2, nullptr, nullptr, flag, 0, 2);
    k_cqr_hr<<<1, CQ_T, 0, s>>>(b, Q1, mp, Z2, R1, R2, perm, swp, Wp, ldw, tau, Tout, Zc,
                                 pivot ? jpvt : nullptr, pivot ? swaps : nullptr, flag);
    if (mp > b) {
        const unsigned g2 = (unsigned)((mp - b + 63) / 64);
        k_cqr_rmul<<<g2, CQ_T, 0, s>>>(mp - b, b, Q1 + b, mp, Zc, Wp + b, ldw, flag, nullptr);
    }
    SKL_CHECK_LAUNCH();
    return flag;
}

} // namespace skl
