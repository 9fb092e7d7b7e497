// k_dht.hip -- randomized discrete Hartley transform and hashing sketch.
//
// The reference applies y = (1/sqrt n) D .* A(:, c), an unnormalised FFTW
// DHT, then the hashing matrix S (src/fast_unitary_transform.cpp:88-100,
// src/fftw_r2r.cpp:45-50, src/sketching.cpp:39-62).  Here two real columns
// share one complex FFT (z = x1 + i x2; with Z_k = a+ib, Z_{n-k} = c+id:
// H1_k = ((a+c) - (b-d))/2, H2_k = ((b+d) + (a-c))/2).  The FFT is a
// mixed-radix Stockham DIF held in LDS; each stage keeps its inputs in
// registers between two barriers so one LDS buffer suffices.
//
//  * n <= LDS capacity: one workgroup per column pair does load -> FFT ->
//    DHT -> bucket sums straight from LDS -> SA, reading A once.
//  * larger smooth n = n1 n2: four-step FFT (pass 1: length-n1 FFTs on
//    strided rows + twiddles; pass 2: length-n2 FFTs on pairs of mirrored
//    rows, which also yields the DHT), then the bucket gather.
//  * other n (a prime factor above 5): Bluestein's algorithm on a
//    {2,3,5}-smooth length of at least 2n - 1.
//
// The LDS FFT uses radices 2, 3, 4, 5 and 8, eight complex elements per
// thread per stage, and 512 (FP64) / 1024 (FP32) threads per workgroup: at
// most about 140 VGPRs, so several waves per SIMD hide the memory latency.
// Loads are batched (all of a thread's elements in flight before the first
// use), and workgroups that share cache lines are mapped to one XCD.
//
// Bucket sums run with contraction off, in increasing source row, matching
// the reference's scatter order (src/SpMat.cpp:53-55).
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>

#include "fft_roots.h"
#include "k_common.h"

#ifndef SKL_PHASE
#define SKL_PHASE(kernel, mark)
#endif

namespace skl {

constexpr int FFT_MAXE = 8;    // complex elements per thread held per stage
constexpr int MAX_STAGES = 24;

struct FftDims {
    int L;                  // transform length
    int nst;                // stages
    int radix[MAX_STAGES];
};

template <class T> struct FftPlanDev {
    long n = 0;
    int mode = 0;           // 1 single pass, 2 four-step
    int n1 = 0, n2 = 0;
    FftDims d1{}, d2{};     // single: d1 only
    cplx<T> *tw1 = nullptr; // W_{n1}^e
    cplx<T> *tw2 = nullptr; // W_{n2}^e
    cplx<T> *twn = nullptr; // W_n^e (four-step twiddles)
    int J2 = 0, K1 = 0;     // sequences per workgroup in the two passes
    int G = 1;              // column pairs per pass-2 workgroup (and per H group)
};

// ---------------------------------------------------------------------------
// radix-R DIF stage on S interleaved sequences in LDS (element j of sequence s
// at buf[j*S + s]); len = current length, stride = product of earlier radices.
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void dft4(cplx<T> p0, cplx<T> p1, cplx<T> p2, cplx<T> p3, cplx<T> &o0,
                                     cplx<T> &o1, cplx<T> &o2, cplx<T> &o3) {
    cplx<T> t0 = cadd(p0, p2), t1 = csub(p0, p2), t2 = cadd(p1, p3), t3 = csub(p1, p3);
    o0 = cadd(t0, t2);
    o2 = csub(t0, t2);
    o1 = {t1.x + t3.y, t1.y - t3.x}; // t1 - i t3
    o3 = {t1.x - t3.y, t1.y + t3.x}; // t1 + i t3
}

// radix-5 DFT with W = exp(-2 pi i / 5): y_k = sum_j x_j W^{jk}, from the
// sums and differences of the mirrored inputs (t1 = x1 + x4, t2 = x2 + x3,
// d1 = x1 - x4, d2 = x2 - x3): y_0 = x0 + t1 + t2; y_{1,4} = a -/+ i c and
// y_{2,3} = b -/+ i e with a = x0 + c1 t1 + c2 t2, b = x0 + c2 t1 + c1 t2,
// c = s1 d1 + s2 d2, e = s2 d1 - s1 d2 (c_k = cos(2 pi k / 5), s_k =
// sin(2 pi k / 5)): 12 multiplies instead of the 25 complex products of the
// direct sum.
template <class T>
__device__ __forceinline__ void dft5(const cplx<T> (&x)[5], cplx<T> (&y)[5]) {
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    const cplx<T> t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
    const cplx<T> d1 = csub(x[1], x[4]), d2 = csub(x[2], x[3]);
    y[0] = cadd(x[0], cadd(t1, t2));
    const cplx<T> a = {x[0].x + c1 * t1.x + c2 * t2.x, x[0].y + c1 * t1.y + c2 * t2.y};
    const cplx<T> b = {x[0].x + c2 * t1.x + c1 * t2.x, x[0].y + c2 * t1.y + c1 * t2.y};
    const cplx<T> c = {s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y};
    const cplx<T> e = {s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y};
    // -i c = (c.y, -c.x)
    y[1] = {a.x + c.y, a.y - c.x};
    y[4] = {a.x - c.y, a.y + c.x};
    y[2] = {b.x + e.y, b.y - e.x};
    y[3] = {b.x - e.y, b.y + e.x};
}

// S and stride enter as FastDiv: the butterfly's (sequence, group) indices
// come from a multiply-high rather than integer division.
template <class T, int R, int NT>
__device__ __forceinline__ void fft_stage(cplx<T> *buf, int L, const FastDiv &fS, int len,
                                          const FastDiv &fst, const cplx<T> *__restrict__ tw) {
    constexpr int NB = (FFT_MAXE + R - 1) / R;
    const int S = (int)fS.d, stride = (int)fst.d;
    const int m = len / R;
    const int nbfly = (L / R) * S;
    const int tws = L / len;
    cplx<T> v[NB][R];
    int bs[NB], bq[NB], bp[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        int id = threadIdx.x + b * NT;
        const int t = (int)fS.div((unsigned)id), tq = (int)fst.div((unsigned)t);
        bs[b] = id - t * S;
        bq[b] = t - tq * stride;
        bp[b] = tq;
        if (id < nbfly) {
            const int s = bs[b], q = bq[b], p = bp[b];
#pragma unroll
            for (int j = 0; j < R; ++j) v[b][j] = buf[(q + stride * (p + j * m)) * S + s];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        int id = threadIdx.x + b * NT;
        if (id < nbfly) {
            const int s = bs[b], q = bq[b], p = bp[b];
            cplx<T> y[R];
            if constexpr (R == 2) {
                y[0] = cadd(v[b][0], v[b][1]);
                y[1] = csub(v[b][0], v[b][1]);
            } else if constexpr (R == 4) {
                dft4(v[b][0], v[b][1], v[b][2], v[b][3], y[0], y[1], y[2], y[3]);
            } else if constexpr (R == 5) {
                dft5(v[b], y);
            } else if constexpr (R == 8) {
                // even/odd split: y_k = a_k + W8^k b_k, y_{k+4} = a_k - W8^k b_k
                cplx<T> a[4], c[4];
                dft4(v[b][0], v[b][2], v[b][4], v[b][6], a[0], a[1], a[2], a[3]);
                dft4(v[b][1], v[b][3], v[b][5], v[b][7], c[0], c[1], c[2], c[3]);
                const T h = (T)0.70710678118654752440;
                cplx<T> w1 = {h * (c[1].x + c[1].y), h * (c[1].y - c[1].x)};   // (1-i)/sqrt2
                cplx<T> w2 = {c[2].y, -c[2].x};                                 // -i
                cplx<T> w3 = {h * (c[3].y - c[3].x), -h * (c[3].x + c[3].y)};  // -(1+i)/sqrt2
                y[0] = cadd(a[0], c[0]);
                y[4] = csub(a[0], c[0]);
                y[1] = cadd(a[1], w1);
                y[5] = csub(a[1], w1);
                y[2] = cadd(a[2], w2);
                y[6] = csub(a[2], w2);
                y[3] = cadd(a[3], w3);
                y[7] = csub(a[3], w3);
            } else {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    T sr = 0, si = 0;
#pragma unroll
                    for (int j = 0; j < R; ++j) {
                        const int e = (j * k) % R;
                        const T cr = (T)RootTab<R>::c[e], ci = (T)RootTab<R>::s[e];
                        sr += v[b][j].x * cr - v[b][j].y * ci;
                        si += v[b][j].x * ci + v[b][j].y * cr;
                    }
                    y[k] = {sr, si};
                }
            }
#pragma unroll
            for (int k = 0; k < R; ++k) {
                cplx<T> o = y[k];
                if (k > 0) {
                    // p < len / R and k < R, so p k < len: no reduction
                    o = cmul(o, tw[p * k * tws]);
                }
                buf[(q + stride * (R * p + k)) * S + s] = o;
            }
        }
    }
    __syncthreads();
}

template <class T, int NT>
__device__ void lds_fft(cplx<T> *buf, int S, const FftDims &D, const cplx<T> *__restrict__ tw) {
    int len = D.L, stride = 1;
    const FastDiv fS((unsigned)S);
    for (int st = 0; st < D.nst; ++st) {
        int r = D.radix[st];
        const FastDiv fst((unsigned)stride);
        switch (r) {
        case 2: fft_stage<T, 2, NT>(buf, D.L, fS, len, fst, tw); break;
        case 3: fft_stage<T, 3, NT>(buf, D.L, fS, len, fst, tw); break;
        case 4: fft_stage<T, 4, NT>(buf, D.L, fS, len, fst, tw); break;
        case 5: fft_stage<T, 5, NT>(buf, D.L, fS, len, fst, tw); break;
        case 8: fft_stage<T, 8, NT>(buf, D.L, fS, len, fst, tw); break;
        default: break;
        }
        len /= r;
        stride *= r;
    }
}

template <class T> __device__ inline T load_scaled(const double *X, long idx, const double *D,
                                                   long j, double scale) {
    // (scale * D[j]) * A(j, c) as in fast_unitary_transform.cpp:91
    double f = D ? scale * D[j] : scale;
    return (T)(f * X[idx]);
}
template <> __device__ inline float load_scaled<float>(const double *X, long idx, const double *D,
                                                       long j, double scale) {
    double f = D ? scale * D[j] : scale;
    return (float)f * (float)X[idx];
}
template <class T> __device__ inline T scaled(double f, double x) { return (T)(f * x); }
template <> __device__ inline float scaled<float>(double f, double x) { return (float)f * (float)x; }

// The scaled entries (scale D[j]) A(j, c) of a column pair (c1 only when
// has1; zero otherwise) at the rows row(u) >= 0 of the thread's FFT_MAXE
// elements (zero where row(u) < 0), the same products as load_scaled.  The
// raw values go to registers first, each load behind nothing but its row
// test, and the products follow: a product next to its load would make the
// compiler wait on each load in turn, one memory latency per element.
template <class T, class ROW>
__device__ inline void load_pair_scaled(const double *__restrict__ X, long ldx, long c0, long c1,
                                        bool has1, const double *__restrict__ D, double scale,
                                        ROW row, cplx<T> (&v)[FFT_MAXE]) {
    double x0[FFT_MAXE], x1[FFT_MAXE], dd[FFT_MAXE];
    const long c1c = has1 ? c1 : c0; // a valid column either way
#pragma unroll
    for (int u = 0; u < FFT_MAXE; ++u) {
        const long j = row(u);
        x0[u] = 0.0;
        x1[u] = 0.0;
        dd[u] = 0.0;
        if (j >= 0) {
            x0[u] = X[j + c0 * ldx];
            x1[u] = X[j + c1c * ldx];
            if (D) dd[u] = D[j];
        }
    }
#pragma unroll
    for (int u = 0; u < FFT_MAXE; ++u) {
        const double f = D ? scale * dd[u] : scale;
        v[u].x = scaled<T>(f, x0[u]);
        v[u].y = has1 ? scaled<T>(f, x1[u]) : T(0);
    }
}

// ---------------------------------------------------------------------------
// Single pass: one workgroup per column pair (c0 = 2 blockIdx.x, c1 = c0 + 1).
// out_mode 0: hash into SA (m x ncols); 1: write the DHT into Y (n x ncols).
// ---------------------------------------------------------------------------
template <class T, int NT>
__global__ __launch_bounds__(NT) void k_dht_single(long n, long ncols, const double *__restrict__ X,
                                                   long ldx, const double *__restrict__ D,
                                                   double scale, FftDims dims,
                                                   const cplx<T> *__restrict__ tw, int out_mode,
                                                   long m, const int32_t *__restrict__ bptr,
                                                   const int32_t *__restrict__ bsrc,
                                                   const T *__restrict__ bcoef, T *__restrict__ out,
                                                   long ldo) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const long c0 = 2L * blockIdx.x, c1 = c0 + 1;
    const bool has1 = c1 < ncols;
    {
        cplx<T> v[FFT_MAXE]; // n <= NT * FFT_MAXE; all loads in flight together
        load_pair_scaled<T>(X, ldx, c0, c1, has1, D, scale, [&](int u) {
            const long j = threadIdx.x + (long)u * NT;
            return j < n ? j : -1L;
        }, v);
#pragma unroll
        for (int u = 0; u < FFT_MAXE; ++u) {
            const long j = threadIdx.x + (long)u * NT;
            if (j < n) buf[j] = v[u];
        }
    }
    __syncthreads();
    lds_fft<T, NT>(buf, 1, dims, tw);
    // DHT extraction through registers
    T h1[FFT_MAXE], h2[FFT_MAXE];
#pragma unroll
    for (int e = 0; e < FFT_MAXE; ++e) {
        long k = threadIdx.x + (long)e * NT;
        if (k < n) {
            long nk = (k == 0) ? 0 : n - k;
            cplx<T> a = buf[k], c = buf[nk];
            h1[e] = T(0.5) * ((a.x + c.x) - (a.y - c.y));
            h2[e] = T(0.5) * ((a.y + c.y) + (a.x - c.x));
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < FFT_MAXE; ++e) {
        long k = threadIdx.x + (long)e * NT;
        if (k < n) buf[k] = {h1[e], h2[e]};
    }
    __syncthreads();
    if (out_mode == 1) {
        for (long k = threadIdx.x; k < n; k += NT) {
            out[k + c0 * ldo] = buf[k].x;
            if (has1) out[k + c1 * ldo] = buf[k].y;
        }
        return;
    }
    {
#pragma clang fp contract(off)
        for (long h = threadIdx.x; h < m; h += NT) {
            T s1 = 0, s2 = 0;
            for (int q = bptr[h]; q < bptr[h + 1]; ++q) {
                cplx<T> hv = buf[bsrc[q]];
                T cf = bcoef[q];
                s1 = s1 + cf * hv.x;
                s2 = s2 + cf * hv.y;
            }
            out[h + c0 * ldo] = s1;
            if (has1) out[h + c1 * ldo] = s2;
        }
    }
}

constexpr int TW4_LO = 256; // low part of the split four-step twiddle exponent

// ---------------------------------------------------------------------------
// Four-step, pass 1: workgroup (pair, block of J2 consecutive j2):
// Y[k1*n2 + j2] = W_n^{j2 k1} * FFT_{n1}( x[j1*n2 + j2] )[k1]
// ---------------------------------------------------------------------------
template <class T, int NT>
__global__ __launch_bounds__(NT, 2048 / NT) void k_fft4_pass1(long n, int n1, int n2, int J2, long ncols,
                                                   const double *__restrict__ X, long ldx,
                                                   const double *__restrict__ D, double scale,
                                                   FftDims dims, const cplx<T> *__restrict__ tw1,
                                                   const cplx<T> *__restrict__ twn,
                                                   cplx<T> *__restrict__ Y) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const long lb = xcd_local_block(blockIdx.x + (long)gridDim.x * blockIdx.y,
                                    (long)gridDim.x * gridDim.y);
    const long pair = lb / gridDim.x;
    const long c0 = 2 * pair, c1 = c0 + 1;
    const bool has1 = c1 < ncols;
    const int j2b = (int)(lb % gridDim.x) * J2;
    const int nJ = min(J2, n2 - j2b);
    const int tot = n1 * J2; // <= NT * FFT_MAXE
    const FastDiv fJ((unsigned)J2);
    SKL_PHASE(6, 0);
    // every load of the thread is issued before the first is used: the
    // elements are independent, and a loop that consumed each load before
    // issuing the next would pay the memory latency once per element
    cplx<T> v[FFT_MAXE];
    load_pair_scaled<T>(X, ldx, c0, c1, has1, D, scale, [&](int u) {
        const int e = threadIdx.x + u * NT;
        const int j1 = (int)fJ.div((unsigned)e), jj = e - j1 * J2;
        return (e < tot && jj < nJ) ? (long)j1 * n2 + j2b + jj : -1L;
    }, v);
#pragma unroll
    for (int u = 0; u < FFT_MAXE; ++u) {
        const int e = threadIdx.x + u * NT;
        if (e < tot) buf[e] = v[u]; // e = j1 J2 + jj
    }
    // the stage twiddles in LDS after the data: an LDS read per butterfly
    // output instead of a cache round trip per stage; then the two halves of
    // the four-step twiddles W_n^a = W_n^{256 (a >> 8)} W_n^{a & 255}: the
    // elements' exponents are scattered over all of W_n, and a read of the
    // full table per element moved a cache line from L2 for 16 bytes
    cplx<T> *twl = buf + tot;
    cplx<T> *twlo = twl + n1, *twhi = twlo + TW4_LO;
    const int nhi = (int)((n + TW4_LO - 1) / TW4_LO);
    for (int e = threadIdx.x; e < n1; e += NT) twl[e] = tw1[e];
    for (int e = threadIdx.x; e < TW4_LO; e += NT) twlo[e] = twn[e];
    for (int e = threadIdx.x; e < nhi; e += NT) twhi[e] = twn[(long)TW4_LO * e];
    __syncthreads();
    SKL_PHASE(6, 1);
    lds_fft<T, NT>(buf, J2, dims, twl);
    SKL_PHASE(6, 2);
    cplx<T> *Yp = Y + pair * n;
#pragma unroll
    for (int u = 0; u < FFT_MAXE; ++u) {
        const int e = threadIdx.x + u * NT;
        const int k1 = (int)fJ.div((unsigned)e), jj = e - k1 * J2;
        if (e < tot && jj < nJ) {
            // j2 k1 <= (n2 - 1)(n1 - 1) < n: no reduction needed
            const int a = (j2b + jj) * k1;
            const cplx<T> w = cmul(twhi[a / TW4_LO], twlo[a % TW4_LO]);
            Yp[(long)k1 * n2 + j2b + jj] = cmul(buf[k1 * J2 + jj], w);
        }
    }
    SKL_PHASE(6, 3);
}

// ---------------------------------------------------------------------------
// Four-step, pass 2: workgroup (group of G pairs, block of K1 primary rows k1
// in [0, n1/2] plus their mirrors n1 - k1, of each pair):
// X[k1 + n1 k2] = FFT_{n2}(Y[k1, :])[k2], then H at k and n - k from the
// primary/mirror rows.
// ---------------------------------------------------------------------------
// pairs_out: H holds the group's transforms interleaved by pair, each pair's
// two columns side by side (H[2 ((grp n + k) G + g) + {0, 1}] for pair
// G grp + g), for the bucket gather of 2 G columns in one line per member;
// otherwise column c at H + c ldh.
template <class T, int NT>
__global__ __launch_bounds__(NT) void k_fft4_pass2(long n, int n1, int n2, int K1, int G,
                                                   long ncols, FftDims dims,
                                                   const cplx<T> *__restrict__ tw2,
                                                   const cplx<T> *__restrict__ Y,
                                                   T *__restrict__ H, long ldh, int pairs_out) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const long lb = xcd_local_block(blockIdx.x + (long)gridDim.x * blockIdx.y,
                                    (long)gridDim.x * gridDim.y);
    const long grp = lb / gridDim.x;
    const long npairs = (ncols + 1) / 2;
    const int half = n1 / 2;                 // primary rows 0..half
    const int kb = (int)(lb % gridDim.x) * K1;
    const int nK = min(K1, half + 1 - kb);   // primary rows in this block
    const int S2 = 2 * K1;                   // a pair's slots: [0,K1) primary, [K1,2K1) mirror
    const int S = S2 * G;                    // sequence slots: pair g at [g S2, (g + 1) S2)
    const int tot = n2 * S; // <= NT * FFT_MAXE
    const FastDiv fS((unsigned)S), fS2((unsigned)S2);
    SKL_PHASE(7, 0);
    {
        cplx<T> v[FFT_MAXE]; // all loads in flight before the first use
#pragma unroll
        for (int u = 0; u < FFT_MAXE; ++u) {
            const int e = threadIdx.x + u * NT;
            const int j2 = (int)fS.div((unsigned)e), sl = e - j2 * S;
            const int g = (int)fS2.div((unsigned)sl), w = sl - g * S2;
            const int kk = w < K1 ? w : w - K1;
            const long pair = grp * G + g;
            v[u] = {T(0), T(0)};
            if (e < tot && kk < nK && pair < npairs) {
                const int k1 = kb + kk;
                const int row = (w < K1 || k1 == 0) ? k1 : n1 - k1;
                v[u] = Y[pair * n + (long)row * n2 + j2];
            }
        }
#pragma unroll
        for (int u = 0; u < FFT_MAXE; ++u) {
            const int e = threadIdx.x + u * NT;
            if (e < tot) buf[e] = v[u]; // e = j2 S + sl
        }
    }
    cplx<T> *twl = buf + tot; // the stage twiddles in LDS after the data
    for (int e = threadIdx.x; e < n2; e += NT) twl[e] = tw2[e];
    __syncthreads();
    SKL_PHASE(7, 1);
    lds_fft<T, NT>(buf, S, dims, twl);
    SKL_PHASE(7, 2);
    // outputs: for primary row k1 and column k2: k = k1 + n1 k2,
    //   partner n-k = (n1-k1) + n1 (n2-1-k2) (mirror slot) for k1 > 0,
    //               = n1 ((n2-k2) % n2)      (same slot)   for k1 = 0.
    // mirror row k1' = n1 - k1 (k1 > 0, k1' != k1): k' = k1' + n1 k2,
    //   partner = k1 + n1 (n2-1-k2) (primary slot).
    for (int e = threadIdx.x; e < tot; e += NT) {
        const int k2 = (int)fS.div((unsigned)e), sl = e - k2 * S;
        const int g = (int)fS2.div((unsigned)sl), w = sl - g * S2;
        const int kk = w < K1 ? w : w - K1;
        const long pair = grp * G + g;
        if (kk >= nK || pair >= npairs) continue;
        int k1 = kb + kk;
        bool mirror = w >= K1;
        if (mirror && (k1 == 0 || 2 * k1 == n1)) continue; // self-mirrored rows done once
        int row = mirror ? n1 - k1 : k1;
        cplx<T> a = buf[k2 * S + sl];
        cplx<T> c;
        if (k1 == 0) {
            c = buf[(k2 == 0 ? 0 : n2 - k2) * S + sl];
        } else if (2 * k1 == n1) {
            c = buf[(n2 - 1 - k2) * S + sl];
        } else {
            int osl = g * S2 + (mirror ? kk : kk + K1);
            c = buf[(n2 - 1 - k2) * S + osl];
        }
        long k = row + (long)n1 * k2;
        T h1 = T(0.5) * ((a.x + c.x) - (a.y - c.y));
        T h2 = T(0.5) * ((a.y + c.y) + (a.x - c.x));
        const long c0 = 2 * pair, c1 = c0 + 1;
        const bool has1 = c1 < ncols;
        if (pairs_out) {
            reinterpret_cast<cplx<T> *>(H)[(grp * n + k) * G + g] = {h1, has1 ? h2 : T(0)};
        } else {
            H[k + c0 * ldh] = h1;
            if (has1) H[k + c1 * ldh] = h2;
        }
    }
    SKL_PHASE(7, 3);
}

// ---------------------------------------------------------------------------
// Bucket gather: Y[h, c] = sum_{q in bucket h} coef[q] X[bsrc[q], c]
// ---------------------------------------------------------------------------
// One workgroup per (column, 256 buckets); the workgroups of a column run on
// one XCD, so the column's random reads are served by that XCD's L2 after
// one fetch of the column.
template <class T>
__global__ __launch_bounds__(256) void k_hash_gather(long m, long ncols, long nbh,
                                                     const int32_t *__restrict__ bptr,
                                                     const int32_t *__restrict__ bsrc,
                                                     const T *__restrict__ bcoef,
                                                     const T *__restrict__ X, long ldx,
                                                     T *__restrict__ Y, long ldy) {
#pragma clang fp contract(off)
    const long lb = xcd_local_block(blockIdx.x, gridDim.x);
    const long c = lb / nbh;
    const long h = (lb % nbh) * 256 + threadIdx.x;
    if (h >= m || c >= ncols) return;
    const int q0 = bptr[h], q1 = bptr[h + 1];
    const T *x = X + c * ldx;
    T s = 0;
    int q = q0;
    for (; q + 4 <= q1; q += 4) { // four reads in flight, summed in source order
        const int i0 = bsrc[q], i1 = bsrc[q + 1], i2 = bsrc[q + 2], i3 = bsrc[q + 3];
        const T w0 = bcoef[q], w1 = bcoef[q + 1], w2 = bcoef[q + 2], w3 = bcoef[q + 3];
        const T x0 = x[i0], x1 = x[i1], x2 = x[i2], x3 = x[i3];
        s = s + w0 * x0;
        s = s + w1 * x1;
        s = s + w2 * x2;
        s = s + w3 * x3;
    }
    for (; q < q1; ++q) s = s + bcoef[q] * x[bsrc[q]];
    Y[h + c * ldy] = s;
}

// The same for the grouped pair layout of pass 2 (X2[(grp n + i) G + g] =
// (column 2 p, column 2 p + 1) of row i for pair p = G grp + g): one bucket
// member's row of a group is 16 G contiguous bytes, one line read for 2 G
// columns (a lone 16-byte read per member and pair moved a whole line from
// L2 for 16 bytes, and the gather ran at the rate of those lines).  NTH
// buckets per workgroup; the workgroups of a group are consecutive on one
// XCD.  A bucket's members go in batches -- the indices and coefficients
// first (the last member repeated past the end), then the rows -- and each
// column's sum keeps the source order.
constexpr int GATHER_NTH = 64, GATHER_GMAX = 10;
template <class T, int G>
__global__ __launch_bounds__(GATHER_NTH) void k_hash_gather_groups(
    long m, long n, long ncols, long nbh, const int32_t *__restrict__ bptr,
    const int32_t *__restrict__ bsrc, const T *__restrict__ bcoef,
    const cplx<T> *__restrict__ X2, T *__restrict__ Y, long ldy) {
#pragma clang fp contract(off)
    const long ngrp = ((ncols + 1) / 2 + G - 1) / G;
    const long lb = xcd_local_block(blockIdx.x, gridDim.x);
    const long gr = lb / nbh;
    const long h = (lb % nbh) * GATHER_NTH + threadIdx.x;
    if (h >= m || gr >= ngrp) return;
    const int q0 = bptr[h], q1 = bptr[h + 1];
    const cplx<T> *x = X2 + gr * n * G;
    T s0[G], s1[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        s0[g] = 0;
        s1[g] = 0;
    }
    constexpr int MB = G <= 5 ? 4 : 2; // members per batch (registers: MB G rows)
    for (int q = q0; q < q1; q += MB) {
        int ix[MB];
        T w[MB];
#pragma unroll
        for (int k = 0; k < MB; ++k) {
            const int qk = min(q + k, q1 - 1);
            ix[k] = bsrc[qk];
            w[k] = bcoef[qk];
        }
        cplx<T> xv[MB][G];
#pragma unroll
        for (int k = 0; k < MB; ++k)
#pragma unroll
            for (int g = 0; g < G; ++g) xv[k][g] = x[(long)ix[k] * G + g];
#pragma unroll
        for (int k = 0; k < MB; ++k)
            if (q + k < q1) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
                    s0[g] = s0[g] + w[k] * xv[k][g].x;
                    s1[g] = s1[g] + w[k] * xv[k][g].y;
                }
            }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const long c = 2 * (gr * G + g);
        if (c < ncols) Y[h + c * ldy] = s0[g];
        if (c + 1 < ncols) Y[h + (c + 1) * ldy] = s1[g];
    }
}

template <class T>
void hash_gather(long m, long ncols, const int32_t *bptr, const int32_t *bsrc, const T *bcoef,
                 const T *X, long ldx, T *Y, long ldy) {
    if (m <= 0 || ncols <= 0) return;
    const long nbh = (m + 255) / 256;
    for (long c0 = 0; c0 < ncols;) { // 1-D grids of at most 2^31 - 1 workgroups
        const long nc = std::min<long>(ncols - c0, 2147483647L / nbh);
        k_hash_gather<T><<<(unsigned)(nbh * nc), 256, 0, ctx().stream>>>(
            m, nc, nbh, bptr, bsrc, bcoef, X + c0 * ldx, ldx, Y + c0 * ldy, ldy);
        SKL_CHECK_LAUNCH();
        c0 += nc;
    }
}
template void hash_gather<double>(long, long, const int32_t *, const int32_t *, const double *,
                                  const double *, long, double *, long);
template void hash_gather<float>(long, long, const int32_t *, const int32_t *, const float *,
                                 const float *, long, float *, long);

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
static bool factor_dims(long L, FftDims &d) {
    d.L = (int)L;
    d.nst = 0;
    long x = L;
    const int pref[] = {8, 4, 2, 3, 5};
    for (int r : pref) {
        while (x % r == 0) {
            if (d.nst >= MAX_STAGES) return false;
            d.radix[d.nst++] = r;
            x /= r;
        }
    }
    return x == 1;
}

template <class T> static std::vector<cplx<T>> twiddle_values(long L) {
    std::vector<cplx<T>> h((size_t)L);
    for (long e = 0; e < L; ++e) {
        long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)e /
                          (long double)L;
        h[(size_t)e] = {(T)cosl(ang), (T)sinl(ang)};
    }
    return h;
}

template <class T> static cplx<T> *make_twiddles(long L) {
    std::vector<cplx<T>> h = twiddle_values<T>(L);
    cplx<T> *dptr = nullptr;
    AllocLock lk(capture_mutex());
    SKL_HIP(hipMalloc(&dptr, sizeof(cplx<T>) * (size_t)std::max<long>(L, 1)));
    SKL_HIP(hipMemcpy(dptr, h.data(), sizeof(cplx<T>) * (size_t)L, hipMemcpyHostToDevice));
    return dptr;
}

// the same table in arena memory (no-allocation mode; not cached)
template <class T> static cplx<T> *arena_twiddles(long L, Arena &ar) {
    std::vector<cplx<T>> h = twiddle_values<T>(L);
    cplx<T> *dptr = ar.alloc_n<cplx<T>>((size_t)std::max<long>(L, 1));
    hipStream_t s = ctx().stream;
    SKL_HIP(hipMemcpyAsync(dptr, h.data(), sizeof(cplx<T>) * (size_t)L, hipMemcpyHostToDevice, s));
    SKL_HIP(hipStreamSynchronize(s)); // h goes out of scope
    return dptr;
}

template <class T> constexpr int fft_nt() { return sizeof(T) == 8 ? 512 : 1024; }
template <class T> constexpr long lds_elems() { return (long)fft_nt<T>() * FFT_MAXE; } // 4096 / 8192

// Lengths the LDS FFT runs directly: radices 2, 3, 4, 5, 8 only.  Larger
// radices (a generic R-point DFT per butterfly) would set the register
// allocation of every FFT kernel -- one kernel serves all lengths -- and cut
// its occupancy; lengths with other prime factors go through Bluestein.
static bool smooth(long x) {
    for (int r : {2, 3, 5})
        while (x % r == 0) x /= r;
    return x == 1;
}

// ---------------------------------------------------------------------------
// Generic batched complex FFT (in place), used by Bluestein's algorithm for
// lengths with a prime factor above 13.  Optional pointwise pre-op
// (multiply by a table, then conjugate) and post-op (conjugate, scale,
// multiply by a table; only the first nout outputs are written to `out`).
// ---------------------------------------------------------------------------
template <class T> struct CfftPlan {
    long L = 0;
    int mode = 0;            // 1 single pass, 2 four-step
    int L1 = 0, L2 = 0, J2 = 0, K1 = 0;
    FftDims d1{}, d2{};
    cplx<T> *tw1 = nullptr, *tw2 = nullptr, *twL = nullptr;
};

template <class T> struct CfftOps {
    const cplx<T> *pre_tab = nullptr;
    int pre_conj = 0;
    int post_conj = 0;
    T post_scale = T(1);
    const cplx<T> *post_tab = nullptr;
};

template <class T>
__device__ inline cplx<T> pre_op(cplx<T> v, long j, const CfftOps<T> &o) {
    if (o.pre_tab) v = cmul(v, o.pre_tab[j]);
    if (o.pre_conj) v.y = -v.y;
    return v;
}
template <class T>
__device__ inline cplx<T> post_op(cplx<T> v, long k, const CfftOps<T> &o) {
    if (o.post_conj) v.y = -v.y;
    v.x *= o.post_scale;
    v.y *= o.post_scale;
    if (o.post_tab) v = cmul(v, o.post_tab[k]);
    return v;
}

template <class T, int NT>
__global__ __launch_bounds__(NT) void k_cfft_single(int L, const cplx<T> *__restrict__ Z, long zs,
                                                    FftDims dims, const cplx<T> *__restrict__ tw,
                                                    CfftOps<T> ops, cplx<T> *__restrict__ out,
                                                    long os, long nout) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const cplx<T> *z = Z + (long)blockIdx.x * zs;
    for (int j = threadIdx.x; j < L; j += NT) buf[j] = pre_op(z[j], j, ops);
    __syncthreads();
    lds_fft<T, NT>(buf, 1, dims, tw);
    cplx<T> *o = out + (long)blockIdx.x * os;
    for (long k = threadIdx.x; k < nout; k += NT) o[k] = post_op(buf[k], k, ops);
}

template <class T, int NT>
__global__ __launch_bounds__(NT) void k_cfft4_p1(long L, int L1, int L2, int J2,
                                                 const cplx<T> *__restrict__ Z, long zs, FftDims dims,
                                                 const cplx<T> *__restrict__ tw1,
                                                 const cplx<T> *__restrict__ twL, CfftOps<T> ops,
                                                 cplx<T> *__restrict__ Y) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const long b = blockIdx.y;
    const int j2b = blockIdx.x * J2;
    const int nJ = min(J2, L2 - j2b);
    const cplx<T> *z = Z + b * zs;
    for (int e = threadIdx.x; e < L1 * J2; e += NT) {
        int jj = e % J2, j1 = e / J2;
        cplx<T> v = {T(0), T(0)};
        if (jj < nJ) {
            long j = (long)j1 * L2 + j2b + jj;
            v = pre_op(z[j], j, ops);
        }
        buf[j1 * J2 + jj] = v;
    }
    __syncthreads();
    lds_fft<T, NT>(buf, J2, dims, tw1);
    cplx<T> *y = Y + b * L;
    for (int e = threadIdx.x; e < L1 * J2; e += NT) {
        int jj = e % J2, k1 = e / J2;
        if (jj < nJ) {
            int j2 = j2b + jj;
            y[(long)k1 * L2 + j2] = cmul(buf[k1 * J2 + jj], twL[((long)j2 * k1) % L]);
        }
    }
}

template <class T, int NT>
__global__ __launch_bounds__(NT) void k_cfft4_p2(long L, int L1, int L2, int K1,
                                                 const cplx<T> *__restrict__ Y, FftDims dims,
                                                 const cplx<T> *__restrict__ tw2, CfftOps<T> ops,
                                                 cplx<T> *__restrict__ out, long os, long nout) {
    extern __shared__ unsigned char smem_raw[];
    cplx<T> *buf = reinterpret_cast<cplx<T> *>(smem_raw);
    const long b = blockIdx.y;
    const int kb = blockIdx.x * K1;
    const int nK = min(K1, L1 - kb);
    const cplx<T> *y = Y + b * L;
    for (int e = threadIdx.x; e < L2 * K1; e += NT) {
        int kk = e % K1, j2 = e / K1;
        cplx<T> v = {T(0), T(0)};
        if (kk < nK) v = y[(long)(kb + kk) * L2 + j2];
        buf[j2 * K1 + kk] = v;
    }
    __syncthreads();
    lds_fft<T, NT>(buf, K1, dims, tw2);
    cplx<T> *o = out + b * os;
    for (int e = threadIdx.x; e < L2 * K1; e += NT) {
        int kk = e % K1, k2 = e / K1;
        if (kk >= nK) continue;
        long k = (kb + kk) + (long)L1 * k2;
        if (k < nout) o[k] = post_op(buf[k2 * K1 + kk], k, ops);
    }
}

static bool smooth5(long x) {
    for (int r : {2, 3, 5})
        while (x % r == 0) x /= r;
    return x == 1;
}

template <class T> static bool make_cplan(long L, CfftPlan<T> &p) {
    const long cap = lds_elems<T>();
    p.L = L;
    if (L <= cap) {
        p.mode = 1;
        if (!factor_dims(L, p.d1)) return false;
        p.tw1 = make_twiddles<T>(L);
        return true;
    }
    long best = -1;
    double bs = 1e300;
    for (long a = 2; a <= L; ++a) {
        if (L % a) continue;
        long b = L / a;
        if (a > cap / 4 || b > cap / 4) continue;
        if (!smooth(a) || !smooth(b)) continue;
        double sc = std::fabs(std::log((double)a) - std::log((double)b));
        if (sc < bs) {
            bs = sc;
            best = a;
        }
    }
    if (best < 0) return false;
    p.mode = 2;
    p.L1 = (int)best;
    p.L2 = (int)(L / best);
    factor_dims(p.L1, p.d1);
    factor_dims(p.L2, p.d2);
    p.tw1 = make_twiddles<T>(p.L1);
    p.tw2 = make_twiddles<T>(p.L2);
    p.twL = make_twiddles<T>(L);
    p.J2 = (int)std::min<long>(32, cap / p.L1);
    p.K1 = (int)std::min<long>(32, cap / p.L2);
    return true;
}

// Batched in-place-capable complex FFT: out[b][k] = post(FFT(pre(Z[b])))[k], k < nout.
template <class T>
static void cfft(const CfftPlan<T> &P, long nseq, const cplx<T> *Z, long zs, const CfftOps<T> &ops,
                 cplx<T> *out, long os, long nout, Arena &ar) {
    constexpr int NT = fft_nt<T>();
    hipStream_t s = ctx().stream;
    if (nseq <= 0) return;
    if (P.mode == 1) {
        size_t sh = sizeof(cplx<T>) * (size_t)P.L;
        k_cfft_single<T, NT><<<(unsigned)nseq, NT, sh, s>>>((int)P.L, Z, zs, P.d1, P.tw1, ops, out,
                                                            os, nout);
    } else {
        cplx<T> *Y = ar.alloc_n<cplx<T>>((size_t)nseq * P.L);
        dim3 g1((unsigned)((P.L2 + P.J2 - 1) / P.J2), (unsigned)nseq);
        k_cfft4_p1<T, NT><<<g1, NT, sizeof(cplx<T>) * (size_t)P.L1 * P.J2, s>>>(
            P.L, P.L1, P.L2, P.J2, Z, zs, P.d1, P.tw1, P.twL, ops, Y);
        dim3 g2((unsigned)((P.L1 + P.K1 - 1) / P.K1), (unsigned)nseq);
        k_cfft4_p2<T, NT><<<g2, NT, sizeof(cplx<T>) * (size_t)P.L2 * P.K1, s>>>(
            P.L, P.L1, P.L2, P.K1, Y, P.d2, P.tw2, ops, out, os, nout);
    }
    SKL_CHECK_LAUNCH();
}

// Bluestein: X_k = w_k conj( FFT_L( conj( FFT_L(a) .* B ) ) )_k / L with
// a_j = z_j w_j (j < n, zero-padded to L), w_j = exp(-i pi j^2 / n),
// B = FFT_L(b), b_j = b_{L-j} = conj(w_j).
template <class T> struct BsPlan {
    long n = 0, L = 0;
    CfftPlan<T> cp;
    cplx<T> *w = nullptr; // n
    cplx<T> *B = nullptr; // L
};

template <class T>
__global__ void k_bs_pack(long n, long L, long ncols, const double *__restrict__ X, long ldx,
                          const double *__restrict__ D, double scale, long pair0,
                          const cplx<T> *__restrict__ w, cplx<T> *__restrict__ Z) {
    const long p = blockIdx.y;
    const long c0 = 2 * (pair0 + p), c1 = c0 + 1;
    const bool has1 = c1 < ncols;
    for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < L; j += (long)gridDim.x * blockDim.x) {
        cplx<T> v = {T(0), T(0)};
        if (j < n) {
            T re = load_scaled<T>(X, j + c0 * ldx, D, j, scale);
            T im = has1 ? load_scaled<T>(X, j + c1 * ldx, D, j, scale) : T(0);
            v = cmul(cplx<T>{re, im}, w[j]);
        }
        Z[p * L + j] = v;
    }
}

// H (two real columns per complex sequence) from X (n per pair): DHT extraction
template <class T>
__global__ void k_dht_extract(long n, long ncols, long pair0, long npc, const cplx<T> *__restrict__ X,
                              T *__restrict__ H, long ldh) {
    const long p = blockIdx.y;
    const long c0 = 2 * (pair0 + p), c1 = c0 + 1;
    const bool has1 = c1 < ncols;
    const cplx<T> *x = X + p * n;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x) {
        long nk = k == 0 ? 0 : n - k;
        cplx<T> a = x[k], c = x[nk];
        H[k + c0 * ldh] = T(0.5) * ((a.x + c.x) - (a.y - c.y));
        if (has1) H[k + c1 * ldh] = T(0.5) * ((a.y + c.y) + (a.x - c.x));
    }
}

template <class T> static const BsPlan<T> &get_bs_plan(long n) {
    static std::map<long, BsPlan<T>> cache;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(n);
    if (it != cache.end()) return it->second;
    BsPlan<T> bp;
    bp.n = n;
    long L = 2 * n - 1;
    while (true) {
        if (smooth5(L) && make_cplan<T>(L, bp.cp)) break;
        ++L;
        SKL_REQUIRE(L < 16 * n + 64, SKL_E_UNSUPPORTED, "no Bluestein length found");
    }
    bp.L = L;
    std::vector<cplx<T>> w((size_t)n), b((size_t)L, cplx<T>{T(0), T(0)});
    const long double PI = 3.14159265358979323846264338327950288L;
    for (long j = 0; j < n; ++j) {
        long e = (long)(((__int128)j * j) % (2 * n));
        long double ang = -PI * (long double)e / (long double)n;
        w[(size_t)j] = {(T)cosl(ang), (T)sinl(ang)};
        cplx<T> cj = {(T)cosl(ang), (T)-sinl(ang)};
        b[(size_t)j] = cj;
        if (j > 0) b[(size_t)(L - j)] = cj;
    }
    {
        AllocLock lk(capture_mutex());
        SKL_HIP(hipMalloc(&bp.w, sizeof(cplx<T>) * (size_t)n));
        SKL_HIP(hipMalloc(&bp.B, sizeof(cplx<T>) * (size_t)L));
        SKL_HIP(hipMemcpy(bp.w, w.data(), sizeof(cplx<T>) * (size_t)n, hipMemcpyHostToDevice));
        SKL_HIP(hipMemcpy(bp.B, b.data(), sizeof(cplx<T>) * (size_t)L, hipMemcpyHostToDevice));
    }
    {
        // B = FFT_L(b) once per length, in place
        Arena tmp;
        CfftOps<T> ops;
        cfft<T>(bp.cp, 1, bp.B, L, ops, bp.B, L, L, tmp);
        SKL_HIP(hipStreamSynchronize(ctx().stream));
    }
    return cache.emplace(n, bp).first->second;
}

// DHT of scaled column pairs through Bluestein, into H (n x ncols, T)
template <class T>
static void dht_bluestein(long n, long ncols, const double *X, long ldx, const double *D,
                          double scale, T *H, long ldh, Arena &ar) {
    const BsPlan<T> &bp = get_bs_plan<T>(n);
    const long L = bp.L;
    long npairs = (ncols + 1) / 2;
    // pairs per chunk: about 512 MB of complex scratch
    long per = (long)std::max<size_t>(1, ((size_t)512 << 20) / (sizeof(cplx<T>) * (size_t)(3 * L + n)));
    hipStream_t s = ctx().stream;
    for (long p0 = 0; p0 < npairs; p0 += per) {
        long npc = std::min(per, npairs - p0);
        size_t mark = ar.used();
        cplx<T> *Z = ar.alloc_n<cplx<T>>((size_t)npc * L);
        cplx<T> *Xc = ar.alloc_n<cplx<T>>((size_t)npc * n);
        dim3 gp((unsigned)std::min<long>((L + 255) / 256, 1024), (unsigned)npc);
        k_bs_pack<T><<<gp, 256, 0, s>>>(n, L, ncols, X, ldx, D, scale, p0, bp.w, Z);
        SKL_CHECK_LAUNCH();
        CfftOps<T> o1;
        cfft<T>(bp.cp, npc, Z, L, o1, Z, L, L, ar);
        CfftOps<T> o2;
        o2.pre_tab = bp.B;
        o2.pre_conj = 1;
        o2.post_conj = 1;
        o2.post_scale = T(1) / (T)L;
        o2.post_tab = bp.w;
        cfft<T>(bp.cp, npc, Z, L, o2, Xc, n, n, ar);
        dim3 ge((unsigned)std::min<long>((n + 255) / 256, 1024), (unsigned)npc);
        k_dht_extract<T><<<ge, 256, 0, s>>>(n, ncols, p0, npc, Xc, H, ldh);
        SKL_CHECK_LAUNCH();
        SKL_HIP(hipStreamSynchronize(s));
        ar.rewind(mark);
    }
}

// Transform plan of length n, cached per length with its tables in device
// memory of its own.  In no-allocation mode a length not cached yet gets an
// uncached plan whose tables live in the arena (lengths with a prime factor
// above 5, whose Bluestein tables are larger, are refused there).
template <class T> static const FftPlanDev<T> &get_plan(long n, Arena &ar) {
    static std::map<long, FftPlanDev<T>> cache;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(n);
    if (it != cache.end()) return it->second;
    const bool scratch = ctx().no_alloc;
    if (scratch && !smooth(n))
        throw Error(SKL_E_UNSUPPORTED,
                    "no-allocation entry point: transform length " + std::to_string(n) +
                        " has a prime factor above 5");
    auto tables = [&](long L) { return scratch ? arena_twiddles<T>(L, ar) : make_twiddles<T>(L); };
    FftPlanDev<T> p;
    p.n = n;
    const long cap = lds_elems<T>();
    if (!smooth(n)) {
        p.mode = 3; // Bluestein
    } else if (n <= cap) {
        p.mode = 1;
        factor_dims(n, p.d1);
        p.tw1 = tables(n);
    } else {
        // n = n1 n2, both smooth, n1 <= cap/4 (J2 >= 4), n2 <= cap/8 (K1 >= 4)
        long best1 = -1;
        double bestscore = 1e300;
        for (long a = 2; a * a <= n * 64 && a <= n; ++a) {
            if (n % a) continue;
            long b = n / a;
            if (!smooth(a) || !smooth(b)) continue;
            if (a > cap / 4 || b > cap / 8) continue;
            double score = std::fabs(std::log((double)a) - std::log((double)b));
            if (score < bestscore) {
                bestscore = score;
                best1 = a;
            }
        }
        if (best1 < 0)
            throw Error(SKL_E_UNSUPPORTED, "DHT length " + std::to_string(n) + " too large");
        p.mode = 2;
        p.n1 = (int)best1;
        p.n2 = (int)(n / best1);
        factor_dims(p.n1, p.d1);
        factor_dims(p.n2, p.d2);
        p.tw1 = tables(p.n1);
        p.tw2 = tables(p.n2);
        p.twn = tables(n);
        p.J2 = (int)std::min<long>(32, cap / p.n1);
        // pass 2 takes G pairs at once (K1 rows of each) and writes H in
        // groups of G pairs: a row of a group is 16 G contiguous bytes, so
        // the bucket gather reads whole lines (n2 <= cap / 8, so G >= 4;
        // the workgroup's 2 K1 G sequences fill the LDS tile best with
        // K1 = 1 and G as large as the gather's registers allow)
        p.G = (int)std::min<long>(GATHER_GMAX, cap / (2L * p.n2));
        p.K1 = (int)std::max<long>(1, std::min<long>(16, cap / (2L * p.n2 * p.G)));
    }
    if (scratch) {
        thread_local FftPlanDev<T> tmp;
        tmp = p;
        return tmp;
    }
    auto res = cache.emplace(n, p);
    return res.first->second;
}

// dynamic LDS above 64 KB needs the kernel's attribute raised first
static void lds_allow(const void *fn, size_t bytes) {
    if (bytes > 65536) SKL_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// four-step into H (n x ncols, T)
template <class T>
static void dht_fourstep(const FftPlanDev<T> &P, long ncols, const double *X, long ldx,
                         const double *D, double scale, T *H, long ldh, Arena &ar,
                         int pairs_out = 0) {
    constexpr int NT = fft_nt<T>();
    long npairs = (ncols + 1) / 2;
    cplx<T> *Y = ar.alloc_n<cplx<T>>((size_t)npairs * P.n);
    hipStream_t s = ctx().stream;
    {
        dim3 g((unsigned)((P.n2 + P.J2 - 1) / P.J2), (unsigned)npairs);
        const size_t nhi = (size_t)((P.n + TW4_LO - 1) / TW4_LO);
        size_t sh = sizeof(cplx<T>) * ((size_t)P.n1 * P.J2 + P.n1 + TW4_LO + nhi);
        lds_allow((const void *)k_fft4_pass1<T, NT>, sh);
        k_fft4_pass1<T, NT><<<g, NT, sh, s>>>(P.n, P.n1, P.n2, P.J2, ncols, X, ldx, D, scale, P.d1,
                                              P.tw1, P.twn, Y);
        SKL_CHECK_LAUNCH();
    }
    {
        int half = P.n1 / 2;
        const long ngrp = (npairs + P.G - 1) / P.G;
        dim3 g((unsigned)((half + 1 + P.K1 - 1) / P.K1), (unsigned)ngrp);
        size_t sh = sizeof(cplx<T>) * ((size_t)P.n2 * 2 * P.K1 * P.G + P.n2);
        lds_allow((const void *)k_fft4_pass2<T, NT>, sh);
        k_fft4_pass2<T, NT><<<g, NT, sh, s>>>(P.n, P.n1, P.n2, P.K1, P.G, ncols, P.d2, P.tw2, Y, H,
                                              ldh, pairs_out);
        SKL_CHECK_LAUNCH();
    }
}

template <class T>
void dht_sketch(long n, long ncols, const double *X, long ldx, const double *D, double scale,
                long m, const int32_t *bptr, const int32_t *bsrc, const T *bcoef, T *SA,
                long ldsa, Arena &ar) {
    const FftPlanDev<T> &P = get_plan<T>(n, ar);
    constexpr int NT = fft_nt<T>();
    hipStream_t s = ctx().stream;
    long npairs = (ncols + 1) / 2;
    if (P.mode == 1) {
        size_t sh = sizeof(cplx<T>) * (size_t)n;
        k_dht_single<T, NT><<<(unsigned)npairs, NT, sh, s>>>(n, ncols, X, ldx, D, scale, P.d1,
                                                             P.tw1, 0, m, bptr, bsrc, bcoef, SA,
                                                             ldsa);
        SKL_CHECK_LAUNCH();
    } else {
        // the transformed columns pass through H (n x cols, in groups of G
        // pairs) and the four-step pass's complex pairs; in no-allocation
        // mode the columns go in chunks of an even count, a multiple of 2 G
        // where that fits (the pairs stay as in one pass, so the result is
        // the same), sized to a quarter of the scratch left
        const long cg = 2L * P.G; // columns per group (mode 2)
        long chunk = ncols;
        if (ctx().no_alloc) {
            const size_t per_col = 2 * sizeof(T) * (size_t)n + 64;
            const size_t room = (ar.capacity() - std::min(ar.used(), ar.capacity())) / 4;
            const long fit = (long)(room / per_col);
            chunk = fit >= cg ? fit / cg * cg : std::max<long>(2, fit & ~1L);
        }
        for (long c0 = 0; c0 < ncols; c0 += chunk) {
            const long nc = std::min(chunk, ncols - c0);
            const size_t mark = ar.used();
            const long ngrp = ((nc + 1) / 2 + P.G - 1) / P.G;
            T *H = ar.alloc_n<T>(std::max((size_t)n * (nc + 1), (size_t)n * ngrp * cg));
            if (P.mode == 2) {
                // groups of G pairs interleaved, gathered 2 G columns per line
                dht_fourstep<T>(P, nc, X + c0 * ldx, ldx, D, scale, H, n, ar, 1);
                const long nbh = (m + GATHER_NTH - 1) / GATHER_NTH;
                SKL_REQUIRE(nbh * ngrp <= 2147483647L, SKL_E_UNSUPPORTED,
                            "bucket gather: grid too large");
                const unsigned gg = (unsigned)(nbh * ngrp);
                const cplx<T> *H2 = reinterpret_cast<const cplx<T> *>(H);
                auto gather = [&](auto Gc) {
                    constexpr int G = decltype(Gc)::value;
                    k_hash_gather_groups<T, G><<<gg, GATHER_NTH, 0, s>>>(
                        m, n, nc, nbh, bptr, bsrc, bcoef, H2, SA + c0 * ldsa, ldsa);
                };
                static_assert(GATHER_GMAX == 10, "the cases below");
                switch (P.G) {
                case 4: gather(std::integral_constant<int, 4>{}); break;
                case 5: gather(std::integral_constant<int, 5>{}); break;
                case 6: gather(std::integral_constant<int, 6>{}); break;
                case 7: gather(std::integral_constant<int, 7>{}); break;
                case 8: gather(std::integral_constant<int, 8>{}); break;
                case 9: gather(std::integral_constant<int, 9>{}); break;
                case 10: gather(std::integral_constant<int, 10>{}); break;
                default: SKL_REQUIRE(false, SKL_E_UNSUPPORTED, "bucket gather: group size");
                }
                SKL_CHECK_LAUNCH();
            } else {
                dht_bluestein<T>(n, nc, X + c0 * ldx, ldx, D, scale, H, n, ar);
                hash_gather<T>(m, nc, bptr, bsrc, bcoef, H, n, SA + c0 * ldsa, ldsa);
            }
            if (chunk < ncols) ar.rewind(mark);
        }
    }
}
template void dht_sketch<double>(long, long, const double *, long, const double *, double, long,
                                 const int32_t *, const int32_t *, const double *, double *, long,
                                 Arena &);
template void dht_sketch<float>(long, long, const double *, long, const double *, double, long,
                                const int32_t *, const int32_t *, const float *, float *, long,
                                Arena &);

// plain DHT of columns: X given in double, output in T
template <class T>
void dht_cols_from_double(long n, long ncols, const double *X, long ldx, T *Y, long ldy,
                          Arena &ar) {
    const FftPlanDev<T> &P = get_plan<T>(n, ar);
    constexpr int NT = fft_nt<T>();
    long npairs = (ncols + 1) / 2;
    if (P.mode == 1) {
        size_t sh = sizeof(cplx<T>) * (size_t)n;
        k_dht_single<T, NT><<<(unsigned)npairs, NT, sh, ctx().stream>>>(
            n, ncols, X, ldx, nullptr, 1.0, P.d1, P.tw1, 1, 0, nullptr, nullptr, nullptr, Y, ldy);
        SKL_CHECK_LAUNCH();
    } else if (P.mode == 2) {
        dht_fourstep<T>(P, ncols, X, ldx, nullptr, 1.0, Y, ldy, ar);
    } else {
        dht_bluestein<T>(n, ncols, X, ldx, nullptr, 1.0, Y, ldy, ar);
    }
}
template void dht_cols_from_double<double>(long, long, const double *, long, double *, long,
                                           Arena &);
template void dht_cols_from_double<float>(long, long, const double *, long, float *, long,
                                          Arena &);

} // namespace skl
