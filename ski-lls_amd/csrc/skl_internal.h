// skl_internal.h -- shared declarations of the MI355X least-squares library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/skl.h"

namespace skl {

// ---------------------------------------------------------------------------
// Errors: internal code throws skl::Error; the C ABI catches and maps it.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throw_hip(hipError_t e, const char *what, const char *file, int line);

#define SKL_HIP(call)                                                   \
    do {                                                                \
        hipError_t e_ = (call);                                         \
        if (e_ != hipSuccess) ::skl::throw_hip(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define SKL_CHECK_LAUNCH() SKL_HIP(hipGetLastError())

#define SKL_REQUIRE(cond, code, msg)                 \
    do {                                             \
        if (!(cond)) throw ::skl::Error((code), (msg)); \
    } while (0)

// ---------------------------------------------------------------------------
// Device allocations, frees, stream/event creation and graph instantiation
// of all worker threads take this lock, one at a time (hipFree waits for
// every stream of the device).  No stream is ever captured: the LSQR graphs
// are built node by node (k_lsqr.hip), because another thread's runtime calls
// invalidated stream captures mid-recording and left the stream unusable.
// ---------------------------------------------------------------------------
std::shared_mutex &capture_mutex();
using AllocLock = std::unique_lock<std::shared_mutex>;

// ---------------------------------------------------------------------------
// Device memory: a grow-only arena per context; buffers are carved with a
// bump pointer and released all at once by reset().  A solve that outgrows the
// reservation gets extra blocks; the next reset() replaces base and extras by
// one block of the high-water size, so repeated solves of one shape allocate
// nothing after the first.
// ---------------------------------------------------------------------------
class Arena {
  public:
    ~Arena();
    void reserve(size_t bytes);        // ensure capacity (reallocates when empty)
    void *alloc(size_t bytes);         // 256-byte aligned
    template <class T> T *alloc_n(size_t n) { return static_cast<T *>(alloc(n * sizeof(T))); }
    void reset();                      // all buffers dead; caller's stream work is done
    // release everything carved after `mark` (stream-ordered reuse: later
    // kernels on the same stream may take the memory over)
    void rewind(size_t mark) {
        if (mark <= used_) used_ = mark;
    }
    size_t used() const { return used_; }
    size_t capacity() const { return cap_; }
    void release();
    // Fixed mode: the arena carves from [base, base + bytes), memory it does
    // not own; it never allocates, and a request beyond the end throws
    // Error(SKL_E_UNSUPPORTED).
    void attach_fixed(void *base, size_t bytes);
    bool fixed() const { return fixed_; }

  private:
    bool fixed_ = false;
    char *base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    size_t over_ = 0, peak_ = 0;   // bytes in extra blocks; high-water of used_ + over_
    std::vector<void *> overflow_; // extra blocks when the reservation was too small
};

// Arguments of the bounded grid barrier (k_common.h:grid_barrier)
struct GridSync {
    unsigned *ctr;  // arrival counter, zero at launch
    unsigned *err;  // nonzero once a wait gave up (sticky until the host clears it)
    unsigned spins; // polls before giving up
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    Arena arena;
    // error word of the grid barriers launched from this context, and the
    // number of polls after which a barrier wait gives up
    unsigned *dev_err = nullptr;
    unsigned barrier_spins = 1u << 28;
    // no-allocation mode (skl_dense_solve_dev_noalloc): solver scratch comes
    // from `fixed`, attached to the device's static scratch array; no device
    // memory, stream, event or graph is created while it is set
    bool no_alloc = false;
    Arena fixed;
    Arena &cur_arena() { return no_alloc ? fixed : arena; }
    // device bytes the current solve was admitted with (a server slot; 0:
    // no allowance, decisions that depend on free memory ask the device)
    size_t mem_allowance = 0;
    // device bytes this context keeps between solves (arena reservation and
    // split-K scratch)
    size_t retained_bytes() const;
    int num_cu = 0;
    int coop_ok = 0;
    size_t lds_max = 0;
    void *pinned = nullptr; // small pinned host buffer for scalars
    // zeroed grid-barrier counters, one per launch (grid_sync_fresh)
    unsigned *ctr_pool = nullptr;
    int ctr_next = 0;
    // grow-only scratch for split-K partial products (stream-ordered reuse)
    void *ws = nullptr;
    size_t ws_cap = 0;
    void *workspace(size_t bytes);
    // second stream for work that overlaps the first (trailing updates of the
    // blocked QR), with its own split-K scratch and two events for the
    // hand-offs between the streams; made on first use (make_side)
    hipStream_t side = nullptr;
    void *ws_side = nullptr;
    size_t ws_side_cap = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    void make_side();
    // third stream for the row-major copy of A that the one-pass LSQR kernel
    // reads, made while the QR runs on the other two; its events mark the
    // copy's start (after A is on the device) and end
    hipStream_t aux = nullptr;
    hipEvent_t aux_ev[2] = {nullptr, nullptr};
    void make_aux();
};

// While alive, ctx().stream (and the split-K scratch) is the context's side
// stream; everything issued through ctx() goes there.
class OnSideStream {
  public:
    OnSideStream();
    ~OnSideStream();
    OnSideStream(const OnSideStream &) = delete;
    OnSideStream &operator=(const OnSideStream &) = delete;
};

Context &ctx();                       // default per-thread context
// GridSync for a launch from the calling thread's context, counter ctr
GridSync grid_sync(unsigned *ctr);
// A grid barrier whose counter is already zero: taken from the context's pool
// of zeroed counters (re-zeroed in one fill when used up), or, in
// no-allocation mode, carved from `ar` and zeroed on the stream.
GridSync grid_sync_fresh(Arena &ar);
// Waits for ctx().stream and throws Error(SKL_E_HIP) when a grid barrier
// launched from this context gave up since the last check (clearing the word).
void check_grid_fault();
// Resolves the device address of the static scratch array of the
// no-allocation entry point for the current device (once per device; called
// from the context set-up with that device current).
void noalloc_setup(int device);
// true when the calling thread holds a context (skl_init or an earlier call)
bool has_context();
// sets the message skl_last_error() returns on the calling thread
void set_last_error(const std::string &msg);
void ensure_init();
void release_thread_context();        // frees the calling thread's context
void barrier_tokens_stream_drained(hipStream_t s);

// Kernels with a grid barrier need every workgroup resident at once, and each
// of their workgroups fills a CU (registers or LDS), so the workgroups of all
// such kernels in flight on the device must fit on its CUs together.  A launch
// reserves one token per workgroup from a process-wide pool of num_cu tokens
// (an exclusive launch, for grids over half the device, reserves them all) and
// returns them when an event recorded after it completes (one event may serve
// several launches); no host thread waits for a launch unless the pool is
// short.
class BarrierLaunch {
  public:
    explicit BarrierLaunch(int grid, bool exclusive = false);
    ~BarrierLaunch(); // the completion event on ctx().stream, now or later
    BarrierLaunch(const BarrierLaunch &) = delete;
    BarrierLaunch &operator=(const BarrierLaunch &) = delete;

  private:
    int tokens_ = 0;
};

double wall();

// ---------------------------------------------------------------------------
// Hashing pattern (host): libc-compatible generator and bucket lists.
// ---------------------------------------------------------------------------
struct Glibc {
    int32_t s[31];
    int f, r;
    int toggle;
    double b2;
    void seed(unsigned sd);
    inline int next() {
        uint32_t v = (uint32_t)s[f] + (uint32_t)s[r];
        s[f] = (int32_t)v;
        if (++f >= 31) f = 0;
        if (++r >= 31) r = 0;
        return (int)(v >> 1);
    }
    double normal();
    // out[0..count) = the next count outputs of next(), generated in parallel
    // chunks for large counts (each chunk's starting window by jump-ahead);
    // the generator is left as after count calls of next()
    void fill(int *out, long count);
};

struct Pattern {
    long m = 0, n = 0, k = 0;
    std::vector<double> D;      // signs (dense path), n
    std::vector<int32_t> rows;  // n*k target rows
    std::vector<double> vals;   // n*k values +-1/sqrt(k)
    // bucket lists (transpose of the pattern): members of bucket h are
    // bsrc[bptr[h] .. bptr[h+1]) in increasing source index, bslot = k*i+t
    std::vector<int32_t> bptr, bsrc, bslot;
};

void gen_signs(Glibc &g, long n, std::vector<double> &D);
void gen_hashing(Glibc &g, long m, long n, long k, Pattern &p);
void build_buckets(Pattern &p);
void gen_subsample(Glibc &g, long n, long m, Pattern &p);
void gen_gaussian(unsigned seed, long rows, long cols, std::vector<double> &G);

// ---------------------------------------------------------------------------
// Kernel-side launchers (defined in the .hip files).  All run on ctx().stream.
// ---------------------------------------------------------------------------
// DHT sketch: SA[:, c] = alpha * S * H * (scale D .* X[:, c]) for c < ncols
template <class T>
void dht_sketch(long n, long ncols, const double *X, long ldx, const double *D, double scale,
                long m, const int32_t *bptr, const int32_t *bsrc, const T *bcoef, T *SA,
                long ldsa, Arena &ar);
// Plain DHT of columns (unnormalised) into Y (real, same precision as T)
template <class T>
void dht_cols(long n, long ncols, const T *X, long ldx, T *Y, long ldy, Arena &ar);
// bucket-gather hashing of dense columns (already transformed)
template <class T>
void hash_gather(long m, long ncols, const int32_t *bptr, const int32_t *bsrc,
                 const T *bcoef, const T *X, long ldx, T *Y, long ldy);
// sparse sketch: SA = S A (CSC, int32 indices on device)
void sparse_sketch_dev(long m, long n, long d, long k, const int32_t *rows, const double *vals,
                       const int64_t *colptr, const int32_t *rowind, const double *val,
                       double *SA, long ldsa, Arena &ar);

// QR family (column-major, W is m x ncol with ld; nfac leading columns factored)
template <class T>
void qr_householder(long m, long ncol, long nfac, T *W, long ldw, T *tau, Arena &ar);
template <class T>
void cpqr_greedy(long m, long ncol, long nfac, T *W, long ldw, int32_t *jpvt, T *tau,
                 Arena &ar);
template <class T>
void cpqr_hqrrp(long m, long ncol, long nfac, T *W, long ldw, int32_t *jpvt, T *tau,
                const double *G_host, Arena &ar);

// triangular inverse and norms
// G = R^T R (R upper, zeros below); false when G is not numerically
// positive definite.  G is overwritten.
bool cholesky_upper(long d, double *G, long ldg, double *R, long ldr, Arena &ar);
void trinv_upper(long r, const double *R, long ldr, double perturb, double *M, long ldm,
                 Arena &ar);
template <class T>
void upcast_upper(long r, const T *R, long ldr, double *Rd, long ldd);
// ||diag(rs) R diag(cs)||_inf of an upper triangular R (rs, cs may be null)
double norm_inf_upper(long r, const double *R, long ldr, Arena &ar, const double *rs = nullptr,
                      const double *cs = nullptr);
// column 2-norms of an upper triangular R and their reciprocals
void colnorm_upper(long r, const double *R, long ldr, double *D, double *invD);

// GEMM: C = alpha op(A) op(B) + beta C (column-major)
template <class T>
void gemm(char ta, char tb, long M, long N, long K, T alpha, const T *A, long lda, const T *B,
          long ldb, T beta, T *C, long ldc, long batch = 1, long sA = 0, long sB = 0,
          long sC = 0, bool upper = false);

// LSQR on the preconditioned operator A(:, E) M
struct LsqrProblem {
    long n = 0, d = 0, r = 0;
    // dense A (n x d, lda) or sparse CSR + CSC (int32 indices)
    const double *A = nullptr;
    long lda = 0;
    const int32_t *csr_ptr = nullptr, *csr_col = nullptr;
    const double *csr_val = nullptr;
    const int32_t *csc_ptr = nullptr, *csc_row = nullptr;
    const double *csc_val = nullptr;
    long nnz = 0;                // sparse: stored entries
    const int32_t *E = nullptr; // r entries or null (identity)
    const double *M = nullptr;  // r x r upper, ldm
    long ldm = 0;
    // dense: a row-major copy of A (row stride ldt) under way on the aux
    // stream, complete when AT_ready fires (lsqr_start_rows); when null,
    // lsqr_solve makes the copy itself if its one-pass kernel is used
    const double *AT = nullptr;
    long ldt = 0;
    hipEvent_t AT_ready = nullptr;
};
struct LsqrResult {
    int flag = 0;
    long it = 0;
    double nrm_ar = 0, nrm_a = 0, nrm_r = 0;
};
LsqrResult lsqr_solve(const LsqrProblem &p, const double *b, double tol, long max_it,
                      double *y, bool use_graph, int chunk, skl_lsqr_step *hist,
                      long hist_len, Arena &ar);
// Starts the row-major copy of dense A (p.n, p.d, p.A, p.lda) for the one-pass
// LSQR kernel on the aux stream, after the work queued so far on the main
// stream, and sets p.AT / p.ldt / p.AT_ready; false (p unchanged) when the
// one-pass kernel will not be used.  The copy lives in the arena.
bool lsqr_start_rows(LsqrProblem &p, Arena &ar);
// whether LSQR on a dense n x d A runs its one-pass kernel over a row-major
// copy (by shape; no device query); the copy's row stride; the copy itself (AT: n x lsqr_rows_ld(d)),
// queued on stream s
bool lsqr_rows_wanted(long n, long d);
// Launch shapes of the LSQR kernels per problem-shape bucket (half octaves of
// n, d and, sparse, nnz / n): lsqr_autotune times the candidates on a
// synthetic problem of the shape and records the fastest (ms per iteration
// returned; *default_ms the default shape's); later solves in the bucket use
// it.  lsqr_shape_bucket: the bucket's id.
double lsqr_autotune(bool sparse, long n, long d, long nnz, int iters, double *default_ms);
bool lsqr_shape_tuned(bool sparse, long n, long d, long nnz);
long lsqr_shape_entries();
void lsqr_shape_clear();
long long lsqr_shape_bucket(bool sparse, long n, long d, long nnz);
void csc_to_csr(long n, long d, long nnz, const int32_t *colptr, const int32_t *rowind,
                const double *val, int32_t *rptr, int32_t *ccol, double *cval, Arena &ar);
long lsqr_rows_ld(long d);
void lsqr_make_rows(long n, long d, const double *A, long lda, double *AT, hipStream_t s);

// small vector helpers
void dev_gemv_n_dense(long n, long d, const double *A, long lda, const double *x, double alpha,
                      double beta, double *y, Arena &ar);
void dev_csc_gemv_n(long n, long d, const int64_t *colptr, const int32_t *rowind,
                    const double *val, const double *x, double alpha, double beta,
                    double *y, Arena &ar);
void dev_gemv_t_dense(long n, long d, const double *A, long lda, const double *u, double alpha,
                      double beta, double *y, Arena &ar);
double dev_nrm2(long n, const double *x, Arena &ar);

} // namespace skl
