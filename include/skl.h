/*
 * skl.h -- C ABI of the MI355X sketch-and-precondition least-squares library.
 *
 * Plain pointers and sizes only; every matrix is column-major with an explicit
 * leading dimension; indices are `long` (64-bit), like the reference's ILP64
 * build (include/config_sk.h:5).  All entry points return an int status
 * (SKL_OK == 0); skl_last_error() gives the message of the last failure on
 * the calling thread.
 *
 * The three solver entry points replace the reference's public API
 * (include/ski-lls.h:15-28); the C++ drop-in header ski-lls.h forwards to
 * them.  The component entry points expose each hot-path stage (pattern,
 * transform, sketch, QR, triangular inverse, LSQR) so that each can be
 * checked on its own.
 */
#ifndef SKL_H
#define SKL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SKL_OK 0
#define SKL_E_SHAPE 1       /* "Matrix is not overdetermined." (n < d) */
#define SKL_E_RANK0 2       /* "SPQR Factorization fails" (rank 0) */
#define SKL_E_NOMEM 3       /* device allocation failed */
#define SKL_E_HIP 4         /* HIP runtime error */
#define SKL_E_ARG 5         /* invalid argument */
#define SKL_E_UNSUPPORTED 6 /* size outside the implemented range */
#define SKL_E_NODEVICE 7    /* no usable GPU */

/* precision modes */
#define SKL_FP64 0  /* sketch, transform, QR and LSQR all in FP64 */
#define SKL_MIXED 1 /* preconditioner from a CholeskyQR of the column-
                       equilibrated sketch (FP32 MFMA Gram, FP64 Cholesky);
                       LSQR in FP64.  Falls back to the FP64 factorization of
                       the same sketch when the Cholesky breaks down, the rank
                       falls short, a Lanczos estimate of (SA M)^T (SA M) (run
                       when an equilibrated pivot is below sqrt(max(m,d) u_32))
                       has an eigenvalue <= 0.1, or the inf-norm condition
                       number of the column-equilibrated R exceeds
                       escalate_cond.  The
                       solvers with a sketch-and-solve start sketch in FP64
                       and refine the start against that sketch; the _nocpqr
                       solver sketches in FP32. */
#define SKL_CHOLQR64 2 /* ablation of SKL_MIXED's precision: the same
                          CholeskyQR preconditioner, start refinement and
                          fallbacks, with the equilibrated sketch and its Gram
                          matrix in FP64 (FP64 MFMA) and an FP64 sketch on
                          every solver.  No counterpart in the reference. */

typedef struct skl_opts {
    double gamma;           /* over-sampling ratio, m = ceil(gamma d)          */
    long k;                 /* nonzeros per column of the hashing matrix       */
    double abs_tol;         /* early exit when ||b - A x0|| < abs_tol          */
    double rcond;           /* dense: rank = #leading |R_ii| > rcond           */
    double it_tol;          /* LSQR relative tolerance                         */
    long max_it;            /* LSQR iteration limit                            */
    double rcond_threshold; /* sparse: R11 "well conditioned" threshold        */
    double perturb;         /* sparse: pivot perturbation when ill conditioned */
    int ordering;           /* accepted for compatibility, ignored             */
    int debug;              /* accepted for compatibility, ignored             */
    int wisdom;             /* accepted for compatibility, ignored             */
    unsigned seed;          /* seed of the libc-compatible generator           */
    int precision;          /* SKL_FP64, SKL_MIXED or SKL_CHOLQR64             */
    double escalate_cond;   /* MIXED: condition estimate that forces FP64 QR   */
    int use_graph;          /* capture the LSQR loop in a hipGraph (1)         */
    int graph_chunk;        /* LSQR iterations per graph launch                */
    int reserved[6];
} skl_opts;

typedef struct skl_info {
    long rank;
    int flag;           /* 0 converged / early exit, 1 max_it reached   */
    long it;            /* LSQR iterations                              */
    int success;        /* sparse: rcond(R11) > rcond_threshold         */
    int escalated;      /* MIXED: FP64 QR fallback taken                */
    double rcond;       /* reciprocal inf-norm condition number of R11  */
    double cond_est;    /* MIXED: ||R D^-1||_inf ||D R^-1||_inf of the FP32
                           factor, D = its column norms (escalation test) */
    double x0_residual; /* ||b - A x0|| of the sketch-and-solve start   */
    double lsqr_nrm_ar, lsqr_nrm_a, lsqr_nrm_r; /* LSQR final estimates */
    double t_pattern, t_sketch, t_qr, t_x0, t_lsqr, t_total; /* seconds */
    long m;             /* sketch rows                                  */
    double precond_lmin, precond_lmax; /* MIXED: Lanczos estimate of the extreme
                           eigenvalues of (SA M)^T (SA M), when the FP32 factor's
                           pivots called for it (0 when not estimated)  */
} skl_info;

/* LSQR per-iteration scalars, for step-by-step comparison */
typedef struct skl_lsqr_step {
    double alpha, beta, rhobar, phibar, nrm_a;
} skl_lsqr_step;

void skl_default_opts(skl_opts *o);
int skl_init(int device);
/* releases the calling thread's stream, device buffers and cached graphs; the
   next call on this thread sets them up again */
void skl_finalize(void);
const char *skl_error_string(int code);
const char *skl_last_error(void);
int skl_device_count(void);

/* ---- solvers: host pointers ------------------------------------------------
 * ls_dense_hashing_blendenpik (src/ls_dense_hashing_blendenpik.cpp:23-149).
 * A (n x d, lda) is read-only here: the column permutation the reference
 * applies in place (dlapmt, :123,147) is folded into the kernels' indexing. */
int skl_dense_solve(long n, long d, const double *A, long lda, const double *b, double *x,
                    const skl_opts *opts, skl_info *info);
/* ls_dense_hashing_blendenpik_noCPQR (src/ls_dense_hashing_blendenpik_noCPQR.cpp:23-60) */
int skl_dense_solve_nocpqr(long n, long d, const double *A, long lda, const double *b,
                           double *x, const skl_opts *opts, skl_info *info);
/* ls_dense_blendenpik (src/ls_dense_blendenpik.cpp:21-71, declared in
 * include/blendenpik.hpp:31-33): the randomized DHT, uniform sampling of
 * m = ceil(gamma d) of its rows (src/dense_helpers.cpp:19-59), Householder QR
 * and LSQR from zero; opts.k, abs_tol and rcond are not used. */
int skl_dense_solve_blendenpik(long n, long d, const double *A, long lda, const double *b,
                               double *x, const skl_opts *opts, skl_info *info);
/* ls_sparse_spqr (src/ls_sparse_spqr.cpp:79-214); A in CSC (cs_dl layout:
 * colptr d+1, rowind/val nnz) */
int skl_sparse_solve(long n, long d, long nnz, const long *colptr, const long *rowind,
                     const double *val, const double *b, double *x, const skl_opts *opts,
                     skl_info *info);

/* ---- solvers: device-resident inputs (pointers into the current device) ---- */
int skl_dense_solve_dev(long n, long d, const double *dA, long lda, const double *db,
                        double *dx, const skl_opts *opts, skl_info *info);
int skl_dense_solve_nocpqr_dev(long n, long d, const double *dA, long lda, const double *db,
                               double *dx, const skl_opts *opts, skl_info *info);
int skl_dense_solve_blendenpik_dev(long n, long d, const double *dA, long lda, const double *db,
                                   double *dx, const skl_opts *opts, skl_info *info);
int skl_sparse_solve_dev(long n, long d, long nnz, const long *dcolptr, const long *drowind,
                         const double *dval, const double *db, double *dx,
                         const skl_opts *opts, skl_info *info);

/* ls_dense_hashing_blendenpik without device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory or create
 * streams (code being captured into a HIP graph, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * skl_dense_solve_dev, computed by the same kernels, with every buffer carved
 * from a static 64 MiB device array of the library (one per GPU, held until
 * the process ends) instead of the per-thread arena; the LSQR operator runs
 * as its two-pass kernels, without graph replay, and the QR without its
 * second stream.  skl_init must have been called on the calling thread.
 * Limits: opts->precision SKL_FP64 only (SKL_E_UNSUPPORTED otherwise);
 * problems whose scratch exceeds 64 MiB (roughly 8 (m (d + 1) + 4 d^2 + 8 n)
 * bytes, m = ceil(gamma d)) are refused with SKL_E_UNSUPPORTED before any
 * work, as are transform lengths n with a prime factor above 5 that no
 * earlier call of the library has planned.  Calls share the array: calls
 * from several threads of one process take turns (they would race on it
 * otherwise), and each call leaves its scratch contents behind. */
int skl_dense_solve_dev_noalloc(long n, long d, const double *dA, long lda, const double *db,
                                double *dx, const skl_opts *opts, skl_info *info);

/* Kernel timing of the LSQR operator products on this thread: when enabled,
 * the LSQR loop runs without graph replay and HIP events on the library
 * stream bracket the four products of the second iteration of every chunk
 * (chunks that end the solve are left out).  slot 0: A z, 1: A^T u, 2: M v,
 * 3: M^T t.  bytes_per_launch is the algorithmic traffic of one launch. */
void skl_profile_enable(int on);
void skl_profile_reset(void);
int skl_profile_read(int slot, double *avg_ms, long *launches, double *bytes_per_launch);

/* Poll limit of the factorization kernels' grid barriers on this thread (no
 * counterpart in the reference).  A barrier whose wait exceeds it gives up,
 * the kernels of that factorization stop early, and the call that launched
 * them returns SKL_E_HIP instead of a result computed from stale data.
 * spins < 0 restores the default (2^28 polls, seconds); 0 makes a barrier give
 * up unless every workgroup has already arrived, which tests that error path. */
void skl_set_barrier_spins(long spins);

/* LSQR per-iteration history of the next solve on this thread (optional):
 * steps must hold max_steps entries; pass NULL to switch it off. */
void skl_set_history(skl_lsqr_step *steps, long max_steps);

/* ---- HBM-resident problems (handles) ----------------------------------------
 * No counterpart in the reference, whose calls take A each time; these are the
 * "keep A resident" entry points for callers that solve on one matrix more
 * than once or feed the serving loop.  An upload copies A to the calling
 * thread's device once, together with what every solve would rebuild (dense:
 * the row-major copy of A for the one-pass LSQR kernel when LSQR would use it
 * and it fits; CSC: int32 indices and a CSR copy).  Each solve on a handle
 * computes exactly what the corresponding solver computes on the same A, b
 * and opts (same sketch for the same seed, same results).  Solves only read
 * the handle: several may run on it at once, from any threads of the
 * process that use its device.  skl_release frees it; no solve may still be
 * running on it. */
typedef struct skl_handle skl_handle;
#define SKL_METHOD_DENSE_CPQR 0       /* ls_dense_hashing_blendenpik */
#define SKL_METHOD_DENSE_NOCPQR 1     /* ls_dense_hashing_blendenpik_noCPQR */
#define SKL_METHOD_DENSE_BLENDENPIK 2 /* ls_dense_blendenpik */
#define SKL_METHOD_SPARSE_SPQR 3      /* ls_sparse_spqr (CSC handles only) */
/* dense A (host or device pointer, n x d, lda >= n) */
int skl_upload_dense(long n, long d, const double *A, long lda, skl_handle **h);
int skl_upload_dense_dev(long n, long d, const double *dA, long lda, skl_handle **h);
/* CSC A (host arrays in the cs_dl layout, checked: colptr monotone from 0 to
 * nnz, row indices in [0, n)) */
int skl_upload_csc(long n, long d, long nnz, const long *colptr, const long *rowind,
                   const double *val, skl_handle **h);
/* shape, stored entries (-1 for dense) and device bytes held */
int skl_handle_query(const skl_handle *h, long *n, long *d, long *nnz, size_t *device_bytes);
/* b (n) and x (d) in host memory / device memory of the handle's device */
int skl_solve_handle(skl_handle *h, int method, const double *b, double *x, const skl_opts *opts,
                     skl_info *info);
int skl_solve_handle_dev(skl_handle *h, int method, const double *db, double *dx,
                         const skl_opts *opts, skl_info *info);
int skl_release(skl_handle *h);

/* ---- LSQR launch shapes: autotuner --------------------------------------------
 * No counterpart in the reference.  The LSQR kernels' launch shapes (the
 * workgroup count the two-pass products aim at, workgroups per CU of the
 * one-pass dense kernel, z staged in LDS for the sparse product, and for
 * rank <= 256 with at most 128 rows per workgroup on half the CUs whether the
 * iterations run as one persistent launch and on how many workgroups) come from a
 * process-wide table keyed by problem-shape bucket (half octaves of n, d and,
 * for CSC, nnz/n); buckets not in it use the defaults.  skl_autotune_lsqr
 * times every candidate shape on a synthetic problem of the given shape (32
 * LSQR iterations replayed as graphs, on the calling thread's device) and
 * records the fastest; best_ms / default_ms: ms per iteration of the chosen
 * and the default shape.  Shapes change the order of partial sums, so x can
 * differ in the last bits (within the LSQR tolerance) between a tuned and an
 * untuned bucket; with a fixed table solves stay bitwise reproducible. */
int skl_autotune_lsqr(int sparse, long n, long d, long nnz, double *best_ms, double *default_ms);
long skl_autotune_entries(void);
void skl_autotune_clear(void);

/* ---- serving: continuous batching over worker slots -------------------------
 * No counterpart in the reference (it solves one problem per call); this is
 * the request loop that callers with a stream of problems use.  Each slot is
 * a worker with its own HIP stream and device arena; queued requests are
 * admitted in arrival order when a slot is free and their device-memory
 * footprint fits mem_budget_bytes (0: 85% of free device memory).  Input
 * arrays must stay valid until skl_server_wait returns for the ticket. */
typedef struct skl_server skl_server;
skl_server *skl_server_create(int slots, size_t mem_budget_bytes, int device);
long skl_server_submit_dense(skl_server *s, int cpqr, long n, long d, const double *A, long lda,
                             const double *b, double *x, const skl_opts *opts);
long skl_server_submit_sparse(skl_server *s, long n, long d, long nnz, const long *colptr,
                              const long *rowind, const double *val, const double *b, double *x,
                              const skl_opts *opts);
/* device-pointer requests (the server's device); the footprint leaves out
 * the inputs, which are already resident */
long skl_server_submit_dense_dev(skl_server *s, int cpqr, long n, long d, const double *dA,
                                 long lda, const double *db, double *dx, const skl_opts *opts);
long skl_server_submit_sparse_dev(skl_server *s, long n, long d, long nnz, const long *dcolptr,
                                  const long *drowind, const double *dval, const double *db,
                                  double *dx, const skl_opts *opts);
/* a solve on a handle (b, x host pointers); the handle must stay alive until
 * skl_server_wait returns for the ticket */
long skl_server_submit_handle(skl_server *s, skl_handle *h, int method, const double *b,
                              double *x, const skl_opts *opts);
/* blocks until the request finished; queue_s / run_s: time queued / solving */
int skl_server_wait(skl_server *s, long ticket, skl_info *info, double *queue_s, double *run_s);
int skl_server_pending(skl_server *s);
/* on != 0: a slot that takes a request whose LSQR shape bucket the table lacks
 * runs skl_autotune_lsqr for it first (once per bucket; requests of a bucket
 * being tuned wait for it), so every request of a bucket uses one shape */
void skl_server_set_autotune(skl_server *s, int on);
void skl_server_destroy(skl_server *s);

/* ---- component entry points (host pointers; each runs one stage) -----------
 * Hashing pattern in the reference's draw order: when dense != 0 the first n
 * draws give the signs D (src/dense_helpers.cpp:7-16, D may be NULL to skip
 * them), then n*k row draws and n*k sign draws (src/sparse_helpers.cpp:90-116). */
int skl_hash_pattern(unsigned seed, int dense, long m, long n, long k, double *D, long *rows,
                     double *vals);
/* Blendenpik's sampling: the n signs D (src/dense_helpers.cpp:7-16), then,
 * from the same seed again (src/dense_helpers.cpp:25), the m sampled rows idx
 * of a partial Fisher-Yates shuffle of 0..n-1 (src/sparse_helpers.cpp:74-87) */
int skl_subsample_pattern(unsigned seed, long m, long n, double *D, long *idx);
/* Y = DHT(X) per column, unnormalised FFTW_DHT (src/fftw_r2r.cpp:17-54);
 * precision SKL_FP64 or SKL_MIXED (FP32 arithmetic, FP64 in/out) */
int skl_dht(long n, long ncols, const double *X, long ldx, double *Y, long ldy, int precision);
/* Y (m x ncols) = alpha S X with S the hashing matrix (rows/vals, k per
 * column), summed per bucket in increasing source index (SpMat.cpp:53-55) */
int skl_hash_apply(long m, long n, long k, const long *rows, const double *vals, double alpha,
                   long ncols, const double *X, long ldx, double *Y, long ldy, int precision);
/* SA (m x (d + has_b)) = (1/sqrt k) S H D [A, b]; the randomized transform and
 * hashing of rand_hashing_dct (src/dense_helpers.cpp:83-112) */
int skl_dense_sketch(long n, long d, const double *A, long lda, const double *b, long m,
                     long k, const double *D, const long *rows, const double *vals,
                     double *SA, long ldsa, int precision);
/* SA (m x d, dense) = S A_csc, and bs = S b when b != NULL; CXSparse
 * cs_multiply / cs_gaxpy summation order (src/sketching.cpp:90-111) */
int skl_sparse_sketch(long m, long n, long d, long k, const long *rows, const double *vals,
                      const long *colptr, const long *rowind, const double *val,
                      const double *b, double *SA, long ldsa, double *bs);
/* Householder QR (dgeqrf semantics, src/dense_qr.cpp:8-33) of A (m x n) in
 * place; the first nfac columns are factored and the rest receive Q^T. */
int skl_qr(long m, long n, long nfac, double *A, long lda, double *tau, int precision);
/* column-pivoted QR of the first nfac columns (dgeqp4_INT dispatch,
 * hqrrp/.../NoFLA_HQRRP_WY_blk_var4_skint.c:131-310: greedy dgeqp3 pivoting
 * when min(m,nfac) < 1000, else HQRRP); remaining columns receive Q^T.
 * jpvt is 1-based on output. */
int skl_cpqr(long m, long n, long nfac, double *A, long lda, long *jpvt, double *tau,
             int precision);
/* Minv = (R + perturb I)^{-1} for upper triangular R (r x r, ldr) */
int skl_trinv(long r, const double *R, long ldr, double perturb, double *Minv, long ldm);
/* LSQR on op = A(:, E[0:r]) M (M = R^{-1} upper, r x r), from y = 0
 * (src/IterSolver.cpp:160-284); E may be NULL (identity). */
int skl_lsqr_dense(long n, long d, const double *A, long lda, long r, const long *E,
                   const double *M, long ldm, const double *b, double tol, long max_it,
                   double *y, int *flag, long *it, skl_lsqr_step *hist);

#ifdef __cplusplus
}
#endif
#endif /* SKL_H */
