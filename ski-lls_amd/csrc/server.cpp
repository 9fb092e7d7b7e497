// server.cpp -- continuous-batching request loop over the solvers.
//
// A server owns `slots` worker threads; each worker has its own HIP stream
// and device arena (the per-thread context), so solves admitted together
// overlap on the device.  Requests wait in arrival order and are admitted
// when a slot is free and their device-memory footprint fits the budget;
// completion frees both.  Every solve reduces in a fixed order on its own
// buffers, so a result does not depend on what else is running.
//
// Memory accounting: a slot holds device memory between requests (its arena
// keeps the high-water size of the solves it ran, and its split-K scratch),
// so each slot is charged max(what it retains, the footprint of the request
// it runs).  A request is admitted by a free slot when the charges of the
// other slots plus its own charge fit the budget (a request that fits no
// budget runs when every other slot is idle).  The admitted footprint is
// also the solve's memory allowance: choices that depend on free memory
// (the row-major copy for one-pass LSQR) follow it instead of asking the
// device, whose free memory the other slots change meanwhile.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <string>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "skl_internal.h"

namespace skl {
void release_thread_context();
}

namespace {

using clk = std::chrono::steady_clock;

enum Kind { DENSE_CPQR = 0, DENSE_NOCPQR = 1, SPARSE = 2, HANDLE = 3 };

struct Request {
    long ticket = 0;
    int kind = 0;
    bool dev = false; // device pointers
    long n = 0, d = 0, nnz = 0;
    const double *A = nullptr;
    long lda = 0;
    const long *colptr = nullptr, *rowind = nullptr;
    const double *val = nullptr;
    skl_handle *h = nullptr;
    int method = 0;
    const double *b = nullptr;
    double *x = nullptr;
    skl_opts opts{};
    size_t footprint = 0;
    clk::time_point submitted, started, finished;
    int rc = -1;
    std::string err; // skl_last_error() of the worker when rc != SKL_OK
    skl_info info{};
    bool done = false;
};

// Device bytes one solve needs beyond what is already resident: the inputs
// (host requests), per-solve copies of A (int32 CSC + CSR for sparse A, the
// row-major copy for one-pass dense LSQR), the sketch and its QR, R and its
// inverse, the LSQR vectors and partial sums, and the dense transform's
// output.  A little above the solvers' own high-water marks.
size_t footprint_of(const Request &r) {
    const bool sparse = r.kind == SPARSE || (r.kind == HANDLE && r.method == SKL_METHOD_SPARSE_SPQR);
    const bool handle = r.kind == HANDLE;
    const double n = (double)r.n, d = (double)r.d, nnz = (double)r.nnz;
    const double m = std::ceil(d * r.opts.gamma);
    double b = 16.0 * n + 16.0 * (d + 1); // b, x and the residual copy
    if (sparse) {
        if (!r.dev && !handle) b += 16.0 * nnz + 8.0 * (d + 1); // int64 indices + values
        if (!handle) b += 16.0 * nnz + 4.0 * (n + d + 2);       // int32 CSC + CSR copy
    } else {
        if (!r.dev && !handle) b += 8.0 * n * d;
        b += 8.0 * n * (d + 1) * 2; // transform output and scratch
        if (!handle && skl::lsqr_rows_wanted(r.n, r.d))
            b += 8.0 * n * (double)skl::lsqr_rows_ld(r.d);
    }
    b += 8.0 * m * (d + 1) * 2;         // sketch + QR copies
    b += 8.0 * d * d * 3;               // R, R^{-1}, products
    b += 8.0 * n * 24 + 8.0 * d * 64;   // LSQR vectors and partials
    return (size_t)b + ((size_t)64 << 20);
}

} // namespace

struct skl_server {
    int slots = 1;
    size_t budget = 0;
    // per slot: device bytes retained between requests, footprint of the
    // request it runs (0 when idle)
    std::vector<size_t> retained, running;
    int busy = 0;
    // LSQR shape autotuning of new buckets before their first request
    bool autotune = false;
    std::set<long long> tuning; // buckets a slot is tuning now
    bool stop = false;
    long next_ticket = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::map<long, Request *> all;
    std::vector<std::thread> workers;
    int device = 0;
};

// device bytes charged to the slots other than `self`
static size_t charged_except(const skl_server *S, int self) {
    size_t t = 0;
    for (int i = 0; i < S->slots; ++i)
        if (i != self) t += std::max(S->retained[(size_t)i], S->running[(size_t)i]);
    return t;
}

static int run_request(Request *r) {
    switch (r->kind) {
    case DENSE_CPQR:
        return r->dev ? skl_dense_solve_dev(r->n, r->d, r->A, r->lda, r->b, r->x, &r->opts, &r->info)
                      : skl_dense_solve(r->n, r->d, r->A, r->lda, r->b, r->x, &r->opts, &r->info);
    case DENSE_NOCPQR:
        return r->dev ? skl_dense_solve_nocpqr_dev(r->n, r->d, r->A, r->lda, r->b, r->x, &r->opts,
                                                   &r->info)
                      : skl_dense_solve_nocpqr(r->n, r->d, r->A, r->lda, r->b, r->x, &r->opts,
                                               &r->info);
    case SPARSE:
        return r->dev ? skl_sparse_solve_dev(r->n, r->d, r->nnz, r->colptr, r->rowind, r->val, r->b,
                                             r->x, &r->opts, &r->info)
                      : skl_sparse_solve(r->n, r->d, r->nnz, r->colptr, r->rowind, r->val, r->b,
                                         r->x, &r->opts, &r->info);
    default:
        return skl_solve_handle(r->h, r->method, r->b, r->x, &r->opts, &r->info);
    }
}

// The request's LSQR shape bucket is tuned before it runs: by this slot when
// no slot is tuning it, else after waiting for the slot that is.
static void tune_for(skl_server *S, Request *r) {
    const bool sparse = r->kind == SPARSE || (r->kind == HANDLE && r->method == SKL_METHOD_SPARSE_SPQR);
    if (skl::lsqr_shape_tuned(sparse, r->n, r->d, r->nnz)) return;
    const long long key = skl::lsqr_shape_bucket(sparse, r->n, r->d, r->nnz);
    {
        std::unique_lock<std::mutex> lk(S->mu);
        if (S->tuning.count(key)) {
            S->cv_done.wait(lk, [&] { return S->tuning.count(key) == 0; });
            return;
        }
        S->tuning.insert(key);
    }
    double best = 0, def = 0;
    (void)skl_autotune_lsqr(sparse ? 1 : 0, r->n, r->d, r->nnz, &best, &def);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->tuning.erase(key);
    }
    S->cv_done.notify_all();
}

static void worker_main(skl_server *S, int self) {
    skl_init(S->device);
    while (true) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_work.wait(lk, [&] {
                if (S->stop) return true;
                if (S->queue.empty()) return false;
                Request *h = S->queue.front();
                // admit in arrival order when the footprint fits (an oversized
                // request runs alone)
                const size_t mine = std::max(S->retained[(size_t)self], h->footprint);
                return charged_except(S, self) + mine <= S->budget || S->busy == 0;
            });
            if (S->stop && S->queue.empty()) break;
            if (S->queue.empty()) continue;
            r = S->queue.front();
            S->queue.pop_front();
            S->running[(size_t)self] = r->footprint;
            S->busy++;
        }
        if (S->autotune) tune_for(S, r);
        r->started = clk::now();
        skl::ctx().mem_allowance = r->footprint;
        r->rc = run_request(r);
        skl::ctx().mem_allowance = 0;
        r->finished = clk::now();
        if (r->rc != SKL_OK) r->err = skl_last_error();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->retained[(size_t)self] = skl::ctx().retained_bytes();
            S->running[(size_t)self] = 0;
            S->busy--;
            r->done = true;
        }
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
    skl::release_thread_context();
}

static long submit(skl_server *S, Request *r) {
    r->footprint = footprint_of(*r);
    r->submitted = clk::now();
    long t;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        t = S->next_ticket++;
        r->ticket = t;
        S->all[t] = r;
        S->queue.push_back(r);
    }
    S->cv_work.notify_all();
    return t;
}

extern "C" {

skl_server *skl_server_create(int slots, size_t mem_budget_bytes, int device) {
    if (slots < 1) slots = 1;
    skl_server *S = new skl_server();
    S->slots = slots;
    S->device = device;
    if (mem_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        if (skl_init(device) == SKL_OK && hipMemGetInfo(&fr, &tot) == hipSuccess)
            mem_budget_bytes = (size_t)(0.85 * (double)fr);
        else
            mem_budget_bytes = (size_t)64 << 30;
    }
    S->budget = mem_budget_bytes;
    S->retained.assign((size_t)slots, 0);
    S->running.assign((size_t)slots, 0);
    for (int i = 0; i < slots; ++i) S->workers.emplace_back(worker_main, S, i);
    return S;
}

static Request *dense_request(int cpqr, bool dev, long n, long d, const double *A, long lda,
                              const double *b, double *x, const skl_opts *opts) {
    Request *r = new Request();
    r->kind = cpqr ? DENSE_CPQR : DENSE_NOCPQR;
    r->dev = dev;
    r->n = n;
    r->d = d;
    r->A = A;
    r->lda = lda;
    r->b = b;
    r->x = x;
    if (opts)
        r->opts = *opts;
    else
        skl_default_opts(&r->opts);
    return r;
}

static Request *sparse_request(bool dev, long n, long d, long nnz, const long *colptr,
                               const long *rowind, const double *val, const double *b, double *x,
                               const skl_opts *opts) {
    Request *r = new Request();
    r->kind = SPARSE;
    r->dev = dev;
    r->n = n;
    r->d = d;
    r->nnz = nnz;
    r->colptr = colptr;
    r->rowind = rowind;
    r->val = val;
    r->b = b;
    r->x = x;
    if (opts)
        r->opts = *opts;
    else
        skl_default_opts(&r->opts);
    return r;
}

long skl_server_submit_dense(skl_server *S, int cpqr, long n, long d, const double *A, long lda,
                             const double *b, double *x, const skl_opts *opts) {
    if (!S || !A || !b || !x) return -SKL_E_ARG;
    return submit(S, dense_request(cpqr, false, n, d, A, lda, b, x, opts));
}

long skl_server_submit_sparse(skl_server *S, long n, long d, long nnz, const long *colptr,
                              const long *rowind, const double *val, const double *b, double *x,
                              const skl_opts *opts) {
    if (!S || !colptr || !b || !x) return -SKL_E_ARG;
    return submit(S, sparse_request(false, n, d, nnz, colptr, rowind, val, b, x, opts));
}

long skl_server_submit_dense_dev(skl_server *S, int cpqr, long n, long d, const double *dA,
                                 long lda, const double *db, double *dx, const skl_opts *opts) {
    if (!S || !dA || !db || !dx) return -SKL_E_ARG;
    return submit(S, dense_request(cpqr, true, n, d, dA, lda, db, dx, opts));
}

long skl_server_submit_sparse_dev(skl_server *S, long n, long d, long nnz, const long *dcolptr,
                                  const long *drowind, const double *dval, const double *db,
                                  double *dx, const skl_opts *opts) {
    if (!S || !dcolptr || !db || !dx) return -SKL_E_ARG;
    return submit(S, sparse_request(true, n, d, nnz, dcolptr, drowind, dval, db, dx, opts));
}

long skl_server_submit_handle(skl_server *S, skl_handle *h, int method, const double *b,
                              double *x, const skl_opts *opts) {
    if (!S || !h || !b || !x) return -SKL_E_ARG;
    long n = 0, d = 0, nnz = 0;
    skl_handle_query(h, &n, &d, &nnz, nullptr);
    Request *r = new Request();
    r->kind = HANDLE;
    r->h = h;
    r->method = method;
    r->n = n;
    r->d = d;
    r->nnz = nnz < 0 ? 0 : nnz;
    r->b = b;
    r->x = x;
    if (opts)
        r->opts = *opts;
    else
        skl_default_opts(&r->opts);
    return submit(S, r);
}

int skl_server_wait(skl_server *S, long ticket, skl_info *info, double *queue_s, double *run_s) {
    Request *r = nullptr;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->all.find(ticket);
        if (it == S->all.end()) return SKL_E_ARG;
        r = it->second;
        S->cv_done.wait(lk, [&] { return r->done; });
        S->all.erase(it);
    }
    if (info) *info = r->info;
    using sec = std::chrono::duration<double>;
    if (queue_s) *queue_s = sec(r->started - r->submitted).count();
    if (run_s) *run_s = sec(r->finished - r->started).count();
    int rc = r->rc;
    if (rc != SKL_OK) skl::set_last_error(r->err);
    delete r;
    return rc;
}

void skl_server_set_autotune(skl_server *S, int on) {
    if (!S) return;
    std::lock_guard<std::mutex> lk(S->mu);
    S->autotune = on != 0;
}

int skl_server_pending(skl_server *S) {
    std::lock_guard<std::mutex> lk(S->mu);
    return (int)S->queue.size() + S->busy;
}

void skl_server_destroy(skl_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &t : S->workers) t.join();
    for (auto &kv : S->all) delete kv.second;
    delete S;
}

} // extern "C"
