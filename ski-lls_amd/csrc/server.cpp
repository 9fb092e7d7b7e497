// server.cpp -- continuous-batching request loop over the solvers.
//
// A server owns `slots` worker threads; each worker has its own HIP stream
// and device arena (the per-thread context), so solves admitted together
// overlap on the device.  Requests wait in arrival order and are admitted
// when a slot is free and their device-memory footprint fits the budget;
// completion frees both.  Every solve reduces in a fixed order on its own
// buffers, so a result does not depend on what else is running.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <string>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "skl_internal.h"

namespace skl {
void release_thread_context();
}

namespace {

using clk = std::chrono::steady_clock;

struct Request {
    long ticket = 0;
    int kind = 0; // 0 dense robust, 1 dense noCPQR, 2 sparse
    long n = 0, d = 0, nnz = 0;
    const double *A = nullptr;
    long lda = 0;
    const long *colptr = nullptr, *rowind = nullptr;
    const double *val = nullptr;
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

// device bytes one solve needs: inputs + sketch/QR workspace + LSQR buffers
size_t footprint_of(const Request &r) {
    double m = std::ceil((double)r.d * r.opts.gamma);
    double b = 0;
    if (r.kind == 2)
        b = 36.0 * r.nnz + 16.0 * r.n + 16.0 * (r.d + 1);
    else
        b = 8.0 * r.n * r.d + 16.0 * r.n;
    b += 8.0 * m * (r.d + 1) * 2;            // sketch + QR copies
    b += 8.0 * r.d * r.d * 3;                // R, R^{-1}, products
    b += 8.0 * r.n * 24 + 8.0 * r.d * 64;    // LSQR vectors and partials
    if (r.kind != 2) b += 8.0 * r.n * (r.d + 1) * 2; // transform scratch
    return (size_t)b + ((size_t)64 << 20);
}

} // namespace

struct skl_server {
    int slots = 1;
    size_t budget = 0, in_use = 0;
    int busy = 0;
    bool stop = false;
    long next_ticket = 1;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> queue;
    std::map<long, Request *> all;
    std::vector<std::thread> workers;
    int device = 0;
};

static void worker_main(skl_server *S) {
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
                return S->in_use + h->footprint <= S->budget || S->busy == 0;
            });
            if (S->stop && S->queue.empty()) break;
            if (S->queue.empty()) continue;
            r = S->queue.front();
            S->queue.pop_front();
            S->in_use += r->footprint;
            S->busy++;
        }
        r->started = clk::now();
        switch (r->kind) {
        case 0:
            r->rc = skl_dense_solve(r->n, r->d, r->A, r->lda, r->b, r->x, &r->opts, &r->info);
            break;
        case 1:
            r->rc = skl_dense_solve_nocpqr(r->n, r->d, r->A, r->lda, r->b, r->x, &r->opts, &r->info);
            break;
        default:
            r->rc = skl_sparse_solve(r->n, r->d, r->nnz, r->colptr, r->rowind, r->val, r->b, r->x,
                                     &r->opts, &r->info);
            break;
        }
        r->finished = clk::now();
        if (r->rc != SKL_OK) r->err = skl_last_error();
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->in_use -= r->footprint;
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
    for (int i = 0; i < slots; ++i) S->workers.emplace_back(worker_main, S);
    return S;
}

long skl_server_submit_dense(skl_server *S, int cpqr, long n, long d, const double *A, long lda,
                             const double *b, double *x, const skl_opts *opts) {
    if (!S || !A || !b || !x) return -SKL_E_ARG;
    Request *r = new Request();
    r->kind = cpqr ? 0 : 1;
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
    return submit(S, r);
}

long skl_server_submit_sparse(skl_server *S, long n, long d, long nnz, const long *colptr,
                              const long *rowind, const double *val, const double *b, double *x,
                              const skl_opts *opts) {
    if (!S || !colptr || !b || !x) return -SKL_E_ARG;
    Request *r = new Request();
    r->kind = 2;
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
