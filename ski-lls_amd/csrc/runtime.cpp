// runtime.cpp -- context, device arena, errors and the host-side hashing
// pattern generator of the MI355X least-squares library.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "skl_internal.h"

namespace skl {

[[noreturn]] void throw_hip(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e,
                  hipGetErrorString(e), what, file, line);
    int code = (e == hipErrorOutOfMemory) ? SKL_E_NOMEM : SKL_E_HIP;
    throw Error(code, buf);
}

double wall() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
Arena::~Arena() { release(); }

void Arena::attach_fixed(void *base, size_t bytes) {
    if (!fixed_) release();
    fixed_ = true;
    base_ = static_cast<char *>(base);
    cap_ = bytes;
    used_ = over_ = peak_ = 0;
}

void Arena::release() {
    if (fixed_) {
        used_ = 0;
        return;
    }
    if (base_ || !overflow_.empty()) {
        AllocLock lk(capture_mutex());
        if (base_) (void)hipFree(base_);
        for (void *p : overflow_) (void)hipFree(p);
    }
    overflow_.clear();
    base_ = nullptr;
    cap_ = used_ = 0;
}

void Arena::reserve(size_t bytes) {
    if (bytes <= cap_ || fixed_) return;
    if (used_ != 0) return; // cannot move live buffers; alloc() falls back
    AllocLock lk(capture_mutex());
    if (base_) SKL_HIP(hipFree(base_));
    base_ = nullptr;
    cap_ = 0;
    for (void *p : overflow_) (void)hipFree(p);
    overflow_.clear();
    size_t want = bytes + (bytes >> 3);
    SKL_HIP(hipMalloc(&base_, want));
    cap_ = want;
    over_ = peak_ = 0;
}

void Arena::reset() {
    if (!fixed_ && !overflow_.empty()) {
        // hipFree waits for the device; the previous solve's buffers are dead
        size_t want = std::max(peak_, cap_);
        release();
        reserve(want);
    }
    used_ = 0;
    over_ = 0;
    peak_ = 0;
}

void *Arena::alloc(size_t bytes) {
    size_t sz = (bytes + 255) & ~size_t(255);
    if (sz == 0) sz = 256;
    if (used_ + sz <= cap_) {
        void *p = base_ + used_;
        used_ += sz;
        peak_ = std::max(peak_, used_ + over_);
        return p;
    }
    if (fixed_)
        throw Error(SKL_E_UNSUPPORTED,
                    "problem needs more device scratch than the no-allocation entry point's " +
                        std::to_string(cap_ >> 20) + " MiB");
    void *p = nullptr;
    {
        AllocLock lk(capture_mutex());
        SKL_HIP(hipMalloc(&p, sz));
    }
    overflow_.push_back(p);
    over_ += sz;
    peak_ = std::max(peak_, used_ + over_);
    return p;
}

size_t Context::retained_bytes() const { return arena.capacity() + ws_cap + ws_side_cap; }

// ---------------------------------------------------------------------------
static thread_local Context *t_ctx = nullptr;
static std::mutex g_init_mu;
static std::atomic<int> g_live_ctx{0};     // host threads holding a context
static int g_device = -1;

void ensure_init() {
    if (t_ctx) return;
    std::lock_guard<std::mutex> lk(g_init_mu);
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) throw Error(SKL_E_NODEVICE, "no HIP device available");
    int dev = g_device >= 0 ? g_device : 0;
    SKL_HIP(hipSetDevice(dev));
    AllocLock alk(capture_mutex());
    Context *c = new Context();
    c->device = dev;
    SKL_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    SKL_HIP(hipGetDeviceProperties(&prop, dev));
    c->num_cu = prop.multiProcessorCount;
    c->coop_ok = prop.cooperativeLaunch;
    c->lds_max = prop.sharedMemPerBlock;
    SKL_HIP(hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault));
    SKL_HIP(hipMalloc(&c->dev_err, 256));
    SKL_HIP(hipMemset(c->dev_err, 0, 256));
    noalloc_setup(dev);
    t_ctx = c;
    g_live_ctx.fetch_add(1);
}

namespace {
struct PendingBarrier {
    hipEvent_t ev;
    int tokens;
};
struct UnrecordedBarrier {
    hipStream_t s;
    int tokens;
};
std::mutex g_tok_mu;
int g_tok_cap = 0, g_tok_free = 0, g_tok_unrec_total = 0;
std::vector<PendingBarrier> g_tok_pending;
std::vector<UnrecordedBarrier> g_tok_unrec; // launches with no event after them yet
std::vector<hipEvent_t> g_tok_events;

// returns the tokens of completed launches (caller holds g_tok_mu)
void reap_barriers() {
    for (size_t i = 0; i < g_tok_pending.size();) {
        if (hipEventQuery(g_tok_pending[i].ev) == hipSuccess) {
            g_tok_free += g_tok_pending[i].tokens;
            g_tok_events.push_back(g_tok_pending[i].ev);
            g_tok_pending[i] = g_tok_pending.back();
            g_tok_pending.pop_back();
        } else {
            (void)hipGetLastError(); // hipErrorNotReady is not an error here
            ++i;
        }
    }
}

// Records one event after the launches that have none yet (one per stream),
// from the event pool, refilled first outside the token lock.  The records
// happen under the lock, so a stream that release_thread_context is about to
// destroy (it takes its entries out under the same lock) is never recorded on.
void flush_unrecorded() {
    size_t want;
    {
        std::lock_guard<std::mutex> lk(g_tok_mu);
        want = g_tok_unrec.size();
        want = want > g_tok_events.size() ? want - g_tok_events.size() : 0;
    }
    std::vector<hipEvent_t> fresh;
    if (want && !ctx().no_alloc) {
        AllocLock alk(capture_mutex());
        for (size_t i = 0; i < want; ++i) {
            hipEvent_t ev = nullptr;
            if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess) fresh.push_back(ev);
            else (void)hipGetLastError();
        }
    }
    std::lock_guard<std::mutex> lk(g_tok_mu);
    g_tok_events.insert(g_tok_events.end(), fresh.begin(), fresh.end());
    for (size_t i = 0; i < g_tok_unrec.size();) {
        const UnrecordedBarrier u = g_tok_unrec[i];
        hipEvent_t ev = nullptr;
        if (!g_tok_events.empty()) {
            ev = g_tok_events.back();
            g_tok_events.pop_back();
        }
        if (ev && hipEventRecord(ev, u.s) == hipSuccess) {
            g_tok_pending.push_back({ev, u.tokens});
        } else {
            // no event to wait on: hand the tokens back after the stream drains
            (void)hipGetLastError();
            (void)hipStreamSynchronize(u.s);
            g_tok_free += u.tokens;
            if (ev) g_tok_events.push_back(ev);
        }
        g_tok_unrec_total -= u.tokens;
        g_tok_unrec[i] = g_tok_unrec.back();
        g_tok_unrec.pop_back();
    }
}
} // namespace

BarrierLaunch::BarrierLaunch(int grid, bool exclusive) {
    const int num_cu = ctx().num_cu;
    std::unique_lock<std::mutex> lk(g_tok_mu);
    if (g_tok_cap == 0) g_tok_cap = g_tok_free = num_cu;
    const int need = exclusive ? g_tok_cap : std::max(1, std::min(grid, g_tok_cap));
    while (g_tok_free < need) {
        reap_barriers();
        if (g_tok_free >= need) break;
        const bool unrec = !g_tok_unrec.empty();
        lk.unlock();
        if (unrec)
            flush_unrecorded();
        else
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
    }
    g_tok_free -= need;
    tokens_ = need;
}

// The tokens wait for an event after the launch, but one event serves every
// launch before it on the stream: the event is recorded only once launches
// without one hold half the tokens, or when a launch waits for tokens (an
// event record costs the stream several microseconds, about as much as the
// short fallback kernels behind these guards).
BarrierLaunch::~BarrierLaunch() {
    bool flush;
    {
        std::lock_guard<std::mutex> lk(g_tok_mu);
        const hipStream_t s = ctx().stream;
        if (!g_tok_unrec.empty() && g_tok_unrec.back().s == s)
            g_tok_unrec.back().tokens += tokens_;
        else
            g_tok_unrec.push_back({s, tokens_});
        g_tok_unrec_total += tokens_;
        flush = 2 * g_tok_unrec_total >= g_tok_cap;
    }
    if (flush) flush_unrecorded();
}

// A drained stream's launches hold no tokens any more (release_thread_context,
// before it destroys the stream).
void barrier_tokens_stream_drained(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_tok_mu);
    for (size_t i = 0; i < g_tok_unrec.size();) {
        if (g_tok_unrec[i].s == s) {
            g_tok_free += g_tok_unrec[i].tokens;
            g_tok_unrec_total -= g_tok_unrec[i].tokens;
            g_tok_unrec[i] = g_tok_unrec.back();
            g_tok_unrec.pop_back();
        } else {
            ++i;
        }
    }
}

Context &ctx() {
    ensure_init();
    return *t_ctx;
}

bool has_context() { return t_ctx != nullptr; }

std::shared_mutex &capture_mutex() {
    static std::shared_mutex mu;
    return mu;
}

void lsqr_release_graphs();

void *Context::workspace(size_t bytes) {
    if (bytes > ws_cap) {
        if (no_alloc) throw Error(SKL_E_UNSUPPORTED, "split-K scratch in no-allocation mode");
        AllocLock lk(capture_mutex());
        SKL_HIP(hipStreamSynchronize(stream));
        if (ws) SKL_HIP(hipFree(ws));
        ws = nullptr;
        ws_cap = 0;
        SKL_HIP(hipMalloc(&ws, bytes));
        ws_cap = bytes;
    }
    return ws;
}

void Context::make_side() {
    if (side) return;
    if (no_alloc) throw Error(SKL_E_UNSUPPORTED, "second stream in no-allocation mode");
    AllocLock lk(capture_mutex());
    // the second stream runs bulk GEMMs beside the critical path's
    // single-workgroup kernels (panel factorizations, the pivot search),
    // which otherwise share CUs with the GEMM's workgroups: the side stream
    // is kept off the first num_cu / 8 CUs of the mask (an XCD's worth; the
    // first 32 bits are one XCD on MI355X), SKL_SIDE_CU_RESERVE = R sets R
    // (0: no mask).  Measured: the sparse 120000 x 5000 FP64 solve's QR
    // 28.9 -> 26.5 ms; the dense 40000 x 4000 solve unchanged (its side GEMM
    // overlaps the pivot search by a third)
    const char *rs = getenv("SKL_SIDE_CU_RESERVE");
    const int reserve = rs ? std::max(0, atoi(rs)) : (num_cu >= 64 ? num_cu / 8 : 0);
    if (reserve > 0 && reserve < num_cu) {
        std::vector<uint32_t> mask((size_t)(num_cu + 31) / 32, 0u);
        for (int c = reserve; c < num_cu; ++c) mask[(size_t)c / 32] |= 1u << (c % 32);
        SKL_HIP(hipExtStreamCreateWithCUMask(&side, (uint32_t)mask.size(), mask.data()));
    } else {
        SKL_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    }
    for (auto &e : ev) SKL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

void Context::make_aux() {
    if (aux) return;
    if (no_alloc) throw Error(SKL_E_UNSUPPORTED, "third stream in no-allocation mode");
    AllocLock lk(capture_mutex());
    SKL_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
    for (auto &e : aux_ev) SKL_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

OnSideStream::OnSideStream() {
    Context &c = ctx();
    c.make_side();
    std::swap(c.stream, c.side);
    std::swap(c.ws, c.ws_side);
    std::swap(c.ws_cap, c.ws_side_cap);
}

OnSideStream::~OnSideStream() {
    Context &c = ctx();
    std::swap(c.stream, c.side);
    std::swap(c.ws, c.ws_side);
    std::swap(c.ws_cap, c.ws_side_cap);
}

// Frees the calling thread's streams, arena and cached graphs (worker exit).
void release_thread_context() {
    if (!t_ctx) return;
    (void)hipStreamSynchronize(t_ctx->stream);
    if (t_ctx->side) (void)hipStreamSynchronize(t_ctx->side);
    if (t_ctx->aux) (void)hipStreamSynchronize(t_ctx->aux);
    barrier_tokens_stream_drained(t_ctx->stream);
    if (t_ctx->side) barrier_tokens_stream_drained(t_ctx->side);
    if (t_ctx->aux) barrier_tokens_stream_drained(t_ctx->aux);
    lsqr_release_graphs();
    t_ctx->arena.release(); // (takes the capture lock itself)
    AllocLock alk(capture_mutex());
    if (t_ctx->ws) (void)hipFree(t_ctx->ws);
    if (t_ctx->ws_side) (void)hipFree(t_ctx->ws_side);
    for (auto &e : t_ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (t_ctx->side) (void)hipStreamDestroy(t_ctx->side);
    for (auto &e : t_ctx->aux_ev)
        if (e) (void)hipEventDestroy(e);
    if (t_ctx->aux) (void)hipStreamDestroy(t_ctx->aux);
    if (t_ctx->pinned) (void)hipHostFree(t_ctx->pinned);
    if (t_ctx->dev_err) (void)hipFree(t_ctx->dev_err);
    if (t_ctx->ctr_pool) (void)hipFree(t_ctx->ctr_pool);
    (void)hipStreamDestroy(t_ctx->stream);
    alk.unlock(); // the arenas' destructors below own nothing any more
    delete t_ctx;
    t_ctx = nullptr;
    g_live_ctx.fetch_sub(1);
}

void set_device(int dev) { g_device = dev; }

GridSync grid_sync(unsigned *ctr) {
    Context &c = ctx();
    return GridSync{ctr, c.dev_err, c.barrier_spins};
}

// Counters of the pool are handed out once each; a barrier launch leaves its
// counter at the grid's arrival total, so the pool is zeroed again (after the
// context's streams drain) when every counter has been used.
static constexpr int CTR_POOL = 8192;

GridSync grid_sync_fresh(Arena &ar) {
    Context &c = ctx();
    if (c.no_alloc) {
        unsigned *p = ar.alloc_n<unsigned>(1);
        SKL_HIP(hipMemsetAsync(p, 0, sizeof(unsigned), c.stream));
        return grid_sync(p);
    }
    if (!c.ctr_pool) {
        AllocLock alk(capture_mutex());
        SKL_HIP(hipMalloc(&c.ctr_pool, sizeof(unsigned) * CTR_POOL));
        SKL_HIP(hipMemsetAsync(c.ctr_pool, 0, sizeof(unsigned) * CTR_POOL, c.stream));
        c.ctr_next = 0;
    } else if (c.ctr_next == CTR_POOL) {
        if (c.side) SKL_HIP(hipStreamSynchronize(c.side));
        if (c.aux) SKL_HIP(hipStreamSynchronize(c.aux));
        SKL_HIP(hipMemsetAsync(c.ctr_pool, 0, sizeof(unsigned) * CTR_POOL, c.stream));
        c.ctr_next = 0;
    }
    return grid_sync(c.ctr_pool + c.ctr_next++);
}

void check_grid_fault() {
    Context &c = ctx();
    // the last 64 bytes of the pinned block are this check's
    unsigned *h = reinterpret_cast<unsigned *>(static_cast<char *>(c.pinned) + 4096 - 64);
    *h = 0;
    SKL_HIP(hipMemcpyAsync(h, c.dev_err, sizeof(unsigned), hipMemcpyDeviceToHost, c.stream));
    SKL_HIP(hipStreamSynchronize(c.stream));
    if (*h != 0) {
        SKL_HIP(hipMemsetAsync(c.dev_err, 0, sizeof(unsigned), c.stream));
        SKL_HIP(hipStreamSynchronize(c.stream));
        throw Error(SKL_E_HIP, "grid barrier gave up waiting for the other workgroups of a "
                               "factorization or LSQR kernel; its results were discarded");
    }
}

// ---------------------------------------------------------------------------
// glibc srandom_r / random_r (TYPE_3: degree 31, separation 3): the generator
// behind libc rand() that the reference draws its hashing pattern and signs
// from (src/dense_helpers.cpp:12-14, src/sparse_helpers.cpp:74-116).
// ---------------------------------------------------------------------------
void Glibc::seed(unsigned sd) {
    if (sd == 0) sd = 1;
    s[0] = (int32_t)sd;
    int32_t word = (int32_t)sd;
    for (int i = 1; i < 31; ++i) {
        long hi = word / 127773;
        long lo = word % 127773;
        word = (int32_t)(16807 * lo - 2836 * hi);
        if (word < 0) word += 2147483647;
        s[i] = word;
    }
    f = 3;
    r = 0;
    for (int i = 0; i < 310; ++i) (void)next();
    toggle = 0;
    b2 = 0.0;
}

// Marsaglia polar method on rand()/RAND_MAX with an alternate-call cache
// (hqrrp/lapack_compatible_sources/NoFLA_HQRRP_WY_blk_var4_skint.c:603-624).
#pragma clang fp contract(off)
double Glibc::normal() {
    if (toggle == 1) {
        toggle = 0;
        return 0.0 + 1.0 * b2;
    }
    double c1, c2, a;
    do {
        c1 = -1.0 + 2.0 * ((double)next() / 2147483647);
        c2 = -1.0 + 2.0 * ((double)next() / 2147483647);
        a = c1 * c1 + c2 * c2;
    } while (a == 0 || a >= 1);
    double factor = std::sqrt((-2 * std::log(a)) / a);
    b2 = c2 * factor;
    toggle = 1;
    return 0.0 + 1.0 * (c1 * factor);
}

// ---- jump-ahead for the generator -------------------------------------------
// With y_0..y_30 the generator's last 31 raw values (oldest first), the raw
// sequence obeys y_i = y_{i-31} + y_{i-3} (mod 2^32), a linear recurrence with
// characteristic polynomial t^31 - t^28 - 1.  If t^N = sum_j c_j t^j modulo
// that polynomial, then y_{N+m} = sum_j c_j y_{j+m} exactly in uint32
// arithmetic, so any later window follows from the current one.
namespace {
typedef std::vector<uint32_t> Poly; // 31 coefficients, degree < 31

// reduce a polynomial of degree < 61 modulo t^31 - t^28 - 1 (t^31 = t^28 + 1)
Poly poly_reduce(std::vector<uint32_t> a) {
    for (int k = (int)a.size() - 1; k >= 31; --k) {
        const uint32_t c = a[(size_t)k];
        if (!c) continue;
        a[(size_t)k] = 0;
        a[(size_t)k - 3] += c;
        a[(size_t)k - 31] += c;
    }
    a.resize(31);
    return a;
}

Poly poly_mul(const Poly &x, const Poly &y) {
    std::vector<uint32_t> a(61, 0u);
    for (int i = 0; i < 31; ++i)
        if (x[(size_t)i])
            for (int j = 0; j < 31; ++j) a[(size_t)(i + j)] += x[(size_t)i] * y[(size_t)j];
    return poly_reduce(std::move(a));
}

// t^N modulo the characteristic polynomial
Poly poly_tpow(uint64_t N) {
    Poly r(31, 0u), b(31, 0u);
    r[0] = 1;
    b[1] = 1;
    while (N) {
        if (N & 1) r = poly_mul(r, b);
        N >>= 1;
        if (N) b = poly_mul(b, b);
    }
    return r;
}

// the window y_N..y_{N+30} from y_0..y_60 (first 31 = current window)
void jump_window(const std::vector<uint32_t> &y, uint64_t N, uint32_t *w) {
    Poly c = poly_tpow(N);
    for (int m = 0; m < 31; ++m) {
        uint32_t v = 0;
        for (int j = 0; j < 31; ++j) v += c[(size_t)j] * y[(size_t)(j + m)];
        w[m] = v;
    }
}
} // namespace

void Glibc::fill(int *out, long count) {
    if (count <= 0) return;
    const long chunk = 1L << 16;
    unsigned hw = std::thread::hardware_concurrency();
    long P = std::min<long>(std::min<long>(8, hw ? (long)hw : 1), count / chunk);
    if (P <= 1) {
        for (long i = 0; i < count; ++i) out[i] = next();
        return;
    }
    // the current window and 30 values beyond it (for shifted windows)
    std::vector<uint32_t> y(61);
    for (int j = 0; j < 31; ++j) y[(size_t)j] = (uint32_t)s[(f + j) % 31];
    for (int j = 31; j < 61; ++j) y[(size_t)j] = y[(size_t)j - 31] + y[(size_t)j - 3];
    const long per = (count + P - 1) / P;
    std::vector<Glibc> part((size_t)P);
    std::vector<std::thread> th;
    for (long p = 0; p < P; ++p) {
        th.emplace_back([&, p] {
            const long b = p * per, e = std::min(count, b + per);
            // output i is raw y_{31 + i}: the window before it starts at y_i
            uint32_t w[31];
            jump_window(y, (uint64_t)b, w);
            Glibc &G = part[(size_t)p];
            for (int j = 0; j < 31; ++j) G.s[j] = (int32_t)w[j];
            G.f = 0;
            G.r = 28;
            for (long i = b; i < e; ++i) out[i] = G.next();
        });
    }
    for (auto &t : th) t.join();
    // continue from the state after the last chunk
    const Glibc &L = part[(size_t)P - 1];
    std::memcpy(s, L.s, sizeof s);
    f = L.f;
    r = L.r;
}

void gen_signs(Glibc &g, long n, std::vector<double> &D) {
    D.resize((size_t)n);
    std::vector<int> R((size_t)n);
    g.fill(R.data(), n);
    for (long i = 0; i < n; ++i) D[(size_t)i] = 1.0 * (double)((R[(size_t)i] % 2) * 2 - 1);
}

// Division by a runtime constant without the hardware divider: the generator
// output is < 2^31, so q = (x * magic) >> 64 with a 64-bit magic is exact.
struct FastMod {
    uint64_t d, magic;
    explicit FastMod(uint64_t dd) : d(dd), magic(dd > 1 ? ~uint64_t(0) / dd + 1 : 0) {}
    inline uint64_t mod(uint64_t x) const {
        if (d == 1) return 0;
        uint64_t q = (uint64_t)(((unsigned __int128)x * magic) >> 64);
        uint64_t r = x - q * d;
        if (r >= d) r -= d; // guard the rare off-by-one of the approximation
        return r;
    }
};

// select_k_from_n on an index array that persists across columns, then the
// k*n sign draws (src/sparse_helpers.cpp:74-116).
void gen_hashing(Glibc &g, long m, long n, long k, Pattern &p) {
    p.m = m;
    p.n = n;
    p.k = k;
    p.rows.resize((size_t)(n * k));
    p.vals.resize((size_t)(n * k));
    std::vector<int32_t> arr((size_t)m);
    for (long i = 0; i < m; ++i) arr[(size_t)i] = (int32_t)i;
    std::vector<FastMod> mods;
    for (long t = 0; t < k; ++t) mods.emplace_back((uint64_t)(m - t));
    int32_t *rows = p.rows.data();
    int32_t *a = arr.data();
    // all draws first (k per source row, then one sign per entry), so that the
    // generator runs in parallel chunks and only the swaps stay sequential
    std::vector<int> R((size_t)(2 * n * k));
    g.fill(R.data(), 2 * n * k);
    const int *rr = R.data();
    for (long i = 0; i < n; ++i) {
        for (long t = 0; t < k; ++t) {
            long j = (long)mods[(size_t)t].mod((uint64_t)rr[k * i + t]) + t;
            int32_t tmp = a[j];
            a[j] = a[t];
            a[t] = tmp;
        }
        for (long t = 0; t < k; ++t) rows[k * i + t] = a[t];
    }
    double scaling = std::sqrt((double)k);
    double pos = 1 / scaling * 1.0, neg = 1 / scaling * -1.0;
    double *v = p.vals.data();
    const int *rs = rr + n * k;
    for (long q = 0; q < n * k; ++q) v[q] = (rs[q] % 2) ? pos : neg;
}

// Bucket lists: a stable counting sort of the pattern by target row, so the
// members of each bucket appear in increasing source index -- the order in
// which the reference's scatter (SpMat.cpp:53-55, cs_gaxpy) accumulates.
void build_buckets(Pattern &p) {
    long m = p.m, nk = p.n * p.k;
    p.bptr.assign((size_t)m + 1, 0);
    for (long q = 0; q < nk; ++q) p.bptr[(size_t)p.rows[(size_t)q] + 1]++;
    for (long h = 0; h < m; ++h) p.bptr[(size_t)h + 1] += p.bptr[(size_t)h];
    p.bsrc.resize((size_t)nk);
    p.bslot.resize((size_t)nk);
    std::vector<int32_t> cur(p.bptr.begin(), p.bptr.end() - 1);
    for (long q = 0; q < nk; ++q) {
        int32_t h = p.rows[(size_t)q];
        int32_t pos = cur[(size_t)h]++;
        p.bsrc[(size_t)pos] = (int32_t)(q / p.k);
        p.bslot[(size_t)pos] = (int32_t)q;
    }
}

// Uniform row sampling of Blendenpik (src/dense_helpers.cpp:19-39): a partial
// Fisher-Yates shuffle of 0..n-1 whose first m entries are the sampled rows
// (select_k_from_n, src/sparse_helpers.cpp:74-87), as m single-source
// buckets with coefficient 1: sketch row i is row idx[i] of the transform.
void gen_subsample(Glibc &g, long n, long m, Pattern &p) {
    p.m = m;
    p.n = n;
    p.k = 1;
    std::vector<int32_t> arr((size_t)n);
    for (long i = 0; i < n; ++i) arr[(size_t)i] = (int32_t)i;
    for (long i = 0; i < m; ++i) {
        const long j = (long)g.next() % (n - i) + i;
        std::swap(arr[(size_t)j], arr[(size_t)i]);
    }
    p.rows.clear();
    p.bptr.resize((size_t)m + 1);
    p.bsrc.assign(arr.begin(), arr.begin() + m);
    p.bslot.resize((size_t)m);
    p.vals.assign((size_t)m, 1.0);
    for (long i = 0; i <= m; ++i) p.bptr[(size_t)i] = (int32_t)i;
    for (long i = 0; i < m; ++i) p.bslot[(size_t)i] = (int32_t)i;
}

void gen_gaussian(unsigned seed, long rows, long cols, std::vector<double> &G) {
    Glibc g;
    g.seed(seed);
    G.resize((size_t)(rows * cols));
    for (long j = 0; j < cols; ++j)
        for (long i = 0; i < rows; ++i) G[(size_t)(i + j * rows)] = g.normal();
}

} // namespace skl
