// ski_lls.cpp -- the ski-lls.h entry points, forwarding to the C ABI.
#include "../../include/ski-lls.h"

#include <cstdlib>
#include <ctime>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

unsigned seed_or(unsigned fallback) {
    const char *e = std::getenv("SKL_SEED");
    if (e && *e) return (unsigned)std::strtoul(e, nullptr, 10);
    return fallback;
}

// SKL_PRECISION=mixed selects the FP32 sketch/QR with FP64 LSQR
int precision_from_env() {
    const char *e = std::getenv("SKL_PRECISION");
    return (e && std::string(e) == "mixed") ? SKL_MIXED : SKL_FP64;
}

void raise(int rc) {
    if (rc == SKL_OK) return;
    if (rc == SKL_E_SHAPE) throw std::runtime_error("Matrix is not overdetermined.");
    if (rc == SKL_E_RANK0) throw std::runtime_error("SPQR Factorization fails");
    std::string msg = skl_error_string(rc);
    const char *last = skl_last_error();
    if (last && *last) msg += std::string(": ") + last;
    throw std::runtime_error(msg);
}

// contiguous copy of a strided vector when needed
std::vector<double> gather(const Vec_d &v) {
    std::vector<double> h((size_t)v.n());
    for (long i = 0; i < v.n(); ++i) h[(size_t)i] = v(i);
    return h;
}

} // namespace

void ls_dense_hashing_blendenpik(Mat_d &A, Vec_d &b, Vec_d &x, long &rank, int &flag, long &it,
                                 double gamma, long k, double abs_tol, double rcond,
                                 double it_tol, long max_it, int debug, int wisdom) {
    assert(A.m() == b.n());
    assert(A.n() == x.n());
    if (A.m() < A.n()) throw std::runtime_error("Matrix is not overdetermined.");
    skl_opts o;
    skl_default_opts(&o);
    o.precision = precision_from_env();
    o.gamma = gamma;
    o.k = k;
    o.abs_tol = abs_tol;
    o.rcond = rcond;
    o.it_tol = it_tol;
    o.max_it = max_it;
    o.debug = debug;
    o.wisdom = wisdom;
    o.seed = seed_or((unsigned)std::time(nullptr));
    std::vector<double> hb = gather(b), hx((size_t)A.n());
    skl_info info;
    raise(skl_dense_solve(A.m(), A.n(), A.data(), A.ld(), hb.data(), hx.data(), &o, &info));
    for (long i = 0; i < x.n(); ++i) x(i) = hx[(size_t)i];
    rank = info.rank;
    flag = info.flag;
    it = info.it;
}

void ls_dense_hashing_blendenpik_noCPQR(Mat_d &A, Vec_d &b, Vec_d &x, long &rank, int &flag,
                                        long &it, double gamma, long k, double it_tol,
                                        long max_it, int debug, int wisdom) {
    assert(A.m() == b.n());
    assert(A.n() == x.n());
    if (A.m() < A.n()) throw std::runtime_error("Matrix is not overdetermined.");
    skl_opts o;
    skl_default_opts(&o);
    o.precision = precision_from_env();
    o.gamma = gamma;
    o.k = k;
    o.it_tol = it_tol;
    o.max_it = max_it;
    o.debug = debug;
    o.wisdom = wisdom;
    o.seed = seed_or((unsigned)std::time(nullptr));
    std::vector<double> hb = gather(b), hx((size_t)A.n());
    skl_info info;
    raise(skl_dense_solve_nocpqr(A.m(), A.n(), A.data(), A.ld(), hb.data(), hx.data(), &o, &info));
    for (long i = 0; i < x.n(); ++i) x(i) = hx[(size_t)i];
    rank = info.rank;
    flag = info.flag;
    it = info.it;
}

void ls_dense_blendenpik(Mat_d &A, Vec_d &b, Vec_d &x, long &rank, int &flag, long &it,
                         double gamma, double it_tol, long max_it, int debug, int wisdom) {
    assert(A.m() == b.n());
    assert(A.n() == x.n());
    if (A.m() < A.n()) throw std::runtime_error("Matrix is not overdetermined.");
    skl_opts o;
    skl_default_opts(&o);
    o.precision = precision_from_env();
    o.gamma = gamma;
    o.it_tol = it_tol;
    o.max_it = max_it;
    o.debug = debug;
    o.wisdom = wisdom;
    o.seed = seed_or((unsigned)std::time(nullptr));
    std::vector<double> hb = gather(b), hx((size_t)A.n());
    skl_info info;
    raise(skl_dense_solve_blendenpik(A.m(), A.n(), A.data(), A.ld(), hb.data(), hx.data(), &o,
                                     &info));
    for (long i = 0; i < x.n(); ++i) x(i) = hx[(size_t)i];
    rank = info.rank;
    flag = info.flag;
    it = info.it;
}

void ls_sparse_spqr(cs_dl &A, Vec_d &b, Vec_d &x, long &rank, int &flag, long &it, double gamma,
                    long k, double abs_tol, int ordering, double it_tol, long max_it,
                    double rcond_threshold, double perturb, int debug) {
    assert(A.m == b.n());
    assert(A.n == x.n());
    if (A.m < A.n) throw std::runtime_error("Matrix is not overdetermined.");
    skl_opts o;
    skl_default_opts(&o);
    o.precision = precision_from_env();
    o.gamma = gamma;
    o.k = k;
    o.abs_tol = abs_tol;
    o.ordering = ordering;
    o.it_tol = it_tol;
    o.max_it = max_it;
    o.rcond_threshold = rcond_threshold;
    o.perturb = perturb;
    o.debug = debug;
    o.seed = seed_or(1u);
    std::vector<double> hb = gather(b), hx((size_t)A.n);
    skl_info info;
    long nnz = A.p[A.n];
    raise(skl_sparse_solve(A.m, A.n, nnz, A.p, A.i, A.x, hb.data(), hx.data(), &o, &info));
    for (long i = 0; i < x.n(); ++i) x(i) = hx[(size_t)i];
    rank = info.rank;
    flag = info.flag;
    it = info.it;
}
