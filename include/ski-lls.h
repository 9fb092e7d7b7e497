/*
 * ski-lls.h -- drop-in replacement of the reference's public header
 * (include/ski-lls.h:15-28) backed by the MI355X library (skl.h).
 *
 * Same three entry points, same argument meaning, same exceptions
 * (std::runtime_error "Matrix is not overdetermined." / "SPQR Factorization
 * fails").  Mat_d / Vec_d are small column-major containers with the subset
 * of the reference's Mat<T>/Vec<T> interface that callers of these entry
 * points use (include/Mat.hpp:39-209, include/Vec.hpp:46-141); cs_dl has the
 * CXSparse layout and is only declared when CXSparse's cs.h is not included.
 *
 * Seeds: the reference seeds the dense path with time(NULL)
 * (src/dense_helpers.cpp:12,92) and leaves the sparse path on the
 * process-wide rand() stream (seed 1 in a fresh process).  These wrappers do
 * the same unless the environment variable SKL_SEED gives a seed.
 *
 * Precision: FP64 throughout, as in the reference, unless the environment
 * variable SKL_PRECISION is "mixed" (FP32 sketch and QR, FP64 LSQR; see
 * SKL_MIXED in skl.h).
 */
#ifndef SKILLS_H
#define SKILLS_H

#include <cassert>
#include <cstddef>
#include <memory>
#include <ostream>
#include <stdexcept>

#include "skl.h"

/* defaults of include/bench_config.hpp:7-20 */
#ifndef IT_TOL
#define IT_TOL 1e-6
#endif
#ifndef MAX_IT
#define MAX_IT 10000
#endif
#ifndef OVER_SAMPLING_RATIO
#define OVER_SAMPLING_RATIO 1.1
#endif
#ifndef NNZ_PER_COLUMN
#define NNZ_PER_COLUMN 3
#endif
#ifndef PERTURB
#define PERTURB 1e-10
#endif
#ifndef RCOND_THRESHOLD
#define RCOND_THRESHOLD 1e-10
#endif
#ifndef WISDOM
#define WISDOM 0
#endif
#ifndef ABS_TOL
#define ABS_TOL 1e-8
#endif
#ifndef SPQR_ORDERING
#define SPQR_ORDERING 2
#endif

#define FFTW_R2R_DCT 1
#define FFTW_R2R_DHT 2

typedef long INT;

template <typename T> class Vec {
  public:
    typedef std::ptrdiff_t idx_t;
    Vec() : n_(0), inc_(1), data_(nullptr) {}
    explicit Vec(idx_t n) : n_(n), inc_(1), own_(new T[n > 0 ? n : 1]()) { data_ = own_.get(); }
    Vec(idx_t n, T *data, idx_t inc = 1) : n_(n), inc_(inc), data_(data) {}
    const idx_t &n() const { return n_; }
    const idx_t &inc() const { return inc_; }
    T *data() const { return data_; }
    T &operator()(idx_t i) const { return data_[i * inc_]; }
    Vec copy() const {
        Vec v(n_);
        for (idx_t i = 0; i < n_; ++i) v(i) = (*this)(i);
        return v;
    }
    Vec &operator=(const T &s) {
        for (idx_t i = 0; i < n_; ++i) (*this)(i) = s;
        return *this;
    }

  private:
    idx_t n_, inc_;
    std::shared_ptr<T> own_;
    T *data_;
};

template <typename T> class Mat {
  public:
    typedef std::ptrdiff_t idx_t;
    Mat() : m_(0), n_(0), ld_(0), data_(nullptr) {}
    Mat(idx_t m, idx_t n) : m_(m), n_(n), ld_(m), own_(new T[(m * n) > 0 ? m * n : 1]()) {
        data_ = own_.get();
    }
    Mat(idx_t m, idx_t n, T *data) : m_(m), n_(n), ld_(m), data_(data) {}
    Mat(idx_t m, idx_t n, T *data, idx_t ld) : m_(m), n_(n), ld_(ld), data_(data) {}
    const idx_t &m() const { return m_; }
    const idx_t &n() const { return n_; }
    const idx_t &ld() const { return ld_; }
    T *data() const { return data_; }
    T &operator()(idx_t i, idx_t j) const { return data_[i + j * ld_]; }
    Mat copy() const {
        Mat c(m_, n_);
        for (idx_t j = 0; j < n_; ++j)
            for (idx_t i = 0; i < m_; ++i) c(i, j) = (*this)(i, j);
        return c;
    }
    Mat submat(idx_t i0, idx_t i1, idx_t j0, idx_t j1) const {
        return Mat(i1 - i0, j1 - j0, data_ + i0 + j0 * ld_, ld_);
    }
    Vec<T> col(idx_t j) const { return Vec<T>(m_, data_ + j * ld_, 1); }
    /* y = alpha op(A) x + beta y  (Mat.hpp:184-187) */
    void mv(const char trans, const T alpha, const Vec<T> &x, const T beta, Vec<T> &y) const {
        if (trans == 'n' || trans == 'N') {
            assert(x.n() == n_ && y.n() == m_);
            for (idx_t i = 0; i < m_; ++i) y(i) = (beta == T(0)) ? T(0) : beta * y(i);
            for (idx_t j = 0; j < n_; ++j) {
                T t = alpha * x(j);
                for (idx_t i = 0; i < m_; ++i) y(i) += t * (*this)(i, j);
            }
        } else {
            assert(x.n() == m_ && y.n() == n_);
            for (idx_t j = 0; j < n_; ++j) {
                T t = T(0);
                for (idx_t i = 0; i < m_; ++i) t += (*this)(i, j) * x(i);
                y(j) = ((beta == T(0)) ? T(0) : beta * y(j)) + alpha * t;
            }
        }
    }

  private:
    idx_t m_, n_, ld_;
    std::shared_ptr<T> own_;
    T *data_;
};

typedef Vec<double> Vec_d;
typedef Mat<double> Mat_d;

template <typename T> std::ostream &operator<<(std::ostream &os, const Vec<T> &v) {
    for (std::ptrdiff_t i = 0; i < v.n(); ++i) os << v(i) << (i + 1 < v.n() ? " " : "");
    return os;
}
template <typename T> std::ostream &operator<<(std::ostream &os, const Mat<T> &A) {
    for (std::ptrdiff_t i = 0; i < A.m(); ++i) {
        for (std::ptrdiff_t j = 0; j < A.n(); ++j) os << A(i, j) << (j + 1 < A.n() ? " " : "");
        os << "\n";
    }
    return os;
}

#ifndef _CS_H
/* CXSparse cs_dl layout */
typedef struct cs_dl_sparse {
    long nzmax;
    long m;
    long n;
    long *p;
    long *i;
    double *x;
    long nz;
} cs_dl;
#endif

/* From blendenpik.hpp */
void ls_dense_hashing_blendenpik(Mat_d &A, Vec_d &b, Vec_d &x, long &rank, int &flag, long &it,
                                 double gamma, long k, double abs_tol, double rcond,
                                 double it_tol, long max_it, int debug, int wisdom);

void ls_dense_hashing_blendenpik_noCPQR(Mat_d &A, Vec_d &b, Vec_d &x, long &rank, int &flag,
                                        long &it, double gamma, long k, double it_tol,
                                        long max_it, int debug, int wisdom);

/* From blendenpik.hpp:31-33 (the reference declares it there, not in
 * ski-lls.h; include/blendenpik.hpp of this repository includes this file) */
void ls_dense_blendenpik(Mat_d &A, Vec_d &b, Vec_d &x, long &rank, int &flag, long &it,
                         double gamma, double it_tol, long max_it, int debug, int wisdom);

/* From lsex_all_in_c.h */
void ls_sparse_spqr(cs_dl &A, Vec_d &b, Vec_d &x, long &rank, int &flag, long &it, double gamma,
                    long k, double abs_tol = ABS_TOL, int ordering = SPQR_ORDERING,
                    double it_tol = IT_TOL, long max_it = MAX_IT,
                    double rcond_threshold = RCOND_THRESHOLD, double perturb = PERTURB,
                    int debug = 0);

#endif /* SKILLS_H */
