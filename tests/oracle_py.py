"""ctypes binding of the CPU oracle (oracle/libskl_oracle.so) for the tests.

The oracle is the checker: tests call it to produce expected values and
compare the HIP library against them.  Nothing in the product imports this.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "libskl_oracle.so")

_lib = None

dp = C.POINTER(C.c_double)
lp = C.POINTER(C.c_long)


class Rng(C.Structure):
    _fields_ = [("state", C.c_int32 * 31), ("f", C.c_int), ("r", C.c_int),
                ("gauss_toggle", C.c_int), ("gauss_b2", C.c_double)]


class LsqrStep(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_double), ("rhobar", C.c_double),
                ("phibar", C.c_double), ("nrm_a", C.c_double)]


class Info(C.Structure):
    _fields_ = [("rank", C.c_long), ("flag", C.c_int), ("it", C.c_long),
                ("success", C.c_int), ("rcond", C.c_double), ("x0_residual", C.c_double),
                ("t_sketch", C.c_double), ("t_qr", C.c_double), ("t_x0", C.c_double),
                ("t_lsqr", C.c_double), ("t_total", C.c_double),
                ("lsqr_nrm_ar", C.c_double), ("lsqr_nrm_a", C.c_double),
                ("lsqr_nrm_r", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run make -C oracle")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def load_variant(path):
    """Binds an oracle library built elsewhere (a planted-mistake variant of
    skl_oracle.c); returns a function hqrrp(A, nb, pp) over it."""
    L = C.CDLL(path)

    def hqrrp_v(A, nb=64, pp=10):
        A = fcopy(A)
        m, n = A.shape
        jpvt = np.zeros(n, dtype=np.int64)
        tau = np.zeros(n)
        L.orc_hqrrp(C.c_long(m), C.c_long(n), _d(A), C.c_long(m), _l(jpvt), _d(tau),
                    C.c_long(nb), C.c_long(pp))
        return A, jpvt, tau
    return hqrrp_v


def _d(a):
    return a.ctypes.data_as(dp)


def _l(a):
    return a.ctypes.data_as(lp)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def fortran(a):
    return np.asfortranarray(a, dtype=np.float64)


def fcopy(a):
    return np.array(a, dtype=np.float64, order="F", copy=True)


# ---- RNG -------------------------------------------------------------------
def rng(seed):
    g = Rng()
    lib().orc_srand(C.byref(g), C.c_uint(seed))
    return g


def rand_seq(seed, count):
    g = rng(seed)
    f = lib().orc_rand
    return np.array([f(C.byref(g)) for _ in range(count)], dtype=np.int64)


def gen_signs(g, n):
    D = np.empty(n)
    lib().orc_gen_signs(C.byref(g), C.c_long(n), C.c_double(1.0), _d(D))
    return D


def gen_hashing(g, m, n, k):
    rows = np.empty(n * k, dtype=np.int64)
    colptr = np.empty(n + 1, dtype=np.int64)
    vals = np.empty(n * k)
    lib().orc_gen_hashing(C.byref(g), C.c_long(m), C.c_long(n), C.c_long(k), _l(rows),
                          _l(colptr), _d(vals))
    return rows, colptr, vals


def gaussian_matrix(g, rows, cols):
    G = np.empty((rows, cols), order="F")
    lib().orc_gaussian_matrix(C.byref(g), C.c_long(rows), C.c_long(cols), _d(G),
                              C.c_long(rows))
    return G


def dense_pattern(seed, n, m, k):
    """Signs D and hashing pattern in the dense solver's draw order
    (dense_helpers.cpp:92,12-14 then sparse_helpers.cpp:90-116)."""
    g = rng(seed)
    D = gen_signs(g, n)
    rows, _, vals = gen_hashing(g, m, n, k)
    return D, rows, vals


def select_k_from_n(g, n, k):
    idx = np.empty(k, dtype=np.int64)
    lib().orc_select_k_from_n(C.byref(g), C.c_long(n), C.c_long(k), _l(idx))
    return idx


def subsample_pattern(seed, n, m):
    """Signs D and sampled rows of ls_dense_blendenpik (dense_helpers.cpp:7-59):
    both drawn from the stream seeded with `seed` (it is seeded twice)."""
    D = gen_signs(rng(seed), n)
    return D, select_k_from_n(rng(seed), n, m)


def sparse_pattern(seed, n, m, k):
    g = rng(seed)
    rows, _, vals = gen_hashing(g, m, n, k)
    return rows, vals


# ---- transforms --------------------------------------------------------------
def fft(z):
    z = np.ascontiguousarray(z, dtype=np.complex128).copy()
    lib().orc_fft(C.c_long(z.size), z.ctypes.data_as(dp))
    return z


def dht_cols(X):
    Y = fcopy(X)
    n, c = Y.shape
    lib().orc_dht_cols(C.c_long(n), C.c_long(c), _d(Y), C.c_long(n))
    return Y


def fut(A, D):
    A = fortran(A)
    n, d = A.shape
    Y = np.empty((n, d), order="F")
    lib().orc_fut(C.c_long(n), C.c_long(d), _d(A), C.c_long(n), _d(f64(D)), _d(Y),
                  C.c_long(n))
    return Y


# ---- sketch ------------------------------------------------------------------
def hash_apply(m, rows, vals, k, alpha, X):
    X = fortran(X)
    n, c = X.shape
    Y = np.empty((m, c), order="F")
    lib().orc_hash_apply(C.c_long(m), C.c_long(n), C.c_long(k), _l(rows), _d(vals),
                         C.c_double(alpha), C.c_long(c), _d(X), C.c_long(n), _d(Y),
                         C.c_long(m))
    return Y


def sketch_sparse(m, rows, vals, k, Ap, Ai, Ax, n, d):
    SA = np.empty((m, d), order="F")
    lib().orc_sketch_sparse(C.c_long(m), C.c_long(n), C.c_long(k), _l(rows), _d(vals),
                            C.c_long(d), _l(Ap), _l(Ai), _d(Ax), _d(SA), C.c_long(m))
    return SA


def sketch_vec(m, rows, vals, k, b):
    y = np.empty(m)
    lib().orc_sketch_vec(C.c_long(m), C.c_long(b.size), C.c_long(k), _l(rows), _d(vals),
                         _d(f64(b)), _d(y))
    return y


# ---- QR ----------------------------------------------------------------------
def geqrf(A):
    A = fcopy(A)
    m, n = A.shape
    tau = np.zeros(min(m, n))
    lib().orc_geqrf(C.c_long(m), C.c_long(n), _d(A), C.c_long(m), _d(tau))
    return A, tau


def geqp3(A):
    A = fcopy(A)
    m, n = A.shape
    jpvt = np.zeros(n, dtype=np.int64)
    tau = np.zeros(n)
    lib().orc_geqp3(C.c_long(m), C.c_long(n), _d(A), C.c_long(m), _l(jpvt), _d(tau))
    return A, jpvt, tau


def hqrrp(A, nb=64, pp=10):
    A = fcopy(A)
    m, n = A.shape
    jpvt = np.zeros(n, dtype=np.int64)
    tau = np.zeros(n)
    lib().orc_hqrrp(C.c_long(m), C.c_long(n), _d(A), C.c_long(m), _l(jpvt), _d(tau),
                    C.c_long(nb), C.c_long(pp))
    return A, jpvt, tau


def cpqr(A):
    A = fcopy(A)
    m, n = A.shape
    jpvt = np.zeros(n, dtype=np.int64)
    tau = np.zeros(n)
    lib().orc_cpqr(C.c_long(m), C.c_long(n), _d(A), C.c_long(m), _l(jpvt), _d(tau))
    return A, jpvt, tau


def ormqr_lt(QR, tau, c):
    QR = fortran(QR)
    m = QR.shape[0]
    c = f64(c).copy()
    lib().orc_ormqr_lt(C.c_long(m), C.c_long(tau.size), _d(QR), C.c_long(m), _d(f64(tau)),
                       _d(c))
    return c


def trsv_un(R, x, perturb=0.0):
    R = fortran(R)
    x = f64(x).copy()
    lib().orc_trsv_un(C.c_long(x.size), _d(R), C.c_long(R.shape[0]), _d(x),
                      C.c_double(perturb))
    return x


def trsv_ut(R, x, perturb=0.0):
    R = fortran(R)
    x = f64(x).copy()
    lib().orc_trsv_ut(C.c_long(x.size), _d(R), C.c_long(R.shape[0]), _d(x),
                      C.c_double(perturb))
    return x


def rcond_inf(R):
    R = fortran(R)
    f = lib().orc_rcond_inf
    f.restype = C.c_double
    return f(C.c_long(R.shape[1]), _d(R), C.c_long(R.shape[0]))


def sparse_precond(SA, bs):
    SA = fcopy(SA)
    m, d = SA.shape
    bs = f64(bs).copy()
    E = np.zeros(d, dtype=np.int64)
    tol = C.c_double(0)
    f = lib().orc_sparse_precond
    f.restype = C.c_long
    rank = f(C.c_long(m), C.c_long(d), _d(SA), C.c_long(m), _d(bs), _l(E), C.byref(tol))
    return SA, bs, E, rank, tol.value


# ---- LSQR --------------------------------------------------------------------
def lsqr_dense(A, R, b, tol, max_it, E=None, hist=False):
    A = fortran(A)
    R = fortran(R)
    n, _ = A.shape
    r = R.shape[1]
    y = np.zeros(r)
    flag = C.c_int(0)
    it = C.c_long(0)
    H = (LsqrStep * max(max_it, 1))() if hist else None
    info = Info()
    Ep = _l(np.ascontiguousarray(E, dtype=np.int64)) if E is not None else None
    lib().orc_lsqr_dense(C.c_long(n), C.c_long(r), _d(A), C.c_long(n), Ep, _d(R),
                         C.c_long(R.shape[0]), _d(f64(b)), C.c_double(tol),
                         C.c_long(max_it), _d(y), C.byref(flag), C.byref(it), H,
                         C.byref(info))
    steps = None
    if hist:
        steps = np.array([[H[i].alpha, H[i].beta, H[i].rhobar, H[i].phibar, H[i].nrm_a]
                          for i in range(it.value)])
    return y, flag.value, it.value, steps, info


# ---- solvers -----------------------------------------------------------------
def lsqr_sparse(Acsc, R, b, tol, max_it, E=None, perturb=0.0):
    """LSQR on A(:, E) (R + perturb I)^{-1} with A in CSC (scipy)."""
    n, d = Acsc.shape
    R = fortran(R)
    r = R.shape[1]
    y = np.zeros(r)
    flag = C.c_int(0)
    it = C.c_long(0)
    info = Info()
    Ep = _l(np.ascontiguousarray(E, dtype=np.int64)) if E is not None else None
    Ap = np.ascontiguousarray(Acsc.indptr, dtype=np.int64)
    Ai = np.ascontiguousarray(Acsc.indices, dtype=np.int64)
    lib().orc_lsqr_sparse(C.c_long(n), C.c_long(d), _l(Ap), _l(Ai), _d(f64(Acsc.data)),
                          C.c_long(r), Ep, _d(R), C.c_long(R.shape[0]), C.c_double(perturb),
                          _d(f64(b)), C.c_double(tol), C.c_long(max_it), _d(y), C.byref(flag),
                          C.byref(it), None, C.byref(info))
    return y, flag.value, it.value, info


def _hist(max_it, want):
    return (LsqrStep * max(max_it, 1))() if want else None


def _steps(H, it):
    if H is None:
        return None
    return np.array([[H[i].alpha, H[i].beta, H[i].rhobar, H[i].phibar, H[i].nrm_a]
                     for i in range(it)])


def ls_dense(A, b, gamma=1.7, k=1, abs_tol=1e-8, rcond=1e-12, it_tol=1e-10,
             max_it=10000, seed=1, hist=False):
    A = fcopy(A)
    n, d = A.shape
    x = np.zeros(d)
    info = Info()
    H = _hist(max_it, hist)
    rc = lib().orc_ls_dense(C.c_long(n), C.c_long(d), _d(A), C.c_long(n), _d(f64(b)), _d(x),
                            C.c_double(gamma), C.c_long(k), C.c_double(abs_tol),
                            C.c_double(rcond), C.c_double(it_tol), C.c_long(max_it),
                            C.c_uint(seed), H, C.byref(info))
    if rc != 0:
        raise RuntimeError("Matrix is not overdetermined." if rc == 1 else f"oracle error {rc}")
    return x, info, _steps(H, info.it)


def ls_dense_nocpqr(A, b, gamma=1.7, k=1, it_tol=1e-10, max_it=10000, seed=1, hist=False):
    A = fortran(A)
    n, d = A.shape
    x = np.zeros(d)
    info = Info()
    H = _hist(max_it, hist)
    rc = lib().orc_ls_dense_nocpqr(C.c_long(n), C.c_long(d), _d(A), C.c_long(n),
                                   _d(f64(b)), _d(x), C.c_double(gamma), C.c_long(k),
                                   C.c_double(it_tol), C.c_long(max_it), C.c_uint(seed), H,
                                   C.byref(info))
    if rc != 0:
        raise RuntimeError("Matrix is not overdetermined." if rc == 1 else f"oracle error {rc}")
    return x, info, _steps(H, info.it)


def ls_dense_blendenpik(A, b, gamma=1.7, it_tol=1e-10, max_it=10000, seed=1, hist=False):
    A = fortran(A)
    n, d = A.shape
    x = np.zeros(d)
    info = Info()
    H = _hist(max_it, hist)
    rc = lib().orc_ls_dense_blendenpik(C.c_long(n), C.c_long(d), _d(A), C.c_long(n),
                                       _d(f64(b)), _d(x), C.c_double(gamma),
                                       C.c_double(it_tol), C.c_long(max_it), C.c_uint(seed), H,
                                       C.byref(info))
    if rc != 0:
        raise RuntimeError("Matrix is not overdetermined." if rc == 1 else f"oracle error {rc}")
    return x, info, _steps(H, info.it)


def ls_sparse(Acsc, b, gamma=1.5, k=3, abs_tol=1e-8, it_tol=1e-10, max_it=10000,
              rcond_threshold=1e-10, perturb=1e-10, seed=1, hist=False):
    """Acsc: scipy.sparse csc_matrix (n x d)."""
    n, d = Acsc.shape
    Ap = np.ascontiguousarray(Acsc.indptr, dtype=np.int64)
    Ai = np.ascontiguousarray(Acsc.indices, dtype=np.int64)
    Ax = f64(Acsc.data)
    x = np.zeros(d)
    info = Info()
    H = _hist(max_it, hist)
    rc = lib().orc_ls_sparse(C.c_long(n), C.c_long(d), _l(Ap), _l(Ai), _d(Ax), _d(f64(b)),
                             _d(x), C.c_double(gamma), C.c_long(k), C.c_double(abs_tol),
                             C.c_double(it_tol), C.c_long(max_it),
                             C.c_double(rcond_threshold), C.c_double(perturb),
                             C.c_uint(seed), H, C.byref(info))
    if rc == 1:
        raise RuntimeError("Matrix is not overdetermined.")
    if rc == 2:
        raise RuntimeError("SPQR Factorization fails")
    return x, info, _steps(H, info.it)
