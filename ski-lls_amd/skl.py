"""Python binding of libskl.so (the MI355X Ski-LLS library) over its C ABI.

Mirrors the reference's solver interface (include/ski-lls.h:15-28):
``ls_dense_hashing_blendenpik``, ``ls_dense_hashing_blendenpik_noCPQR`` and
``ls_sparse_spqr`` take the same arguments with the same meaning and raise
the same ``RuntimeError`` messages ("Matrix is not overdetermined.",
"SPQR Factorization fails").  The component functions expose each hot-path
stage for parity checks.  Loading fails loudly when the HIP library is not
built: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libskl.so")

SKL_FP64 = 0
SKL_MIXED = 1
SKL_CHOLQR64 = 2  # precision ablation: SKL_MIXED's algorithm with an FP64 Gram

# include/bench_config.hpp:7-20
IT_TOL = 1e-6
MAX_IT = 10000
OVER_SAMPLING_RATIO = 1.1
NNZ_PER_COLUMN = 3
PERTURB = 1e-10
RCOND_THRESHOLD = 1e-10
ABS_TOL = 1e-8
SPQR_ORDERING = 2

dp = C.POINTER(C.c_double)
lp = C.POINTER(C.c_long)


class Opts(C.Structure):
    _fields_ = [("gamma", C.c_double), ("k", C.c_long), ("abs_tol", C.c_double),
                ("rcond", C.c_double), ("it_tol", C.c_double), ("max_it", C.c_long),
                ("rcond_threshold", C.c_double), ("perturb", C.c_double),
                ("ordering", C.c_int), ("debug", C.c_int), ("wisdom", C.c_int),
                ("seed", C.c_uint), ("precision", C.c_int), ("escalate_cond", C.c_double),
                ("use_graph", C.c_int), ("graph_chunk", C.c_int), ("reserved", C.c_int * 6)]


class Info(C.Structure):
    _fields_ = [("rank", C.c_long), ("flag", C.c_int), ("it", C.c_long), ("success", C.c_int),
                ("escalated", C.c_int), ("rcond", C.c_double), ("cond_est", C.c_double),
                ("x0_residual", C.c_double), ("lsqr_nrm_ar", C.c_double),
                ("lsqr_nrm_a", C.c_double), ("lsqr_nrm_r", C.c_double),
                ("t_pattern", C.c_double), ("t_sketch", C.c_double), ("t_qr", C.c_double),
                ("t_x0", C.c_double), ("t_lsqr", C.c_double), ("t_total", C.c_double),
                ("m", C.c_long), ("precond_lmin", C.c_double), ("precond_lmax", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LsqrStep(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_double), ("rhobar", C.c_double),
                ("phibar", C.c_double), ("nrm_a", C.c_double)]


# every symbol include/skl.h declares
EXPORTS = [
    "skl_default_opts", "skl_init", "skl_finalize", "skl_error_string", "skl_last_error",
    "skl_device_count", "skl_dense_solve", "skl_dense_solve_nocpqr", "skl_sparse_solve",
    "skl_dense_solve_dev", "skl_dense_solve_nocpqr_dev", "skl_sparse_solve_dev",
    "skl_dense_solve_blendenpik", "skl_dense_solve_blendenpik_dev", "skl_subsample_pattern",
    "skl_set_history", "skl_hash_pattern", "skl_dht", "skl_hash_apply", "skl_dense_sketch",
    "skl_sparse_sketch", "skl_qr", "skl_cpqr", "skl_trinv", "skl_lsqr_dense",
    "skl_profile_enable", "skl_profile_reset", "skl_profile_read", "skl_set_barrier_spins",
    "skl_dense_solve_dev_noalloc",
    "skl_upload_dense", "skl_upload_dense_dev", "skl_upload_csc", "skl_handle_query",
    "skl_solve_handle", "skl_solve_handle_dev", "skl_release",
    "skl_autotune_lsqr", "skl_autotune_entries", "skl_autotune_clear", "skl_server_set_autotune",
    "skl_server_create", "skl_server_submit_dense", "skl_server_submit_sparse",
    "skl_server_submit_dense_dev", "skl_server_submit_sparse_dev", "skl_server_submit_handle",
    "skl_server_wait", "skl_server_pending", "skl_server_destroy",
]

# solver of a handle solve (skl.h SKL_METHOD_*)
METHOD_DENSE_CPQR, METHOD_DENSE_NOCPQR, METHOD_DENSE_BLENDENPIK, METHOD_SPARSE_SPQR = 0, 1, 2, 3

_lib = None


class SklError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        _lib.skl_error_string.restype = C.c_char_p
        _lib.skl_last_error.restype = C.c_char_p
    return _lib


def _check(rc):
    if rc == 0:
        return
    if rc == 1:
        raise SklError(rc, "Matrix is not overdetermined.")
    if rc == 2:
        raise SklError(rc, "SPQR Factorization fails")
    L = lib()
    raise SklError(rc, f"{L.skl_error_string(rc).decode()}: {L.skl_last_error().decode()}")


def set_barrier_spins(spins):
    """Poll limit of the factorization kernels' grid barriers on this thread;
    negative restores the default.  A barrier that gives up makes the call
    that launched it fail with SklError (code 4)."""
    lib().skl_set_barrier_spins(C.c_long(int(spins)))


def default_opts(**kw):
    o = Opts()
    lib().skl_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _d(a):
    return a.ctypes.data_as(dp)


def _l(a):
    return a.ctypes.data_as(lp)


def _f(a):
    return np.asfortranarray(a, dtype=np.float64)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- solvers ------------------------------------------------------------------
def dense_solve(A, b, cpqr=True, history=0, sampling="hashing", **opts):
    """x, info[, steps] for min ||A x - b|| (host arrays).  sampling="uniform"
    runs ls_dense_blendenpik (row sampling of the transform, no CPQR)."""
    A = _f(A)
    b = _c(b)
    n, d = A.shape
    x = np.zeros(d)
    o = default_opts(**opts)
    info = Info()
    H = None
    if history:
        H = (LsqrStep * history)()
        lib().skl_set_history(H, C.c_long(history))
    if sampling == "uniform":
        fn = lib().skl_dense_solve_blendenpik
    else:
        fn = lib().skl_dense_solve if cpqr else lib().skl_dense_solve_nocpqr
    try:
        rc = fn(C.c_long(n), C.c_long(d), _d(A), C.c_long(n), _d(b), _d(x), C.byref(o),
                C.byref(info))
    finally:
        if history:
            lib().skl_set_history(None, C.c_long(0))
    _check(rc)
    if history:
        steps = np.array([[H[i].alpha, H[i].beta, H[i].rhobar, H[i].phibar, H[i].nrm_a]
                          for i in range(min(info.it, history))])
        return x, info, steps
    return x, info


def sparse_solve(Acsc, b, history=0, **opts):
    n, d = Acsc.shape
    Ap = np.ascontiguousarray(Acsc.indptr, dtype=np.int64)
    Ai = np.ascontiguousarray(Acsc.indices, dtype=np.int64)
    Ax = _c(Acsc.data)
    b = _c(b)
    x = np.zeros(d)
    o = default_opts(**opts)
    info = Info()
    H = None
    if history:
        H = (LsqrStep * history)()
        lib().skl_set_history(H, C.c_long(history))
    try:
        rc = lib().skl_sparse_solve(C.c_long(n), C.c_long(d), C.c_long(Ap[-1]), _l(Ap), _l(Ai),
                                    _d(Ax), _d(b), _d(x), C.byref(o), C.byref(info))
    finally:
        if history:
            lib().skl_set_history(None, C.c_long(0))
    _check(rc)
    if history:
        steps = np.array([[H[i].alpha, H[i].beta, H[i].rhobar, H[i].phibar, H[i].nrm_a]
                          for i in range(min(info.it, history))])
        return x, info, steps
    return x, info


def dense_solve_dev(n, d, A_ptr, lda, b_ptr, x_ptr, cpqr=True, sampling="hashing", **opts):
    """Device-resident inputs (raw device pointers, e.g. torch tensor data_ptr())."""
    o = default_opts(**opts)
    info = Info()
    if sampling == "uniform":
        fn = lib().skl_dense_solve_blendenpik_dev
    else:
        fn = lib().skl_dense_solve_dev if cpqr else lib().skl_dense_solve_nocpqr_dev
    rc = fn(C.c_long(n), C.c_long(d), C.c_void_p(A_ptr), C.c_long(lda), C.c_void_p(b_ptr),
            C.c_void_p(x_ptr), C.byref(o), C.byref(info))
    _check(rc)
    return info


def dense_solve_dev_noalloc(n, d, A_ptr, lda, b_ptr, x_ptr, **opts):
    """skl_dense_solve_dev_noalloc: the robust dense FP64 solver on device
    pointers with no device allocation (scratch from the library's static
    array; skl_init must have run on this thread)."""
    o = default_opts(**opts)
    info = Info()
    rc = lib().skl_dense_solve_dev_noalloc(C.c_long(n), C.c_long(d), C.c_void_p(A_ptr),
                                           C.c_long(lda), C.c_void_p(b_ptr), C.c_void_p(x_ptr),
                                           C.byref(o), C.byref(info))
    _check(rc)
    return info


def sparse_solve_dev(n, d, nnz, cp_ptr, ri_ptr, val_ptr, b_ptr, x_ptr, **opts):
    o = default_opts(**opts)
    info = Info()
    rc = lib().skl_sparse_solve_dev(C.c_long(n), C.c_long(d), C.c_long(nnz), C.c_void_p(cp_ptr),
                                    C.c_void_p(ri_ptr), C.c_void_p(val_ptr), C.c_void_p(b_ptr),
                                    C.c_void_p(x_ptr), C.byref(o), C.byref(info))
    _check(rc)
    return info


# ---- reference-interface mirror (include/ski-lls.h:15-28) ----------------------
def ls_dense_hashing_blendenpik(A, b, gamma, k, abs_tol, rcond, it_tol, max_it, debug=0,
                                wisdom=0, seed=None):
    """Returns (x, rank, flag, it) like the reference's output arguments."""
    if seed is None:
        import time
        seed = int(time.time()) & 0xFFFFFFFF
    x, info = dense_solve(A, b, cpqr=True, gamma=gamma, k=k, abs_tol=abs_tol, rcond=rcond,
                          it_tol=it_tol, max_it=max_it, debug=debug, wisdom=wisdom, seed=seed)
    return x, info.rank, info.flag, info.it


def ls_dense_hashing_blendenpik_noCPQR(A, b, gamma, k, it_tol, max_it, debug=0, wisdom=0,
                                       seed=None):
    if seed is None:
        import time
        seed = int(time.time()) & 0xFFFFFFFF
    x, info = dense_solve(A, b, cpqr=False, gamma=gamma, k=k, it_tol=it_tol, max_it=max_it,
                          debug=debug, wisdom=wisdom, seed=seed)
    return x, info.rank, info.flag, info.it


def ls_dense_blendenpik(A, b, gamma, it_tol, max_it, debug=0, wisdom=0, seed=None):
    """include/blendenpik.hpp:31-33: returns (x, rank, flag, it)."""
    if seed is None:
        import time
        seed = int(time.time()) & 0xFFFFFFFF
    x, info = dense_solve(A, b, cpqr=False, sampling="uniform", gamma=gamma, it_tol=it_tol,
                          max_it=max_it, debug=debug, wisdom=wisdom, seed=seed)
    return x, info.rank, info.flag, info.it


def ls_sparse_spqr(A, b, gamma, k, abs_tol=ABS_TOL, ordering=SPQR_ORDERING, it_tol=IT_TOL,
                   max_it=MAX_IT, rcond_threshold=RCOND_THRESHOLD, perturb=PERTURB, debug=0,
                   seed=1):
    x, info = sparse_solve(A, b, gamma=gamma, k=k, abs_tol=abs_tol, ordering=ordering,
                           it_tol=it_tol, max_it=max_it, rcond_threshold=rcond_threshold,
                           perturb=perturb, debug=debug, seed=seed)
    return x, info.rank, info.flag, info.it


# ---- LSQR launch-shape autotuner ---------------------------------------------------
def autotune_lsqr(n, d, nnz=None):
    """Tune the LSQR launch shape of the (n, d[, nnz]) bucket (CSC when nnz is
    given) -> (ms per iteration chosen, ms per iteration of the default)."""
    best, dflt = C.c_double(0), C.c_double(0)
    _check(lib().skl_autotune_lsqr(C.c_int(0 if nnz is None else 1), C.c_long(n), C.c_long(d),
                                   C.c_long(0 if nnz is None else nnz), C.byref(best),
                                   C.byref(dflt)))
    return best.value, dflt.value


def autotune_entries():
    lib().skl_autotune_entries.restype = C.c_long
    return lib().skl_autotune_entries()


def autotune_clear():
    lib().skl_autotune_clear()


# ---- device-resident problems -----------------------------------------------------
class Handle:
    """A problem kept in device memory (skl_upload_dense / skl_upload_csc):
    A is copied once; solve() takes b and returns x like dense_solve /
    sparse_solve on the same A.  Use as a context manager or call release()."""

    def __init__(self, A=None, Acsc=None, dev=None):
        L = lib()
        h = C.c_void_p()
        if Acsc is not None:
            n, d = Acsc.shape
            Ap = np.ascontiguousarray(Acsc.indptr, dtype=np.int64)
            Ai = np.ascontiguousarray(Acsc.indices, dtype=np.int64)
            Ax = _c(Acsc.data)
            rc = L.skl_upload_csc(C.c_long(n), C.c_long(d), C.c_long(Ap[-1]), _l(Ap), _l(Ai),
                                  _d(Ax), C.byref(h))
            self.sparse = True
        elif dev is not None:
            n, d, ptr, lda = dev
            rc = L.skl_upload_dense_dev(C.c_long(n), C.c_long(d), C.c_void_p(ptr), C.c_long(lda),
                                        C.byref(h))
            self.sparse = False
        else:
            A = _f(A)
            n, d = A.shape
            rc = L.skl_upload_dense(C.c_long(n), C.c_long(d), _d(A), C.c_long(n), C.byref(h))
            self.sparse = False
        _check(rc)
        self._h = h
        self.n, self.d = n, d

    def query(self):
        """-> (n, d, nnz or -1, device bytes held)"""
        n, d, nnz, by = C.c_long(), C.c_long(), C.c_long(), C.c_size_t()
        _check(lib().skl_handle_query(self._h, C.byref(n), C.byref(d), C.byref(nnz), C.byref(by)))
        return n.value, d.value, nnz.value, by.value

    def default_method(self, cpqr=True):
        if self.sparse:
            return METHOD_SPARSE_SPQR
        return METHOD_DENSE_CPQR if cpqr else METHOD_DENSE_NOCPQR

    def solve(self, b, method=None, **opts):
        b = _c(b)
        x = np.zeros(self.d)
        o = default_opts(**opts)
        info = Info()
        m = self.default_method() if method is None else method
        _check(lib().skl_solve_handle(self._h, C.c_int(m), _d(b), _d(x), C.byref(o),
                                      C.byref(info)))
        return x, info

    def solve_dev(self, b_ptr, x_ptr, method=None, **opts):
        o = default_opts(**opts)
        info = Info()
        m = self.default_method() if method is None else method
        _check(lib().skl_solve_handle_dev(self._h, C.c_int(m), C.c_void_p(b_ptr), C.c_void_p(x_ptr),
                                          C.byref(o), C.byref(info)))
        return info

    def release(self):
        if self._h:
            _check(lib().skl_release(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.release()


# ---- serving ----------------------------------------------------------------------
class Server:
    """Continuous-batching request loop (skl_server_*).  Submitted arrays are
    kept alive by the Server until their ticket is waited on."""

    def __init__(self, slots=4, mem_budget_bytes=0, device=0):
        L = lib()
        L.skl_server_create.restype = C.c_void_p
        L.skl_server_submit_dense.restype = C.c_long
        L.skl_server_submit_sparse.restype = C.c_long
        L.skl_server_submit_handle.restype = C.c_long
        L.skl_server_submit_dense_dev.restype = C.c_long
        L.skl_server_submit_sparse_dev.restype = C.c_long
        self._h = C.c_void_p(L.skl_server_create(C.c_int(slots), C.c_size_t(mem_budget_bytes),
                                                 C.c_int(device)))
        self._keep = {}

    def submit_dense(self, A, b, cpqr=True, **opts):
        A = _f(A)
        b = _c(b)
        n, d = A.shape
        x = np.zeros(d)
        o = default_opts(**opts)
        t = lib().skl_server_submit_dense(self._h, C.c_int(1 if cpqr else 0), C.c_long(n),
                                          C.c_long(d), _d(A), C.c_long(n), _d(b), _d(x),
                                          C.byref(o))
        if t < 0:
            _check(-t)
        self._keep[t] = (A, b, x, o)
        return t

    def submit_sparse(self, Acsc, b, **opts):
        n, d = Acsc.shape
        Ap = np.ascontiguousarray(Acsc.indptr, dtype=np.int64)
        Ai = np.ascontiguousarray(Acsc.indices, dtype=np.int64)
        Ax = _c(Acsc.data)
        b = _c(b)
        x = np.zeros(d)
        o = default_opts(**opts)
        t = lib().skl_server_submit_sparse(self._h, C.c_long(n), C.c_long(d), C.c_long(Ap[-1]),
                                           _l(Ap), _l(Ai), _d(Ax), _d(b), _d(x), C.byref(o))
        if t < 0:
            _check(-t)
        self._keep[t] = (Ap, Ai, Ax, b, x, o)
        return t

    def submit_handle(self, handle, b, method=None, **opts):
        """A solve on a Handle (its A stays resident); the Handle must outlive
        the wait."""
        b = _c(b)
        x = np.zeros(handle.d)
        o = default_opts(**opts)
        m = handle.default_method() if method is None else method
        t = lib().skl_server_submit_handle(self._h, handle._h, C.c_int(m), _d(b), _d(x),
                                           C.byref(o))
        if t < 0:
            _check(-t)
        self._keep[t] = (handle, b, x, o)
        return t

    def submit_dense_dev(self, n, d, A_ptr, lda, b_ptr, x_ptr, cpqr=True, **opts):
        """Device pointers (on the server's device); wait() returns x = None."""
        o = default_opts(**opts)
        t = lib().skl_server_submit_dense_dev(self._h, C.c_int(1 if cpqr else 0), C.c_long(n),
                                              C.c_long(d), C.c_void_p(A_ptr), C.c_long(lda),
                                              C.c_void_p(b_ptr), C.c_void_p(x_ptr), C.byref(o))
        if t < 0:
            _check(-t)
        self._keep[t] = (o, None, o)
        return t

    def submit_sparse_dev(self, n, d, nnz, cp_ptr, ri_ptr, val_ptr, b_ptr, x_ptr, **opts):
        """Device pointers (int64 CSC arrays, on the server's device)."""
        o = default_opts(**opts)
        t = lib().skl_server_submit_sparse_dev(self._h, C.c_long(n), C.c_long(d), C.c_long(nnz),
                                               C.c_void_p(cp_ptr), C.c_void_p(ri_ptr),
                                               C.c_void_p(val_ptr), C.c_void_p(b_ptr),
                                               C.c_void_p(x_ptr), C.byref(o))
        if t < 0:
            _check(-t)
        self._keep[t] = (o, None, o)
        return t

    def wait(self, ticket):
        """-> (x, info, queue_seconds, run_seconds)"""
        info = Info()
        q = C.c_double(0)
        r = C.c_double(0)
        rc = lib().skl_server_wait(self._h, C.c_long(ticket), C.byref(info), C.byref(q),
                                   C.byref(r))
        keep = self._keep.pop(ticket)
        _check(rc)
        return keep[-2], info, q.value, r.value

    def pending(self):
        return lib().skl_server_pending(self._h)

    def set_autotune(self, on=True):
        """Tune each new LSQR shape bucket before its first request runs."""
        lib().skl_server_set_autotune(self._h, C.c_int(1 if on else 0))

    def close(self):
        if self._h:
            lib().skl_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ---- kernel timing --------------------------------------------------------------
PROFILE_SLOTS = {0: "A z (gemv_n / spmv_n; fused_rows: A z with A^T u)", 1: "A^T u (gemv_t / csc_t)", 2: "M v (trmv_n)",
                 3: "M^T t (trmv_t)"}


def profile(on=True, reset=True):
    if reset:
        lib().skl_profile_reset()
    lib().skl_profile_enable(C.c_int(1 if on else 0))


def profile_read():
    out = {}
    for s in range(4):
        ms = C.c_double(0)
        cnt = C.c_long(0)
        by = C.c_double(0)
        _check(lib().skl_profile_read(C.c_int(s), C.byref(ms), C.byref(cnt), C.byref(by)))
        out[s] = {"name": PROFILE_SLOTS[s], "avg_ms": ms.value, "launches": cnt.value,
                  "bytes": by.value}
    return out


# ---- components -------------------------------------------------------------
def hash_pattern(seed, m, n, k, dense):
    D = np.zeros(n) if dense else None
    rows = np.zeros(n * k, dtype=np.int64)
    vals = np.zeros(n * k)
    _check(lib().skl_hash_pattern(C.c_uint(seed), C.c_int(1 if dense else 0), C.c_long(m),
                                  C.c_long(n), C.c_long(k), _d(D) if D is not None else None,
                                  _l(rows), _d(vals)))
    return D, rows, vals


def subsample_pattern(seed, m, n):
    """Signs D and the m sampled rows of ls_dense_blendenpik for a seed."""
    D = np.zeros(n)
    idx = np.zeros(m, dtype=np.int64)
    _check(lib().skl_subsample_pattern(C.c_uint(seed), C.c_long(m), C.c_long(n), _d(D), _l(idx)))
    return D, idx


def dht(X, precision=SKL_FP64):
    X = _f(X)
    n, c = X.shape
    Y = np.zeros((n, c), order="F")
    _check(lib().skl_dht(C.c_long(n), C.c_long(c), _d(X), C.c_long(n), _d(Y), C.c_long(n),
                         C.c_int(precision)))
    return Y


def hash_apply(m, rows, vals, k, alpha, X, precision=SKL_FP64):
    X = _f(X)
    n, c = X.shape
    Y = np.zeros((m, c), order="F")
    _check(lib().skl_hash_apply(C.c_long(m), C.c_long(n), C.c_long(k),
                                _l(np.ascontiguousarray(rows, dtype=np.int64)), _d(_c(vals)),
                                C.c_double(alpha), C.c_long(c), _d(X), C.c_long(n), _d(Y),
                                C.c_long(m), C.c_int(precision)))
    return Y


def dense_sketch(A, b, m, k, D, rows, vals, precision=SKL_FP64):
    A = _f(A)
    n, d = A.shape
    nc = d + (1 if b is not None else 0)
    SA = np.zeros((m, nc), order="F")
    _check(lib().skl_dense_sketch(C.c_long(n), C.c_long(d), _d(A), C.c_long(n),
                                  _d(_c(b)) if b is not None else None, C.c_long(m), C.c_long(k),
                                  _d(_c(D)), _l(np.ascontiguousarray(rows, dtype=np.int64)),
                                  _d(_c(vals)), _d(SA), C.c_long(m), C.c_int(precision)))
    return SA


def sparse_sketch(m, rows, vals, k, Acsc, b=None):
    n, d = Acsc.shape
    Ap = np.ascontiguousarray(Acsc.indptr, dtype=np.int64)
    Ai = np.ascontiguousarray(Acsc.indices, dtype=np.int64)
    Ax = _c(Acsc.data)
    SA = np.zeros((m, d), order="F")
    bs = np.zeros(m) if b is not None else None
    _check(lib().skl_sparse_sketch(C.c_long(m), C.c_long(n), C.c_long(d), C.c_long(k),
                                   _l(np.ascontiguousarray(rows, dtype=np.int64)), _d(_c(vals)),
                                   _l(Ap), _l(Ai), _d(Ax), _d(_c(b)) if b is not None else None,
                                   _d(SA), C.c_long(m), _d(bs) if bs is not None else None))
    return SA, bs


def qr(A, nfac=None, precision=SKL_FP64):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    nfac = n if nfac is None else nfac
    tau = np.zeros(max(min(m, nfac), 1))
    _check(lib().skl_qr(C.c_long(m), C.c_long(n), C.c_long(nfac), _d(A), C.c_long(m), _d(tau),
                        C.c_int(precision)))
    return A, tau[:min(m, nfac)]


def cpqr(A, nfac=None, precision=SKL_FP64):
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    nfac = n if nfac is None else nfac
    tau = np.zeros(max(min(m, nfac), 1))
    jpvt = np.zeros(nfac, dtype=np.int64)
    _check(lib().skl_cpqr(C.c_long(m), C.c_long(n), C.c_long(nfac), _d(A), C.c_long(m), _l(jpvt),
                          _d(tau), C.c_int(precision)))
    return A, jpvt, tau[:min(m, nfac)]


def trinv(R, perturb=0.0):
    R = _f(R)
    r = R.shape[1]
    M = np.zeros((r, r), order="F")
    _check(lib().skl_trinv(C.c_long(r), _d(R), C.c_long(R.shape[0]), C.c_double(perturb), _d(M),
                           C.c_long(r)))
    return M


def lsqr_dense(A, M, b, tol, max_it, E=None, history=False):
    A = _f(A)
    M = _f(M)
    n, d = A.shape
    r = M.shape[1]
    y = np.zeros(r)
    flag = C.c_int(0)
    it = C.c_long(0)
    H = (LsqrStep * max(max_it, 1))() if history else None
    Ep = _l(np.ascontiguousarray(E, dtype=np.int64)) if E is not None else None
    _check(lib().skl_lsqr_dense(C.c_long(n), C.c_long(d), _d(A), C.c_long(n), C.c_long(r), Ep,
                                _d(M), C.c_long(r), _d(_c(b)), C.c_double(tol), C.c_long(max_it),
                                _d(y), C.byref(flag), C.byref(it), H))
    steps = None
    if history:
        steps = np.array([[H[i].alpha, H[i].beta, H[i].rhobar, H[i].phibar, H[i].nrm_a]
                          for i in range(it.value)])
    return y, flag.value, it.value, steps
