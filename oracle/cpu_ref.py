"""CPU restatement of the Ski-LLS solvers on numpy/scipy BLAS and LAPACK.

TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times beside the GPU
solver (its cpu_baseline leg), nothing in the product imports it.  The plain-C
oracle (skl_oracle.c) is the checker of record; this module restates the same
algorithm with the heavy operations on the host's BLAS/LAPACK so that the CPU
figure is a fair one, at one thread (the reference's configuration,
include/Config.hpp:7, config/make_intel.inc:4-6 mkl_sequential) and at all
cores:

* ls_dense_hashing_blendenpik (src/ls_dense_hashing_blendenpik.cpp:23-149):
  signs and hashing pattern in the reference's libc draw order (through the
  oracle's generator), the randomized DHT by real FFT (scipy.fft, the
  orthonormal cas transform of src/fast_unitary_transform.cpp:42-121), the
  hashing sketch as a sparse product (src/sketching.cpp:17-62), CPQR by LAPACK
  dgeqp3 (the reference's choice below 1000 columns,
  hqrrp/.../NoFLA_HQRRP_WY_blk_var4_skint.c:205-209; above it the reference
  runs HQRRP, which hqrrp/README.md:93-95 reports about 3x faster than
  dgeqp3 on one core), the sketch-and-solve start and Meng's LSQR with the
  recurrences and stopping test of src/IterSolver.cpp:160-284, the triangular
  solves by dtrsv (solve_triangular).
* ls_sparse_spqr (src/ls_sparse_spqr.cpp:79-214): the same sketch on CSC A
  (scipy sparse products), Householder QR of the sketch by LAPACK dgeqrf in
  place of SuiteSparseQR (absent), LSQR with sparse matrix-vector products.
"""
import math
import os
import sys
import time

import numpy as np
import scipy.fft
import scipy.linalg as sl
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests"))
import oracle_py as O  # noqa: E402  (the libc-compatible pattern generator)


def dht_cols(X, workers=1, chunk=512):
    """Orthonormal DHT of each column: (Re F - Im F) / sqrt(n), F the DFT,
    from the real FFT (F[n - k] = conj F[k]); in column chunks."""
    n, d = X.shape
    out = np.empty((n, d), order="F")
    s = 1.0 / math.sqrt(n)
    for c0 in range(0, d, chunk):
        F = scipy.fft.rfft(X[:, c0:c0 + chunk], axis=0, workers=workers)
        h = F.shape[0]
        blk = np.empty((n, F.shape[1]))
        blk[:h] = F.real - F.imag
        # k >= h: F[k] = conj(F[n - k]) -> Re + Im of F[n - k]
        kk = np.arange(h, n)
        blk[h:] = F.real[n - kk] + F.imag[n - kk]
        out[:, c0:c0 + chunk] = blk * s
    return out


def hashing_matrix(m, n, rows, vals, k):
    """S (m x n, CSR) with k entries per column: S[rows[i k + t], i] = vals."""
    cols = np.repeat(np.arange(n, dtype=np.int64), k)
    return sp.csr_matrix((vals, (rows, cols)), shape=(m, n))


def lsqr(matvec, rmatvec, b, r, tol, max_it):
    """Meng's LSQR (src/IterSolver.cpp:160-284): y for min ||op y - b||;
    stops when phibar alpha |c| < tol ||op||_F-estimate phibar."""
    u = b.copy()
    beta = np.linalg.norm(u)
    v = np.zeros(r)
    alpha = 0.0
    if beta > 0:
        u /= beta
        v = rmatvec(u)
        alpha = np.linalg.norm(v)
    if alpha > 0:
        v /= alpha
    w = v.copy()
    x = np.zeros(r)
    if alpha * beta == 0:
        return x, 0, 0
    phibar, rhobar, nrm_a = beta, alpha, 0.0
    for it in range(1, max_it + 1):
        u = matvec(v) - alpha * u
        beta = np.linalg.norm(u)
        if beta > 0:
            u /= beta
            nrm_a = math.sqrt(nrm_a * nrm_a + alpha * alpha + beta * beta)
            v = rmatvec(u) - beta * v
            alpha = np.linalg.norm(v)
            if alpha > 0:
                v /= alpha
        rho = math.sqrt(rhobar * rhobar + beta * beta)
        c, s = rhobar / rho, beta / rho
        theta = s * alpha
        rhobar = -c * alpha
        phi = c * phibar
        phibar = s * phibar
        x += (phi / rho) * w
        w = v - (theta / rho) * w
        if phibar * alpha * abs(c) < tol * nrm_a * phibar:
            return x, 0, it
    return x, 1, max_it


def dense_solve(A, b, gamma=1.7, k=1, abs_tol=1e-8, rcond=1e-12, it_tol=1e-10,
                max_it=10000, seed=1, workers=1, lsqr_it_cap=None):
    """ls_dense_hashing_blendenpik on the host.  Returns x, info (rank, flag,
    it and per-phase seconds).  lsqr_it_cap stops LSQR early (for timing
    samples; the info then says how many iterations ran)."""
    n, d = A.shape
    m = int(math.ceil(d * gamma))
    t0 = time.perf_counter()
    D, rows, vals = O.dense_pattern(seed, n, m, k)
    S = hashing_matrix(m, n, rows, vals * (1.0 / math.sqrt(k)), k)
    HDA = dht_cols(A * D[:, None], workers)
    SA = np.asfortranarray((S @ HDA))
    c = S @ dht_cols((b * D)[:, None], workers)[:, 0]
    del HDA
    t1 = time.perf_counter()
    cQ, R, P = sl.qr_multiply(SA, c[None, :], mode="right", pivoting=True)
    qc = cQ[0]
    dg = np.abs(np.diag(R))
    rank = int(np.argmax(dg <= rcond)) if np.any(dg <= rcond) else d
    R = R[:rank, :rank]
    t2 = time.perf_counter()
    x = np.zeros(d)
    x[P[:rank]] = sl.solve_triangular(R, qc[:rank])
    rs = b - A @ x
    info = {"rank": rank, "x0_residual": float(np.linalg.norm(rs))}
    t3 = time.perf_counter()
    flag, it = 0, 0
    if info["x0_residual"] >= abs_tol:
        Ar = np.asfortranarray(A[:, P[:rank]])  # the reference permutes A in place (dlapmt)
        mv = lambda v: Ar @ sl.solve_triangular(R, v)
        rmv = lambda u: sl.solve_triangular(R, Ar.T @ u, trans="T")
        y, flag, it = lsqr(mv, rmv, rs, rank, it_tol,
                           max_it if lsqr_it_cap is None else lsqr_it_cap)
        x[P[:rank]] += sl.solve_triangular(R, y)
    t4 = time.perf_counter()
    info.update(flag=flag, it=it, t_sketch=t1 - t0, t_qr=t2 - t1, t_x0=t3 - t2,
                t_lsqr=t4 - t3, t_total=t4 - t0)
    return x, info


def sparse_solve(A, b, gamma=1.5, k=3, abs_tol=1e-8, it_tol=1e-10, max_it=10000, seed=1,
                 lsqr_it_cap=None):
    """ls_sparse_spqr on the host with LAPACK dgeqrf in place of SuiteSparseQR
    (natural column order; the rank-deficient refactorisation is left out:
    the timing workloads are full rank)."""
    A = sp.csc_matrix(A)
    n, d = A.shape
    m = int(math.ceil(d * gamma))
    t0 = time.perf_counter()
    rows, vals = O.sparse_pattern(seed, n, m, k)
    S = hashing_matrix(m, n, rows, vals, k)
    SA = (S @ A).toarray()
    c = S @ b
    t1 = time.perf_counter()
    cQ, R = sl.qr_multiply(SA, c[None, :], mode="right")
    qc = cQ[0]
    t2 = time.perf_counter()
    x = sl.solve_triangular(R, qc[:d])
    rs = b - A @ x
    info = {"rank": d, "x0_residual": float(np.linalg.norm(rs))}
    t3 = time.perf_counter()
    flag, it = 0, 0
    if info["x0_residual"] >= abs_tol:
        Ar = A.tocsr()
        mv = lambda v: Ar @ sl.solve_triangular(R, v)
        rmv = lambda u: sl.solve_triangular(R, Ar.T @ u, trans="T")
        y, flag, it = lsqr(mv, rmv, rs, d, it_tol, max_it if lsqr_it_cap is None else lsqr_it_cap)
        x += sl.solve_triangular(R, y)
    t4 = time.perf_counter()
    info.update(flag=flag, it=it, t_sketch=t1 - t0, t_qr=t2 - t1, t_x0=t3 - t2,
                t_lsqr=t4 - t3, t_total=t4 - t0)
    return x, info


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """Cores this process may use (sched_getaffinity), the 'all cores' run."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1
