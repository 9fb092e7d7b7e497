"""Pinning the CPU oracle (no GPU): against libc, the committed golden
fixtures, numpy/scipy/LAPACK routines and closed-form properties."""
import ctypes
import hashlib
import json
import math
import os
import sys

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from conftest import GOLDEN, c1_problem, hqrrp_block_mismatches, read_mtx, rel, sparse_problem

libc = ctypes.CDLL("libc.so.6")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ---------------------------------------------------------------- RNG / pattern
@pytest.mark.parametrize("seed", [0, 1, 12, 1234, 123456, 2**31 + 5, 4000000000])
def test_rand_matches_libc(O, seed):
    libc.srand(ctypes.c_uint(seed))
    ref = [libc.rand() for _ in range(2000)]
    assert np.array_equal(O.rand_seq(seed, 2000), ref)


def test_rand_known_values(O):
    # glibc rand() after srand(1) and srand(12)
    assert list(O.rand_seq(1, 3)) == [1804289383, 846930886, 1681692777]
    assert list(O.rand_seq(12, 2)) == [1687063760, 945274514]


def test_patterns_match_golden(O):
    with open(os.path.join(GOLDEN, "patterns.json")) as fh:
        pats = json.load(fh)
    for name, e in pats.items():
        if e["dense"]:
            D, rows, vals = O.dense_pattern(e["seed"], e["n"], e["m"], e["k"])
            assert D[:64].tolist() == e["D64"] and sha(D) == e["D_sha"], name
        else:
            rows, vals = O.sparse_pattern(e["seed"], e["n"], e["m"], e["k"])
        assert rows[:64].tolist() == e["rows64"], name
        assert vals[:64].tolist() == e["vals64"], name
        assert sha(rows) == e["rows_sha"] and sha(vals) == e["vals_sha"], name


def test_pattern_properties(O):
    m, n, k = 37, 500, 4
    rows, vals = O.sparse_pattern(9, n, m, k)
    R = rows.reshape(n, k)
    assert all(len(set(r)) == k for r in R)          # k distinct rows per column
    assert rows.min() >= 0 and rows.max() < m
    assert set(np.round(vals * math.sqrt(k), 12)) == {-1.0, 1.0}
    counts = np.bincount(rows, minlength=m)          # roughly uniform buckets
    assert counts.min() > 0.5 * n * k / m and counts.max() < 1.6 * n * k / m


def test_gaussian_matches_golden(O):
    with open(os.path.join(GOLDEN, "gaussian.json")) as fh:
        g = json.load(fh)
    G = O.gaussian_matrix(O.rng(12), g["rows"], g["cols"])
    assert G[:, 0].tolist() == g["first_col"]
    assert G[:, 1].tolist() == g["col1"]
    assert sha(np.asfortranarray(G).T) == g["sha"]


# ---------------------------------------------------------------- transforms
@pytest.mark.parametrize("n", [1, 2, 3, 7, 16, 50, 97, 100, 120, 997, 4000, 4001, 40000])
def test_fft_matches_numpy(O, n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(O.fft(z), np.fft.fft(z)) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 5, 50, 97, 4000, 4001])
def test_dht_closed_form(O, n):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, 3))
    H = O.dht_cols(X)
    # FFTW_DHT: Y_k = sum_j x_j cas(2 pi jk / n)
    if n <= 97:
        j = np.arange(n)
        ang = 2 * np.pi * np.outer(j, j) / n
        assert rel(H, (np.cos(ang) + np.sin(ang)) @ X) < 1e-13
    F = np.fft.fft(X, axis=0)
    assert rel(H, F.real - F.imag) < 1e-13
    # involution: DHT(DHT(x)) = n x
    assert rel(O.dht_cols(H), n * X) < 1e-12


def test_fut_orthonormal(O):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((4000, 4))
    D = O.dense_pattern(5, 4000, 10, 1)[0]
    Y = O.fut(A, D)
    assert np.allclose(np.linalg.norm(Y, axis=0), np.linalg.norm(A, axis=0), rtol=1e-13)
    # first entry: (1/sqrt n) sum_j D_j A_j
    assert abs(Y[0, 0] - (D * A[:, 0]).sum() / np.sqrt(4000)) < 1e-12


# ---------------------------------------------------------------- sketch
def dense_S(m, rows, vals, k, n):
    S = np.zeros((m, n))
    for i in range(n):
        for t in range(k):
            S[rows[k * i + t], i] += vals[k * i + t]
    return S


@pytest.mark.parametrize("k", [1, 3])
def test_hash_apply_brute_force(O, k):
    rng = np.random.default_rng(k)
    n, m = 300, 40
    rows, vals = O.sparse_pattern(17, n, m, k)
    X = rng.standard_normal((n, 5))
    alpha = 1 / np.sqrt(k)
    Y = O.hash_apply(m, rows, vals, k, alpha, X)
    assert rel(Y, alpha * dense_S(m, rows, vals, k, n) @ X) < 1e-14
    if k == 1:  # each bucket sum of +-x: exact in increasing source order
        S = dense_S(m, rows, vals, k, n)
        for h in range(m):
            acc = 0.0
            for i in np.nonzero(S[h])[0]:
                acc = acc + S[h, i] * X[i, 0]
            assert Y[h, 0] == acc


def test_sparse_sketch_matches_scipy(O):
    A, b = sparse_problem(1, n=2000, d=30)
    m, k = 45, 3
    rows, vals = O.sparse_pattern(4, 2000, m, k)
    Ap = np.ascontiguousarray(A.indptr, dtype=np.int64)
    Ai = np.ascontiguousarray(A.indices, dtype=np.int64)
    SA = O.sketch_sparse(m, rows, vals, k, Ap, Ai, A.data, 2000, 30)
    S = sp.csc_matrix(dense_S(m, rows, vals, k, 2000))
    assert rel(SA, (S @ A).toarray()) < 1e-14
    assert rel(O.sketch_vec(m, rows, vals, k, b), S @ b) < 1e-14


# ---------------------------------------------------------------- QR family
def test_geqrf_matches_lapack(O):
    rng = np.random.default_rng(0)
    A = rng.standard_normal((90, 40))
    QR, tau = O.geqrf(A)
    (qr_l, tau_l), _ = sl.qr(A, mode="raw")
    assert rel(np.triu(QR[:40]), np.triu(qr_l[:40])) < 1e-13  # same dlarfg signs
    assert rel(tau, tau_l) < 1e-13
    assert rel(np.tril(QR, -1), np.tril(qr_l, -1)) < 1e-12


def test_geqp3_pivots_match_lapack(O):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((120, 60)) * np.logspace(0, -5, 60)
    QR, jpvt, tau = O.geqp3(A)
    _, R, P = sl.qr(A, pivoting=True)
    assert np.array_equal(jpvt - 1, P)
    assert rel(np.triu(QR[:60]), R[:60]) < 1e-12


def q_from(QR, tau):
    m, n = QR.shape
    Q = np.eye(m)
    for i in range(len(tau) - 1, -1, -1):
        v = np.zeros(m)
        v[i] = 1
        v[i + 1:] = QR[i + 1:, i]
        Q = Q - tau[i] * np.outer(v, v @ Q)
    return Q


def test_hqrrp_invariants_and_first_block(O):
    rng = np.random.default_rng(2)
    m, n = 300, 200
    A = rng.standard_normal((m, n)) * np.logspace(0, -3, n)
    QR, jpvt, tau = O.hqrrp(A, nb=16, pp=10)
    assert sorted(jpvt) == list(range(1, n + 1))
    Q = q_from(QR, tau)
    R = np.triu(QR[:n])
    assert np.abs(Q.T @ Q - np.eye(m)).max() < 1e-13
    assert rel(A[:, jpvt - 1], Q[:, :n] @ R) < 1e-13
    # the first block's pivots are the greedy QRP pivots of Y = G A
    G = O.gaussian_matrix(O.rng(12), 26, m)
    _, _, P = sl.qr(G @ A, pivoting=True)
    assert sorted(jpvt[:16] - 1) == sorted(P[:16])


def test_hqrrp_vs_dgeqp3_on_large(O):
    # min(m, n) >= 1000 takes the HQRRP branch of dgeqp4_INT
    rng = np.random.default_rng(3)
    A = rng.standard_normal((1100, 1000))
    QR, jpvt, tau = O.cpqr(A)
    R = np.triu(QR[:1000])
    d = np.abs(np.diag(R))
    # R is the R factor of A P: column norms of R equal those of A P
    assert np.allclose(np.linalg.norm(R, axis=0), np.linalg.norm(A[:, jpvt - 1], axis=0),
                       rtol=1e-12)
    assert d.min() > 0


@pytest.mark.parametrize("m,n,nb,seed", [(400, 300, 16, 7), (600, 520, 32, 8),
                                         (1100, 1000, 64, 9)])
def test_hqrrp_every_block_pivots(O, m, n, nb, seed):
    """Pins the oracle's HQRRP beyond its first block: every block's pivot set
    and panel order follow from the exactly downdated sketch."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) * np.logspace(0, -3, n)
    QR, jpvt, tau = O.hqrrp(A, nb=nb, pp=10)
    assert hqrrp_block_mismatches(O, A, QR, jpvt, tau, nb, 10) == []


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_hqrrp_block_check_catches_planted_downdate_mistakes(O, tmp_path, plant):
    """Each planted mistake in the Y downdate still yields a valid
    factorization A P = Q R (the invariants alone cannot see it), but the
    per-block pivot check above fails on it."""
    import subprocess
    import oracle_py
    src = os.path.join(os.path.dirname(oracle_py.LIB_PATH), "skl_oracle.c")
    so = str(tmp_path / f"liboracle_plant{plant}.so")
    subprocess.check_call(["gcc", "-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off",
                           f"-DORC_PLANT={plant}", "-shared", "-o", so, src, "-lm"])
    hq = oracle_py.load_variant(so)
    rng = np.random.default_rng(7)
    m, n, nb = 400, 300, 16
    A = rng.standard_normal((m, n)) * np.logspace(0, -3, n)
    QR, jpvt, tau = hq(A, nb=nb, pp=10)
    Q = q_from(QR, tau)
    assert rel(A[:, jpvt - 1], Q[:, :n] @ np.triu(QR[:n])) < 1e-13
    assert len(hqrrp_block_mismatches(O, A, QR, jpvt, tau, nb, 10)) > 0


def test_ormqr_and_trsv(O):
    rng = np.random.default_rng(4)
    A = rng.standard_normal((70, 30))
    c = rng.standard_normal(70)
    QR, tau = O.geqrf(A)
    Q = q_from(QR, tau)
    assert rel(O.ormqr_lt(QR, tau, c), Q.T @ c) < 1e-13
    R = np.triu(QR[:30, :30])
    x = rng.standard_normal(30)
    assert rel(O.trsv_un(R, x), sl.solve_triangular(R, x)) < 1e-12
    assert rel(O.trsv_ut(R, x), sl.solve_triangular(R, x, trans="T")) < 1e-12
    Rp = R + 0.5 * np.eye(30)
    assert rel(O.trsv_un(R, x, 0.5), sl.solve_triangular(Rp, x)) < 1e-12
    assert rel(O.trsv_ut(R, x, 0.5), sl.solve_triangular(Rp, x, trans="T")) < 1e-12


def test_rcond_vs_lapack_dtrcon(O):
    rng = np.random.default_rng(5)
    for scale in [0, -3, -8]:
        R = np.triu(rng.standard_normal((40, 40))) + 4 * np.eye(40)
        R = R * np.logspace(0, scale, 40)
        exact = O.rcond_inf(R)
        est, info = sl.lapack.dtrcon(R, norm="I", uplo="U", diag="N")
        # dtrcon's estimate of ||R^-1|| is a lower bound: rcond_est >= exact
        assert exact <= est * (1 + 1e-12) and exact >= est / 20
        Ri = np.linalg.inv(R)
        assert np.isclose(exact, 1 / (np.abs(R).sum(1).max() * np.abs(Ri).sum(1).max()),
                          rtol=1e-10)


# ---------------------------------------------------------------- LSQR
def test_lsqr_iterates_match_scipy(O):
    """LSQR iterates are fixed by the Krylov space: the k-th iterate of the
    reference recurrences equals scipy's LSQR after k steps."""
    rng = np.random.default_rng(6)
    A = rng.standard_normal((400, 30)) * np.logspace(0, -2, 30)
    b = rng.standard_normal(400)
    R = np.eye(30)
    for k in [1, 3, 8]:
        y, flag, it, _, _ = O.lsqr_dense(A, R, b, 0.0, k)
        assert it == k and flag == 1
        ys = spl.lsqr(A, b, atol=0, btol=0, conlim=0, iter_lim=k)[0]
        assert rel(y, ys) < 1e-10


def test_lsqr_estimates(O):
    rng = np.random.default_rng(7)
    A = rng.standard_normal((300, 20))
    b = rng.standard_normal(300)
    R = np.eye(20)
    for k in [2, 5, 9]:
        y, flag, it, steps, info = O.lsqr_dense(A, R, b, 0.0, k, hist=True)
        r = b - A @ y
        # phibar = ||r_k||, phibar alpha |c| = ||A^T r_k|| in exact arithmetic
        assert np.isclose(steps[-1, 3], np.linalg.norm(r), rtol=1e-9)
        assert np.isclose(info.lsqr_nrm_ar, np.linalg.norm(A.T @ r), rtol=1e-6)


def test_lsqr_zero_rhs(O):
    A = np.random.default_rng(8).standard_normal((50, 5))
    y, flag, it, _, _ = O.lsqr_dense(A, np.eye(5), np.zeros(50), 1e-10, 100)
    assert it == 0 and flag == 0 and not y.any()


# ---------------------------------------------------------------- solvers
def test_dense_solver_matches_gelsd(O):
    A, b = c1_problem(0, n=2000, d=100)
    xr = sl.lstsq(A, b, lapack_driver="gelsd")[0]
    x, info, _ = O.ls_dense(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=3)
    assert rel(x, xr) < 1e-8 and info.flag == 0 and info.rank == 100
    x2, info2, _ = O.ls_dense_nocpqr(A, b, gamma=1.7, k=1, it_tol=1e-10, seed=3)
    assert rel(x2, xr) < 1e-8 and info2.flag == 0


def test_subsample_is_the_hashing_draw_for_one_column(O):
    """select_k_from_n is the step gen_hashing_matrix takes per column
    (sparse_helpers.cpp:98-104); with one column the two draw the same rows
    from the same stream, and gen_hashing is pinned by the golden patterns."""
    for seed, n, k in [(1, 4000, 340), (7, 500, 500), (123, 40000, 6800)]:
        idx = O.select_k_from_n(O.rng(seed), n, k)
        rows, _ = O.sparse_pattern(seed, 1, n, k)
        assert np.array_equal(idx, rows), seed
        assert len(set(idx.tolist())) == k and idx.min() >= 0 and idx.max() < n


def test_subsample_pattern_restarts_the_stream(O):
    """ls_dense_blendenpik seeds twice from the same clock: the signs and the
    sampled rows both start at the first draw of the stream."""
    seed, n, m = 5, 1000, 120
    D, idx = O.subsample_pattern(seed, n, m)
    r = O.rand_seq(seed, m)
    assert np.array_equal(D[:m], (r % 2) * 2.0 - 1.0)
    arr = list(range(n))
    for i in range(m):  # brute-force Fisher-Yates on the libc values
        j = int(r[i]) % (n - i) + i
        arr[i], arr[j] = arr[j], arr[i]
    assert np.array_equal(idx, arr[:m])


def test_blendenpik_solver_matches_gelsd(O):
    A, b = c1_problem(2, n=2000, d=100)
    xr = sl.lstsq(A, b, lapack_driver="gelsd")[0]
    x, info, _ = O.ls_dense_blendenpik(A, b, gamma=4.0, it_tol=1e-10, seed=3)
    assert rel(x, xr) < 1e-8 and info.flag == 0 and info.rank == 100
    with pytest.raises(RuntimeError):   # more sampled rows than A has
        O.ls_dense_blendenpik(A[:150], b[:150], gamma=1.7)


def test_blendenpik_sketch_is_sampled_transform(O):
    """The R of the oracle's Blendenpik preconditioner is the R of the sampled
    rows of the randomized DHT: R^T R = (P F D A)^T (P F D A)."""
    rng = np.random.default_rng(4)
    n, d, m = 300, 12, 40
    A = rng.standard_normal((n, d))
    D, idx = O.subsample_pattern(9, n, m)
    F = O.fut(np.eye(n), D)              # the transform applied to the identity
    SA = (F @ A)[idx]
    assert np.allclose(SA, O.fut(A, D)[idx], atol=1e-13)
    R = sl.qr(SA, mode="r")[0][:d]
    assert np.allclose(R.T @ R, SA.T @ SA, rtol=1e-10, atol=1e-10)


def test_sparse_solver_matches_gelsd(O):
    A, b = sparse_problem(0, n=3000, d=80)
    xr = sl.lstsq(A.toarray(), b, lapack_driver="gelsd")[0]
    x, info, _ = O.ls_sparse(A, b, gamma=1.5, k=3, it_tol=1e-10)
    assert rel(x, xr) < 1e-8 and info.flag == 0 and info.rank == 80 and info.success == 1


def test_examples_known_answers(O, kat):
    e = kat["dense_example"]
    A = np.array(e["A"]).reshape(e["n"], e["m"]).T
    x, info, _ = O.ls_dense(A, np.array(e["b"]), gamma=1.1, k=3, rcond=1e-10, it_tol=1e-6,
                            seed=1)
    assert np.allclose(x, e["x"], atol=1e-6)
    s = kat["sparse_example"]
    As = sp.csc_matrix((s["val"], s["rowind"], s["colptr"]), shape=(s["m"], s["n"]))
    x, info, _ = O.ls_sparse(As, np.array(s["b"]), gamma=1.1, k=3, it_tol=1e-6)
    assert np.allclose(x, s["x"], atol=1e-6)


def test_full_rank_fixtures(O, kat):
    for f, r_ref in kat["full_rank_gelsd_residual"].items():
        A = read_mtx(os.path.join(GOLDEN, "data", f))
        b = np.ones(A.shape[0])
        # test/ls_blendenpik_hashing.cpp:60-66 parameters
        x, info, _ = O.ls_dense(A, b, gamma=1.7, k=1, abs_tol=1e-8, rcond=1e-12, it_tol=1e-6,
                                seed=1)
        assert abs(np.linalg.norm(A @ x - b) - r_ref) < 1e-6 * max(1, r_ref), f


def test_rank_deficient_residual_kats(O, kat):
    # test/test_solvers.sh:28-31 through the dense robust solver on the
    # densified matrix (test/ls_blendenpik_hashing_sparse_input.cpp)
    for name, r_ref in kat["rank_deficient"].items():
        A = read_mtx(os.path.join(GOLDEN, "data", "rank_deficient", name + ".mtx"))
        if A.shape[0] * A.shape[1] > 4e6:
            continue
        b = np.ones(A.shape[0])
        x, info, _ = O.ls_dense(A, b, gamma=1.7, k=1, abs_tol=1e-8, rcond=1e-12,
                                it_tol=1e-6, max_it=10000, seed=1)
        r = np.linalg.norm(A @ x - b)
        assert abs(r - r_ref) < 1e-4 * max(1.0, r_ref) + 1e-5, (name, r, r_ref)


def test_overdetermined_error(O):
    A = np.ones((3, 5))
    with pytest.raises(RuntimeError, match="Matrix is not overdetermined."):
        O.ls_dense(A, np.ones(3))
    with pytest.raises(RuntimeError, match="Matrix is not overdetermined."):
        O.ls_sparse(sp.csc_matrix(A), np.ones(3))


def test_cpu_ref_matches_oracle(O):
    """The BLAS/LAPACK restatement that serves as the CPU baseline computes
    what the oracle computes: dense (4000 x 200 configuration) and sparse."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                    "oracle"))
    import cpu_ref
    A, b = c1_problem(3)
    kw = dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=5)
    xo, io, _ = O.ls_dense(A, b, **kw)
    x, info = cpu_ref.dense_solve(A, b, **kw)
    assert info["rank"] == io.rank and info["flag"] == 0
    assert abs(info["it"] - io.it) <= 2
    assert rel(x, xo) < 1e-8
    As, bs = sparse_problem(4, n=3000, d=80, density=0.03)
    kws = dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)
    xo, io, _ = O.ls_sparse(As, bs, **kws)
    x, info = cpu_ref.sparse_solve(As, bs, **kws)
    assert info["flag"] == 0 and abs(info["it"] - io.it) <= 2
    assert rel(x, xo) < 1e-8
