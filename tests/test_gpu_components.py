"""Stage-by-stage parity of the HIP kernels with the oracle, through the C ABI.

Tolerances: integer / pattern work and the sketches whose summation order
follows the reference are bit-exact; FP64 transforms and factorizations are
compared at 1e-12..1e-13 relative (rounding of a different but equally
stable evaluation order); FP32 variants at 1e-5."""
import numpy as np
import pytest
import scipy.linalg as sl

from conftest import rel, sparse_problem

pytestmark = pytest.mark.gpu


def test_device_present(skl):
    assert skl.lib().skl_device_count() >= 1
    assert skl.lib().skl_init(0) == 0


@pytest.mark.parametrize("n", [1, 2, 50, 97, 120, 340, 1001, 2798, 4000, 4096, 5533, 7000, 12000,
                               40000, 131072, 199999])
def test_dht_fp64(skl, O, n):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, 5))
    assert rel(skl.dht(X), O.dht_cols(X)) < 1e-13


@pytest.mark.parametrize("n", [50, 1399, 4000, 7000, 40000])
def test_dht_fp32(skl, O, n):
    rng = np.random.default_rng(n + 1)
    X = rng.standard_normal((n, 4))
    assert rel(skl.dht(X, skl.SKL_MIXED), O.dht_cols(X)) < 2e-6


@pytest.mark.parametrize("k", [1, 3])
def test_hash_apply_bit_exact(skl, O, k):
    rng = np.random.default_rng(k)
    n, m = 4000, 340
    rows, vals = O.sparse_pattern(3, n, m, k)
    X = rng.standard_normal((n, 7))
    a = 1 / np.sqrt(k)
    assert np.array_equal(skl.hash_apply(m, rows, vals, k, a, X), O.hash_apply(m, rows, vals, k, a, X))


@pytest.mark.parametrize("n,d,m,k", [(4000, 20, 340, 1), (4000, 7, 40, 3), (40000, 6, 300, 1),
                                     (40000, 37, 600, 2)])
def test_dense_sketch(skl, O, n, d, m, k):
    rng = np.random.default_rng(n + d)
    A = rng.standard_normal((n, d))
    b = rng.standard_normal(n)
    D, rows, vals = O.dense_pattern(11, n, m, k)
    SA = skl.dense_sketch(A, b, m, k, D, rows, vals)
    ref = O.hash_apply(m, rows, vals, k, 1 / np.sqrt(k), O.fut(np.column_stack([A, b]), D))
    assert rel(SA, ref) < 1e-13
    SA32 = skl.dense_sketch(A, b, m, k, D, rows, vals, precision=skl.SKL_MIXED)
    assert rel(SA32, ref) < 1e-5


def test_sparse_sketch_bit_exact(skl, O):
    A, b = sparse_problem(3, n=6000, d=120)
    m, k = 180, 3
    rows, vals = O.sparse_pattern(1, 6000, m, k)
    SA, bs = skl.sparse_sketch(m, rows, vals, k, A, b)
    Ap = np.ascontiguousarray(A.indptr, dtype=np.int64)
    Ai = np.ascontiguousarray(A.indices, dtype=np.int64)
    assert np.array_equal(SA, O.sketch_sparse(m, rows, vals, k, Ap, Ai, A.data, 6000, 120))
    assert np.array_equal(bs, O.sketch_vec(m, rows, vals, k, b))


@pytest.mark.parametrize("m,n", [(60, 40), (300, 120), (680, 400), (1100, 700)])
def test_qr_fp64(skl, O, m, n):
    rng = np.random.default_rng(m)
    A = rng.standard_normal((m, n + 1)) * np.append(np.logspace(0, -4, n), 1.0)
    W, tau = skl.qr(A, nfac=n)
    Wo, tauo = O.geqrf(A[:, :n])
    assert rel(np.triu(W[:n, :n]), np.triu(Wo[:n])) < 1e-12
    assert rel(tau, tauo) < 1e-12
    # the carried column receives Q^T c
    assert rel(W[:, n], O.ormqr_lt(Wo, tauo, A[:, n])) < 1e-12


def test_qr_fp32(skl, O):
    rng = np.random.default_rng(9)
    A = rng.standard_normal((500, 300))
    W, tau = skl.qr(A, precision=skl.SKL_MIXED)
    Wo, _ = O.geqrf(A)
    assert rel(np.triu(W[:300]), np.triu(Wo[:300])) < 1e-5


@pytest.mark.parametrize("m,n", [(80, 50), (340, 200), (700, 420)])
def test_cpqr_greedy_pivots(skl, O, m, n):
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n)) * np.logspace(0, -4, n)
    W, jpvt, tau = skl.cpqr(A)
    Wo, jo, to = O.geqp3(A)
    assert np.array_equal(jpvt, jo)
    assert rel(np.triu(W[:n]), np.triu(Wo[:n])) < 1e-12


def test_cpqr_hqrrp_pivots(skl, O):
    rng = np.random.default_rng(5)
    m, n = 1200, 1000
    A = rng.standard_normal((m, n)) * np.logspace(0, -3, n)
    W, jpvt, tau = skl.cpqr(A)
    Wo, jo, to = O.cpqr(A)
    assert np.array_equal(jpvt, jo)
    assert rel(np.triu(W[:n]), np.triu(Wo[:n])) < 1e-11


@pytest.mark.parametrize("scale", ["logspace(0,-6)", "ones", "logspace(0,-0.7)"])
def test_hqrrp_leading_column_search_same_bits(skl, O, monkeypatch, scale):
    """HQRRP's block pivots first come from a one-workgroup search over the
    columns of largest initial norm, which stops when a left-out column could
    win; the grid-wide search then runs.  With widely scaled columns it
    finishes every block, with equal norms it stops at once, and with mildly
    scaled ones some blocks stop part way.  In each case the factorization is
    bit for bit the one of the grid-wide search alone (SKL_YQRP_TOP=0), and
    the pivots are the oracle's."""
    rng = np.random.default_rng(21)
    m, n = 1200, 1000
    sc = {"logspace(0,-6)": np.logspace(0, -6, n), "ones": np.ones(n),
          "logspace(0,-0.7)": np.logspace(0, -0.7, n)}[scale]
    A = rng.standard_normal((m, n)) * sc
    monkeypatch.setenv("SKL_YQRP_TOP", "0")
    W0, j0, t0 = skl.cpqr(A)
    monkeypatch.setenv("SKL_YQRP_TOP", "1")
    W1, j1, t1 = skl.cpqr(A)
    assert np.array_equal(j0, j1)
    assert np.array_equal(t0, t1)
    assert np.array_equal(W0, W1)
    Wo, jo, to = O.cpqr(A)
    assert np.array_equal(j1, jo)
    assert rel(np.triu(W1[:n]), np.triu(Wo[:n])) < 1e-11


@pytest.mark.parametrize("r", [1, 31, 32, 33, 200, 1000])
def test_trinv(skl, r):
    rng = np.random.default_rng(r)
    R = np.triu(rng.standard_normal((r, r))) + 3 * np.eye(r)
    M = skl.trinv(R)
    assert rel(M, sl.solve_triangular(R, np.eye(r))) < 1e-12
    Mp = skl.trinv(R, perturb=0.25)
    assert rel(Mp, sl.solve_triangular(R + 0.25 * np.eye(r), np.eye(r))) < 1e-12


def test_lsqr_steps_match_oracle(skl, O):
    rng = np.random.default_rng(12)
    n, d = 5000, 60
    A = rng.standard_normal((n, d)) * np.logspace(0, -3, d)
    b = rng.standard_normal(n)
    # a rough preconditioner so that LSQR takes tens of iterations
    R = np.linalg.qr(A[:: 7], mode="r")
    M = sl.solve_triangular(R, np.eye(d))
    y, flag, it, steps = skl.lsqr_dense(A, M, b, 1e-12, 500, history=True)
    yo, flo, ito, so, _ = O.lsqr_dense(A, R, b, 1e-12, 500, hist=True)
    assert flag == flo == 0 and abs(it - ito) <= 1
    k = min(it, ito) - 1
    assert rel(steps[:k, :4], so[:k, :4]) < 1e-9
    assert rel(y, yo) < 1e-9


def test_lsqr_permuted_columns(skl, O):
    rng = np.random.default_rng(13)
    n, d = 3000, 40
    A = rng.standard_normal((n, d))
    b = rng.standard_normal(n)
    E = rng.permutation(d)[:30]
    R = np.linalg.qr(A[:, E], mode="r")
    M = sl.solve_triangular(R, np.eye(30))
    y, flag, it, _ = skl.lsqr_dense(A, M, b, 1e-12, 200, E=E)
    yo, *_ = O.lsqr_dense(A, R, b, 1e-12, 200, E=E)
    assert rel(y, yo) < 1e-10


def test_grid_barrier_give_up_is_reported(skl, O):
    """A grid barrier whose bounded wait gives up must surface as SKL_E_HIP,
    never as a factor computed from stale data.  The poll limit is forced to
    zero from the host, so the first workgroup to arrive at a barrier gives
    up and every kernel still exits normally (k_common.h:grid_barrier)."""
    rng = np.random.default_rng(11)
    A = rng.standard_normal((3000, 300)) * np.logspace(0, -3, 300)  # multi-workgroup panels
    Ah = rng.standard_normal((1200, 1000)) * np.logspace(0, -3, 1000)  # HQRRP
    x_star = rng.standard_normal(300)
    b = A @ x_star
    skl.set_barrier_spins(0)
    try:
        for call in (lambda: skl.cpqr(A), lambda: skl.qr(A, precision=skl.SKL_MIXED),
                     lambda: skl.cpqr(Ah),
                     lambda: skl.dense_solve(np.vstack([A, A]), np.concatenate([b, b]),
                                             cpqr=True, it_tol=1e-10, seed=3)):
            with pytest.raises(skl.SklError) as ei:
                call()
            assert ei.value.code == 4
            assert "grid barrier" in str(ei.value)
    finally:
        skl.set_barrier_spins(-1)
    # the error word was cleared: the same factorizations succeed again
    W, jpvt, tau = skl.cpqr(A)
    Wo, jo, to = O.geqp3(A)
    assert np.array_equal(jpvt, jo)
    assert rel(np.triu(W[:300]), np.triu(Wo[:300])) < 1e-12
    W, jpvt, tau = skl.cpqr(Ah)
    Wo, jo, to = O.cpqr(Ah)
    assert np.array_equal(jpvt, jo)


def _q_from(W, tau):
    m, n = W.shape
    Q = np.eye(m)
    for i in range(len(tau) - 1, -1, -1):
        v = np.zeros(m)
        v[i] = 1.0
        v[i + 1:] = W[i + 1:, i]
        Q = Q - tau[i] * np.outer(v, v @ Q)
    return Q


@pytest.mark.parametrize("pivot", [False, True])
def test_qr_panel_fallback_on_dependent_columns(skl, pivot):
    """A panel with a repeated column: the CholeskyQR2 panel gives up on it
    (the repeated column's partial norm vanishes) and the column-by-column
    kernel factors the untouched panel; the factorization is still exact:
    Q orthonormal and A P = Q R."""
    rng = np.random.default_rng(17)
    m, n = 1500, 1040 if pivot else 160
    A = rng.standard_normal((m, n)) * np.logspace(0, -2, n)
    A[:, 37] = A[:, 5]
    A[:, n - 3] = 2.0 * A[:, 100]
    if pivot:
        W, jpvt, tau = skl.cpqr(A)
        P = jpvt - 1
        assert sorted(P) == list(range(n))
    else:
        W, tau = skl.qr(A)
        P = np.arange(n)
    Q = _q_from(W, tau)
    R = np.triu(W[:n])
    assert np.abs(Q.T @ Q - np.eye(m)).max() < 1e-12
    assert rel(A[:, P], Q[:, :n] @ R) < 1e-13


def test_csc_duplicate_entries_repeatable(skl, O):
    """CSC input with duplicate (i, j) entries (each split value stored twice in
    its column): the CSR copy orders duplicates by CSC position, so repeated
    solves agree bit for bit; the solution is the one of the matrix the
    duplicates sum to."""
    import scipy.sparse as sp
    from conftest import rel, sparse_problem
    A, b = sparse_problem(31, n=6000, d=200, density=0.02)
    A = sp.csc_matrix(A)
    rng = np.random.default_rng(2)
    cols, rows, vals = [], [], []
    for j in range(A.shape[1]):
        r = A.indices[A.indptr[j]:A.indptr[j + 1]]
        v = A.data[A.indptr[j]:A.indptr[j + 1]]
        split = rng.random(len(r)) < 0.3
        w = rng.uniform(0.2, 0.8, len(r))
        rr = np.concatenate([r, r[split]])
        vv = np.concatenate([np.where(split, v * w, v), (v * (1 - w))[split]])
        rows.append(rr)
        vals.append(vv)
        cols.append(len(rr))
    indptr = np.concatenate([[0], np.cumsum(cols)])
    Ad = sp.csc_matrix((np.concatenate(vals), np.concatenate(rows), indptr), shape=A.shape)
    assert Ad.nnz > A.nnz  # duplicates kept (no sum_duplicates)
    kw = dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)
    x1, i1 = skl.sparse_solve(Ad, b, **kw)
    x2, i2 = skl.sparse_solve(Ad, b, **kw)
    assert np.array_equal(x1, x2) and i1.it == i2.it
    xo, io, _ = O.ls_sparse(A, b, **kw)
    assert rel(x1, xo) < 1e-8
