"""Serving loop: requests solved concurrently on worker slots give the same
result, bit for bit, as the same request solved alone."""
import numpy as np
import pytest

from conftest import c1_problem, sparse_problem

pytestmark = pytest.mark.gpu


def _problems():
    P = []
    for i, (n, d) in enumerate([(3000, 60), (8000, 300), (2500, 25)]):
        A, b = c1_problem(10 + i, n=n, d=d)
        P.append(("dense", A, b, dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=100 + i)))
    for i, (n, d) in enumerate([(5000, 120), (3000, 40)]):
        A, b = sparse_problem(20 + i, n=n, d=d, density=0.02)
        P.append(("sparse", A, b, dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)))
    A, b = c1_problem(30, n=4000, d=100)
    P.append(("dense_mixed", A, b, dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=7,
                                        precision=1)))
    return P


def test_alone_equals_mixed(skl):
    P = _problems()
    alone = []
    for kind, A, b, kw in P:
        if kind == "sparse":
            x, info = skl.sparse_solve(A, b, **kw)
        else:
            x, info = skl.dense_solve(A, b, **kw)
        alone.append((x, info.it))
    with skl.Server(slots=4) as S:
        tickets = []
        for rep in range(2):
            for kind, A, b, kw in P:
                if kind == "sparse":
                    tickets.append(S.submit_sparse(A, b, **kw))
                else:
                    tickets.append(S.submit_dense(A, b, **kw))
        out = [S.wait(t) for t in tickets]
    for i, (x, info, q, r) in enumerate(out):
        xa, ita = alone[i % len(P)]
        diff = float(np.max(np.abs(x - xa)) / np.max(np.abs(xa)))
        assert info.it == ita and np.array_equal(x, xa), (i, P[i % len(P)][0], info.it, ita, diff)
        assert q >= 0 and r > 0


def test_server_admission_small_budget(skl):
    """A budget below one request's footprint still runs requests, one at a time."""
    A, b = c1_problem(40, n=3000, d=50)
    xa, ia = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, seed=3)
    with skl.Server(slots=3, mem_budget_bytes=1) as S:
        ts = [S.submit_dense(A, b, gamma=1.7, k=1, it_tol=1e-10, seed=3) for _ in range(3)]
        for t in ts:
            x, info, q, r = S.wait(t)
            assert np.array_equal(x, xa)


def _c5_pool(seed, count, max_dense_elems=10_000_000, max_sparse_elems=50_000_000):
    """Problems of the serving stream's distribution: d log-uniform in
    [20, 2000], n log-uniform in [10 d, 200000]; even entries dense (4000x200
    family: N(0,1) times logspace(0,-4,d) columns), odd entries CSC at density
    U[0.5%, 2%] (120000x5000 family: logspace(0,-8,d) columns, unit
    diagonal).  n is capped so that a dense problem holds at most
    max_dense_elems entries and a sparse one spans at most max_sparse_elems."""
    import math
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    pool = []
    for i in range(count):
        d = int(round(math.exp(rng.uniform(math.log(20), math.log(2000)))))
        n = int(round(math.exp(rng.uniform(math.log(10 * d), math.log(200000)))))
        if i % 2 == 0:
            n = min(n, max(10 * d, max_dense_elems // d))
            A = np.asfortranarray(rng.standard_normal((n, d)) * np.logspace(0, -4, d))
            b = A @ rng.standard_normal(d) + 1e-8 * rng.standard_normal(n)
            pool.append(("dense", A, b, dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12,
                                             seed=500 + i)))
        else:
            n = min(n, max(10 * d, max_sparse_elems // d))
            A = sp.random(n, d, density=rng.uniform(0.005, 0.02), random_state=rng,
                          format="csc", data_rvs=rng.standard_normal)
            A = sp.csc_matrix(A @ sp.diags(np.logspace(0, -8, d)) + sp.eye(n, d, format="csc"))
            A.sort_indices()
            b = A @ rng.standard_normal(d) + 1e-8 * rng.standard_normal(n)
            pool.append(("sparse", A, b, dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)))
    return pool


def test_serving_stream_scale(skl):
    """Two hundred requests of the serving distribution, FP64 and mixed
    precision interleaved, submitted faster than four slots can take them:
    every result equals, bit for bit, the same request solved alone, and each
    mixed-precision x is within 1e-8 of the FP64 solution of its problem."""
    pool = _c5_pool(4242, 24)
    alone = {}
    for i, (kind, A, b, kw) in enumerate(pool):
        for prec in (skl.SKL_FP64, skl.SKL_MIXED):
            solve = skl.sparse_solve if kind == "sparse" else skl.dense_solve
            x, info = solve(A, b, precision=prec, **kw)
            assert info.flag == 0, (i, kind, A.shape, prec)
            alone[i, prec] = (x, info.it)
        x64, xm = alone[i, skl.SKL_FP64][0], alone[i, skl.SKL_MIXED][0]
        assert np.linalg.norm(xm - x64) <= 1e-8 * np.linalg.norm(x64), (i, kind, A.shape)
    rng = np.random.default_rng(7)
    reqs = [(int(rng.integers(len(pool))), skl.SKL_FP64 if rng.random() < 0.3 else skl.SKL_MIXED)
            for _ in range(200)]
    with skl.Server(slots=4) as S:
        tickets = []
        for i, prec in reqs:
            kind, A, b, kw = pool[i]
            submit = S.submit_sparse if kind == "sparse" else S.submit_dense
            tickets.append(submit(A, b, precision=prec, **kw))
        for (i, prec), t in zip(reqs, tickets):
            x, info, q, r = S.wait(t)
            xa, ita = alone[i, prec]
            assert info.it == ita and np.array_equal(x, xa), (i, prec, info.it, ita)
