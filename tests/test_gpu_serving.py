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
