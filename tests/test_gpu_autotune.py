"""LSQR launch-shape autotuner (skl_autotune_lsqr) and the server's use of it:
the tuned shape is the fastest candidate on the synthetic problem, solves in a
tuned bucket still match the oracle, and with a fixed table they are bitwise
repeatable, alone and from a server that tunes new buckets itself."""
import numpy as np
import pytest

from conftest import c1_problem, rel, sparse_problem

pytestmark = pytest.mark.gpu

DENSE_KW = dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=21)
SPARSE_KW = dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)


def test_autotune_dense_and_sparse(skl, O):
    skl.autotune_clear()
    assert skl.autotune_entries() == 0
    A, b = c1_problem(12, n=6000, d=300)       # one-pass kernel bucket
    A2, b2 = c1_problem(13, n=3000, d=120)     # two-pass bucket (n < 4096)
    As, bs = sparse_problem(14, n=8000, d=300, density=0.02)
    for args in [(6000, 300), (3000, 120), (8000, 300, As.nnz)]:
        best, dflt = skl.autotune_lsqr(*args)
        assert 0 < best <= dflt, (args, best, dflt)
    assert skl.autotune_entries() == 3
    for AA, bb in [(A, b), (A2, b2)]:
        xo, io, _ = O.ls_dense(AA, bb, **DENSE_KW)
        x1, i1 = skl.dense_solve(AA, bb, **DENSE_KW)
        x2, i2 = skl.dense_solve(AA, bb, **DENSE_KW)
        assert np.array_equal(x1, x2) and abs(i1.it - io.it) <= 2
        assert rel(x1, xo) < 1e-8
    xo, io, _ = O.ls_sparse(As, bs, **SPARSE_KW)
    x1, i1 = skl.sparse_solve(As, bs, **SPARSE_KW)
    x2, _ = skl.sparse_solve(As, bs, **SPARSE_KW)
    assert np.array_equal(x1, x2) and abs(i1.it - io.it) <= 2
    assert rel(x1, xo) < 1e-8
    skl.autotune_clear()


def test_server_autotunes_new_buckets(skl):
    skl.autotune_clear()
    P = [c1_problem(20 + i, n=n, d=d) for i, (n, d) in enumerate([(5000, 200), (7000, 90)])]
    S2 = sparse_problem(30, n=9000, d=250, density=0.015)
    with skl.Server(slots=3) as S:
        S.set_autotune(True)
        tk = [S.submit_dense(A, b, **DENSE_KW) for A, b in P for _ in range(3)]
        tk += [S.submit_sparse(S2[0], S2[1], **SPARSE_KW) for _ in range(3)]
        out = [S.wait(t) for t in tk]
    assert skl.autotune_entries() == 3
    # every request of a bucket ran with the bucket's one tuned shape: the
    # repeats agree bit for bit, and with a solve made now
    for g in range(3):
        xs = [out[3 * g + r][0] for r in range(3)]
        assert all(np.array_equal(xs[0], x) for x in xs[1:]), g
    for i, (A, b) in enumerate(P):
        x, _ = skl.dense_solve(A, b, **DENSE_KW)
        assert np.array_equal(x, out[3 * i][0])
    x, _ = skl.sparse_solve(S2[0], S2[1], **SPARSE_KW)
    assert np.array_equal(x, out[6][0])
    skl.autotune_clear()
