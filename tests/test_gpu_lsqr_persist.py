"""The persistent LSQR iteration (every iteration of a solve in one launch,
k_lsqr_persist, taken for rank <= 256 and at most 128 rows per workgroup)
against the oracle and against the launch-per-kernel iteration of the same
problem.  SKL_LSQR_PERSIST=0 turns it off, =P asks for P workgroups (at least
ceil(n / 128)); P decides whether a workgroup holds one or two 64-row chunks
of the operator and the order of the partial sums, so runs with different P
agree to rounding and runs with the same P bit for bit."""
import contextlib
import os

import numpy as np
import pytest
import scipy.linalg as sl

from conftest import c1_problem, rel, sparse_problem

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def persist(p):
    old = os.environ.get("SKL_LSQR_PERSIST")
    os.environ["SKL_LSQR_PERSIST"] = str(p)
    try:
        yield
    finally:
        if old is None:
            del os.environ["SKL_LSQR_PERSIST"]
        else:
            os.environ["SKL_LSQR_PERSIST"] = old


def _preconditioned(seed, n=5000, d=60):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, d)) * np.logspace(0, -3, d)
    b = rng.standard_normal(n)
    # a rough preconditioner so that LSQR takes tens of iterations
    R = np.linalg.qr(A[::7], mode="r")
    return A, b, R, sl.solve_triangular(R, np.eye(d))


@pytest.mark.parametrize("P", [1, 79, 128])
def test_persist_steps_match_oracle(skl, O, P):
    A, b, R, M = _preconditioned(12)
    with persist(P):
        y, flag, it, steps = skl.lsqr_dense(A, M, b, 1e-12, 500, history=True)
    yo, flo, ito, so, _ = O.lsqr_dense(A, R, b, 1e-12, 500, hist=True)
    assert flag == flo == 0 and abs(it - ito) <= 1
    k = min(it, ito) - 1
    assert rel(steps[:k, :4], so[:k, :4]) < 1e-9
    assert rel(y, yo) < 1e-9
    with persist(0):
        y0, f0, it0, s0 = skl.lsqr_dense(A, M, b, 1e-12, 500, history=True)
    assert f0 == flag and abs(it0 - it) <= 1
    assert rel(y, y0) < 1e-10


def test_persist_repeatable(skl):
    A, b, R, M = _preconditioned(14)
    with persist(32):
        y1, f1, it1, s1 = skl.lsqr_dense(A, M, b, 1e-12, 500, history=True)
        y2, f2, it2, s2 = skl.lsqr_dense(A, M, b, 1e-12, 500, history=True)
    assert it1 == it2 and f1 == f2
    assert np.array_equal(y1, y2) and np.array_equal(s1, s2)


def test_persist_permuted_columns(skl, O):
    """E selects 30 of 40 columns in a permuted order (the rank-deficient
    CPQR case): z is scattered through E and t gathered through it."""
    rng = np.random.default_rng(13)
    n, d = 3000, 40
    A = rng.standard_normal((n, d))
    b = rng.standard_normal(n)
    E = rng.permutation(d)[:30]
    R = np.linalg.qr(A[:, E], mode="r")
    M = sl.solve_triangular(R, np.eye(30))
    yo, *_ = O.lsqr_dense(A, R, b, 1e-12, 200, E=E)
    for P in (3, 47):
        with persist(P):
            y, flag, it, _ = skl.lsqr_dense(A, M, b, 1e-12, 200, E=E)
        assert rel(y, yo) < 1e-10


def test_persist_iteration_limit_and_zero_rhs(skl, O):
    A, b, R, M = _preconditioned(15)
    with persist(16):
        y, flag, it, steps = skl.lsqr_dense(A, M, b, 1e-14, 7, history=True)
        assert flag == 1 and it == 7 and steps.shape == (7, 5)
        yo, flo, ito, so, _ = O.lsqr_dense(A, R, b, 1e-14, 7, hist=True)
        assert flo == 1 and ito == 7
        assert rel(y, yo) < 1e-10 and rel(steps[:, :4], so[:, :4]) < 1e-10
        # b = 0: the initial step ends the solve before the persistent launch
        y, flag, it, _ = skl.lsqr_dense(A, M, np.zeros_like(b), 1e-12, 50)
        assert it == 0 and flag == 0 and not np.any(y)


@pytest.mark.parametrize("P", [4, 63])
def test_persist_dense_solver_matches_oracle(skl, O, P):
    """The whole dense solver on the 4000x200 configuration, rank-deficient
    too (column 57 a combination of columns 11 and 3: CPQR rank 199, E a
    proper subset; no two columns are equal, so no pivot is a tie decided by
    rounding)."""
    A, b = c1_problem(3)
    kw = dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, abs_tol=1e-8)
    xo, io, _ = O.ls_dense(A, b, seed=5, **kw)
    with persist(P):
        x, info = skl.dense_solve(A, b, cpqr=True, seed=5, **kw)
    assert info.flag == io.flag == 0 and abs(info.it - io.it) <= 2
    assert rel(x, xo) < 1e-8
    Ad = A.copy()
    Ad[:, 57] = 2.0 * Ad[:, 11] - 0.5 * Ad[:, 3]
    xo, io, _ = O.ls_dense(Ad, b, seed=5, **kw)
    with persist(P):
        x, info = skl.dense_solve(Ad, b, cpqr=True, seed=5, **kw)
    assert info.rank == io.rank == 199
    assert info.flag == io.flag == 0 and abs(info.it - io.it) <= 2
    assert rel(x, xo) < 1e-8


@pytest.mark.parametrize("P", [4, 47])
def test_persist_sparse_solver_matches_oracle(skl, O, P):
    """ls_sparse_spqr at 3000x100 (CSR rows for A z, CSC columns for A^T u),
    with the LSQR history against the launch-per-kernel iteration's."""
    A, b = sparse_problem(4, n=3000, d=100, density=0.02, decades=8)
    kw = dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)
    xo, io, _ = O.ls_sparse(A, b, **kw)
    with persist(P):
        x, info, h = skl.sparse_solve(A, b, history=200, **kw)
    assert info.flag == io.flag == 0 and abs(info.it - io.it) <= 2
    assert rel(x, xo) < 1e-8
    with persist(0):
        x0, info0, h0 = skl.sparse_solve(A, b, history=200, **kw)
    assert abs(info0.it - info.it) <= 1
    # the first steps only: with the sketch's preconditioner the residual
    # reaches the least-squares minimum within a few steps, after which the
    # Lanczos vectors lose orthogonality and alpha and beta follow their
    # rounding, which differs between summation orders (x does not)
    assert rel(h[:4, :4], h0[:4, :4]) < 1e-9
    assert rel(x, x0) < 1e-10
