"""End-to-end parity of the three solvers with the oracle, the reference's
known answers, failure handling, mixed precision and full-size properties.

x is compared to the oracle's x (same algorithm, same seed) at 1e-8
relative, the north-star tolerance; LSQR's stopping estimate must satisfy
the reference's criterion (src/IterSolver.cpp:275-278)."""
import os

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

from conftest import GOLDEN, c1_problem, read_mtx, rel, sparse_problem

pytestmark = pytest.mark.gpu


def ne_residual(A, x, b):
    r = b - A @ x
    return np.linalg.norm(A.T @ r) / (np.linalg.norm(A, 2) * np.linalg.norm(r))


@pytest.mark.parametrize("seed", [0, 1])
def test_dense_c1_matches_oracle(skl, O, seed):
    A, b = c1_problem(seed)
    kw = dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, abs_tol=1e-8, seed=1234 + seed)
    x, info = skl.dense_solve(A, b, cpqr=True, **kw)
    xo, io, _ = O.ls_dense(A, b, **kw)
    assert info.flag == 0 and info.rank == io.rank == 200
    assert abs(info.it - io.it) <= 2, (info.it, io.it)
    assert rel(x, xo) < 1e-8
    assert info.lsqr_nrm_ar < 1e-10 * info.lsqr_nrm_a * info.lsqr_nrm_r
    xr = sl.lstsq(A, b, lapack_driver="gelsd")[0]
    assert rel(x, xr) < 1e-8


def test_dense_nocpqr_c1_matches_oracle(skl, O):
    A, b = c1_problem(2)
    kw = dict(gamma=1.7, k=1, it_tol=1e-10, seed=77)
    x, info = skl.dense_solve(A, b, cpqr=False, **kw)
    xo, io, _ = O.ls_dense_nocpqr(A, b, **kw)
    assert info.flag == 0 and info.rank == 200
    assert abs(info.it - io.it) <= 2, (info.it, io.it)
    assert rel(x, xo) < 1e-8


@pytest.mark.parametrize("n,d,gamma", [(4000, 200, 4.0), (40000, 400, 2.0)])
def test_dense_blendenpik_matches_oracle(skl, O, n, d, gamma):
    """ls_dense_blendenpik: rows of the randomized DHT sampled uniformly
    (n = 4000 takes the single-pass transform, n = 40000 the four-step one),
    Householder QR, LSQR from zero."""
    A, b = c1_problem(4, n=n, d=d)
    kw = dict(gamma=gamma, it_tol=1e-10, seed=91)
    x, info = skl.dense_solve(A, b, cpqr=False, sampling="uniform", **kw)
    xo, io, _ = O.ls_dense_blendenpik(A, b, **kw)
    assert info.flag == 0 and info.rank == d and info.m == int(np.ceil(gamma * d))
    assert abs(info.it - io.it) <= 2, (info.it, io.it)
    assert rel(x, xo) < 1e-8
    assert info.lsqr_nrm_ar < 1e-10 * info.lsqr_nrm_a * info.lsqr_nrm_r
    xm, im = skl.dense_solve(A, b, cpqr=False, sampling="uniform", precision=skl.SKL_MIXED, **kw)
    assert im.flag == 0 and im.escalated == 0 and abs(im.it - info.it) <= 3
    assert rel(xm, x) < 1e-8


def test_dense_blendenpik_rejects_more_samples_than_rows(skl):
    A, b = c1_problem(5, n=300, d=200)
    with pytest.raises(RuntimeError, match="row sampling"):
        skl.dense_solve(A, b, cpqr=False, sampling="uniform", gamma=1.7)


def test_dense_history_matches_oracle(skl, O):
    A, b = c1_problem(3, n=3000, d=120)
    kw = dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=5)
    x, info, st = skl.dense_solve(A, b, cpqr=True, history=400, **kw)
    xo, io, so = O.ls_dense(A, b, hist=True, **kw)
    # alpha, beta of the first iterations agree to the rounding of the
    # preconditioner (R from the two QRs differs at ~kappa eps); later ones
    # drift apart as finite-precision Lanczos vectors do, while x still agrees
    assert rel(st[:10, :2], so[:10, :2]) < 1e-5
    assert abs(info.it - io.it) <= 2
    assert rel(x, xo) < 1e-8


def test_sparse_matches_oracle(skl, O):
    A, b = sparse_problem(4, n=6000, d=150, density=0.02)
    kw = dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)
    x, info = skl.sparse_solve(A, b, **kw)
    xo, io, _ = O.ls_sparse(A, b, **kw)
    assert info.flag == 0 and info.rank == io.rank == 150 and info.success == io.success == 1
    assert abs(info.it - io.it) <= 2, (info.it, io.it)
    assert rel(x, xo) < 1e-8


def test_examples_known_answers(skl, kat):
    e = kat["dense_example"]
    A = np.array(e["A"]).reshape(e["n"], e["m"]).T
    x, info = skl.dense_solve(A, np.array(e["b"]), gamma=1.1, k=3, rcond=1e-10, it_tol=1e-6,
                              seed=1)
    assert np.allclose(x, e["x"], atol=1e-6)
    s = kat["sparse_example"]
    As = sp.csc_matrix((s["val"], s["rowind"], s["colptr"]), shape=(s["m"], s["n"]))
    x, info = skl.sparse_solve(As, np.array(s["b"]), gamma=1.1, k=3, it_tol=1e-6)
    assert np.allclose(x, s["x"], atol=1e-6)


def test_full_rank_fixtures(skl, kat):
    for f, r_ref in kat["full_rank_gelsd_residual"].items():
        A = read_mtx(os.path.join(GOLDEN, "data", f))
        b = np.ones(A.shape[0])
        x, info = skl.dense_solve(A, b, gamma=1.7, k=1, abs_tol=1e-8, rcond=1e-12, it_tol=1e-6,
                                  seed=1)
        assert abs(np.linalg.norm(A @ x - b) - r_ref) < 1e-6 * max(1, r_ref), f
        if "Sparse" in f:
            As = sp.csc_matrix(A)
            xs, _ = skl.sparse_solve(As, b, gamma=1.4, k=2, it_tol=1e-6)
            assert abs(np.linalg.norm(A @ xs - b) - r_ref) < 1e-6 * max(1, r_ref), f


def test_rank_deficient_kats(skl, kat):
    """test/test_solvers.sh:28-31 through the dense robust solver on the
    densified matrix, as test/ls_blendenpik_hashing_sparse_input.cpp does."""
    for name, r_ref in kat["rank_deficient"].items():
        A = read_mtx(os.path.join(GOLDEN, "data", "rank_deficient", name + ".mtx"))
        b = np.ones(A.shape[0])
        x, info = skl.dense_solve(A, b, gamma=1.7, k=1, abs_tol=1e-8, rcond=1e-12,
                                  it_tol=1e-6, seed=1)
        r = np.linalg.norm(A @ x - b)
        assert abs(r - r_ref) < 1e-4 * max(1.0, r_ref) + 1e-5, (name, r, r_ref)


def test_errors_and_flags(skl):
    with pytest.raises(RuntimeError, match="Matrix is not overdetermined."):
        skl.dense_solve(np.ones((3, 5)), np.ones(3))
    with pytest.raises(RuntimeError, match="Matrix is not overdetermined."):
        skl.sparse_solve(sp.csc_matrix(np.ones((3, 5))), np.ones(3))
    A, b = c1_problem(5, n=2000, d=50)
    x, info = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, max_it=1, rcond=1e-12)
    assert info.flag == 1 and info.it == 1
    # b in the range of A: the sketch-and-solve start already meets abs_tol
    xs = np.random.default_rng(1).standard_normal(50)
    x, info = skl.dense_solve(A, A @ xs, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12,
                              abs_tol=1e-6)
    assert info.it == 0 and info.flag == 0 and rel(x, xs) < 1e-8
    x, info = skl.dense_solve(A, np.zeros(2000), gamma=1.7, k=1)
    assert not x.any() and info.flag == 0


def test_deterministic(skl):
    A, b = c1_problem(6)
    x1, i1 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, seed=3)
    x2, i2 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, seed=3)
    assert np.array_equal(x1, x2) and i1.it == i2.it


def test_deterministic_sparse_and_after_other_shapes(skl):
    """Repeated solves in one thread agree bitwise, also after solves of other
    shapes have reused the device arena in between."""
    A, b = sparse_problem(20, n=5000, d=120, density=0.02)
    x1, i1 = skl.sparse_solve(A, b, gamma=1.5, k=3, it_tol=1e-10, seed=1)
    A2, b2 = c1_problem(12, n=2500, d=25)
    skl.dense_solve(A2, b2, gamma=1.7, k=1, it_tol=1e-10, seed=3)
    A3, b3 = sparse_problem(21, n=3000, d=40, density=0.02)
    skl.sparse_solve(A3, b3, gamma=1.5, k=3, it_tol=1e-10, seed=1)
    x2, i2 = skl.sparse_solve(A, b, gamma=1.5, k=3, it_tol=1e-10, seed=1)
    x3, i3 = skl.sparse_solve(A, b, gamma=1.5, k=3, it_tol=1e-10, seed=1)
    assert i1.it == i2.it == i3.it, (i1.it, i2.it, i3.it)
    assert np.array_equal(x1, x2) and np.array_equal(x2, x3)


def test_mixed_precision_dense(skl):
    A, b = c1_problem(7)
    x64, i64 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=9)
    x32, i32 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=9,
                               precision=skl.SKL_MIXED)
    assert i32.flag == 0 and i32.escalated == 0
    assert rel(x32, x64) < 1e-8
    assert i32.lsqr_nrm_ar < 1e-10 * i32.lsqr_nrm_a * i32.lsqr_nrm_r


def test_mixed_precision_dense_iterations_match_fp64(skl):
    """The mixed factor (CholeskyQR of the equilibrated sketch) preconditions
    as well as the FP64 CPQR, and the start refined against the FP64 sketch
    leaves LSQR the same work: iteration counts within 2."""
    A, b = c1_problem(17, n=8000, d=400)
    x64, i64 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=2)
    x32, i32 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=2,
                               precision=skl.SKL_MIXED)
    assert i32.escalated == 0 and i32.flag == 0 and i32.rank == 400
    assert abs(i32.it - i64.it) <= 2, (i32.it, i64.it)
    assert rel(x32, x64) < 1e-8


def test_mixed_precision_dense_rank_deficient_falls_back(skl):
    """Duplicated columns make the Gram singular: the Cholesky breaks down and
    the solve runs the FP64 CPQR of the same FP64 sketch, which the FP64 mode
    also runs -- the same x, bit for bit."""
    A, b = c1_problem(18, n=3000, d=120)
    A = np.asfortranarray(np.hstack([A[:, :100], A[:, :20]]))
    x64, i64 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=4)
    x32, i32 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=4,
                               precision=skl.SKL_MIXED)
    assert i32.escalated == 1 and i32.rank == i64.rank and i64.rank < 120
    assert np.array_equal(x32, x64)


def correlated_problem(seed, n=20000, d=400, decades=3, c=0.5, dep=None):
    """A = G diag(logspace(0,-decades)) U with U unit upper triangular: the
    column-equilibrated sketch has condition number about 2e3 for the
    defaults.  dep = (j, a, b): column j becomes 1e-3 A_a + A_b."""
    rng = np.random.default_rng(seed)
    U = np.eye(d) + c * np.triu(rng.standard_normal((d, d)) / np.sqrt(d), 1)
    A = (rng.standard_normal((n, d)) * np.logspace(0, -decades, d)) @ U
    if dep is not None:
        j, a, bb = dep
        A[:, j] = 1e-3 * A[:, a] + A[:, bb]
    b = A @ rng.standard_normal(d) + 1e-8 * rng.standard_normal(n)
    return np.asfortranarray(A), b


def test_mixed_precision_probe_keeps_ill_conditioned_factor(skl):
    """The FP32 factor's smallest equilibrated pivot (about 5e-3) is below
    sqrt(m u_32), where a dependent column would also land, so the Lanczos
    estimate of SA M runs; it finds SA M well conditioned and the FP32
    factor is kept."""
    A, b = correlated_problem(21)
    x64, i64 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=3)
    x32, i32 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=3,
                               precision=skl.SKL_MIXED)
    assert i64.precond_lmin == 0.0 and i64.precond_lmax == 0.0
    assert i32.escalated == 0 and i32.rank == 400, (i32.escalated, i32.cond_est)
    assert 0.1 < i32.precond_lmin <= i32.precond_lmax < 10.0, (i32.precond_lmin, i32.precond_lmax)
    assert abs(i32.it - i64.it) <= 3, (i32.it, i64.it)
    assert rel(x32, x64) < 1e-8


def test_mixed_precision_dependent_column_falls_back(skl):
    """Column 200 = 1e-3 A_3 + A_7: an FP32 Gram can leave this dependence a
    pivot near sqrt(u_32) instead of 0 (or break down).  Either way the solve
    falls back to the FP64 CPQR of the same FP64 sketch and returns its rank
    and x, bit for bit."""
    A, b = correlated_problem(22, dep=(200, 3, 7))
    x64, i64 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=5)
    x32, i32 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=5,
                               precision=skl.SKL_MIXED)
    assert i64.rank == 399
    assert i32.escalated == 1 and i32.rank == i64.rank
    assert np.array_equal(x32, x64)


@pytest.mark.parametrize("d", [1, 7, 64, 65, 129])
def test_mixed_precision_small_and_ragged_widths(skl, d):
    """Widths around the 64-column blocks of the Cholesky (a single column, a
    partial first block, exactly one block, one column more, two blocks and
    one more): the mixed factor must give the FP64 solution."""
    A, b = c1_problem(30 + d, n=max(50 * d, 400), d=d)
    x64, i64 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=8)
    x32, i32 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=8,
                               precision=skl.SKL_MIXED)
    assert i32.escalated == 0 and i32.rank == d and i32.flag == 0
    assert rel(x32, x64) < 1e-8


def test_mixed_precision_nocpqr(skl):
    A, b = c1_problem(19, n=4000, d=200)
    x64, i64 = skl.dense_solve(A, b, cpqr=False, gamma=1.7, k=1, it_tol=1e-10, seed=6)
    x32, i32 = skl.dense_solve(A, b, cpqr=False, gamma=1.7, k=1, it_tol=1e-10, seed=6,
                               precision=skl.SKL_MIXED)
    assert i32.escalated == 0 and i32.flag == 0
    assert abs(i32.it - i64.it) <= 3, (i32.it, i64.it)
    assert rel(x32, x64) < 1e-8


@pytest.mark.parametrize("unit_diag", ["scaled", False])
def test_mixed_precision_escalates_on_ill_conditioned_sparse(skl, unit_diag):
    """Columns scaled by logspace(0,-10): column norms spread over 1e10, so the
    FP32 factor cannot make the rank decision that the FP64 tolerance
    20 (m + d) eps max_j ||SA_j|| makes, and the solve falls back to FP64."""
    A, b = sparse_problem(8, n=4000, d=100, density=0.02, decades=10, unit_diag=unit_diag)
    assert np.linalg.cond(A.toarray()) > 1e8
    x, info = skl.sparse_solve(A, b, gamma=1.5, k=3, it_tol=1e-10, precision=skl.SKL_MIXED)
    assert info.escalated == 1 and info.flag == 0
    x64, _ = skl.sparse_solve(A, b, gamma=1.5, k=3, it_tol=1e-10)
    assert rel(x, x64) < 1e-8


def test_mixed_precision_keeps_fp32_when_unit_diagonal_dominates(skl):
    """The same scaling with the unit diagonal added after it leaves every
    column of norm >= 1 and cond(A) small: the FP32 factor preconditions it and
    no fallback is taken."""
    A, b = sparse_problem(8, n=4000, d=100, density=0.02, decades=10, unit_diag=True)
    assert np.linalg.cond(A.toarray()) < 1e3
    x, info = skl.sparse_solve(A, b, gamma=1.5, k=3, it_tol=1e-10, precision=skl.SKL_MIXED)
    assert info.escalated == 0 and info.flag == 0
    x64, _ = skl.sparse_solve(A, b, gamma=1.5, k=3, it_tol=1e-10)
    assert rel(x, x64) < 1e-8


@pytest.mark.parametrize("n,d", [(6000, 300), (5003, 1500), (20000, 2100), (9001, 5001)])
def test_fused_lsqr_pass_matches_two_pass(skl, O, monkeypatch, n, d):
    """The one-pass LSQR kernel (A z and A^T u from one read of a row-major
    copy of A) against the two-pass kernels and the oracle; the widths take
    1, 4, 8 and 16 column pairs per thread (the last one odd, with a padding
    column in the copy), and the row counts leave workgroups with unequal
    row counts."""
    A, b = c1_problem(11, n=n, d=d)
    monkeypatch.setenv("SKL_LSQR_FUSED", "0")
    x2, i2 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=4)
    monkeypatch.setenv("SKL_LSQR_FUSED", "1")
    x1, i1 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=4)
    monkeypatch.delenv("SKL_LSQR_FUSED")
    assert i1.flag == 0 and abs(i1.it - i2.it) <= 2
    assert rel(x1, x2) < 1e-8
    if n * d <= 6000 * 300:
        xo, io, _ = O.ls_dense(A, b, gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=4)
        assert rel(x1, xo) < 1e-8


def test_sparse_rank_deficient_wide_refactorises_with_hqrrp(skl):
    """More than 1000 columns with 50 duplicated ones: the natural-order QR
    meets SPQR's tolerance, and the column-pivoted refactorisation (HQRRP at
    this width) finds the rank.  Any least-squares solution has the residual
    of LAPACK's SVD solution (dgelss; dgelsd's divide and conquer leaves a
    residual 1% above the minimum on this exactly rank-deficient matrix)."""
    A, b = sparse_problem(21, n=5000, d=1100, density=0.01)
    A = sp.hstack([A[:, :1050], A[:, :50]]).tocsc()
    A.sort_indices()
    b = b + 0.1 * np.random.default_rng(3).standard_normal(b.shape[0])
    x, info = skl.sparse_solve(A, b, gamma=1.5, k=3, it_tol=1e-10)
    assert info.rank == 1050 and info.flag == 0
    Ad = A.toarray()
    xl = sl.lstsq(Ad, b, lapack_driver="gelss")[0]
    r, rl = np.linalg.norm(Ad @ x - b), np.linalg.norm(Ad @ xl - b)
    assert abs(r - rl) <= 1e-8 * rl


def test_sparse_long_rows_csr_order(skl):
    """Rows with more entries than a wave has lanes (65, 100, 129 and all 300
    columns) go through the CSR row placement in several 64-entry rounds: the
    solution matches the dense least-squares one and repeats bitwise."""
    import scipy.sparse as sp
    A, b = sparse_problem(31, n=3000, d=300, density=0.01)
    A = A.tolil()
    rng = np.random.default_rng(5)
    for i, cnt in [(17, 65), (500, 100), (1201, 129), (2999, 300)]:
        cols = rng.choice(300, size=cnt, replace=False)
        A[i, cols] = rng.standard_normal(cnt)
    A = A.tocsc()
    A.sort_indices()
    b = A @ rng.standard_normal(300) + 1e-3 * rng.standard_normal(3000)
    x1, i1 = skl.sparse_solve(A, b, gamma=1.5, k=3, it_tol=1e-12, seed=1)
    x2, _ = skl.sparse_solve(A, b, gamma=1.5, k=3, it_tol=1e-12, seed=1)
    xl = np.linalg.lstsq(A.toarray(), b, rcond=None)[0]
    assert i1.flag == 0 and i1.rank == 300
    assert rel(x1, xl) < 1e-8
    assert np.array_equal(x1, x2)


def test_fused_lsqr_rank_deficient_matches_two_pass(skl, monkeypatch):
    """Twenty duplicated columns: CPQR finds rank 280 and LSQR runs on the
    pivoted columns E[0:280] of A (z scattered through E, A^T u gathered
    through it).  The one-pass kernel gives the two-pass x, and the residual
    is LAPACK's least-squares minimum."""
    A, b = c1_problem(41, n=6000, d=280)
    A = np.asfortranarray(np.hstack([A, A[:, :20]]))
    monkeypatch.setenv("SKL_LSQR_FUSED", "0")
    x2, i2 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-12, rcond=1e-10, seed=6)
    monkeypatch.delenv("SKL_LSQR_FUSED")
    x1, i1 = skl.dense_solve(A, b, gamma=1.7, k=1, it_tol=1e-12, rcond=1e-10, seed=6)
    assert i1.rank == i2.rank == 280 and i1.flag == 0
    assert rel(x1, x2) < 1e-8
    r_min = np.linalg.norm(A @ sl.lstsq(A, b, lapack_driver="gelss")[0] - b)
    assert abs(np.linalg.norm(A @ x1 - b) - r_min) < 1e-8 * r_min
