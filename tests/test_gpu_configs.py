"""The benchmark configurations as parity tests, through the C ABI.

* 40000x4000 family (row weights lognormal(0,2), columns logspace(0,-6)):
  at sizes the oracle finishes in seconds, FP64 and mixed x against the
  oracle (north-star tolerance 1e-8 relative, iterations within 2); at the
  full size, size-independent properties: LSQR converged below 1e-10, x
  within 1e-8 of a dense Householder QR solution, mixed within 1e-8 of FP64;
  the HQRRP pivots at the sketch's shape, block by block, against the
  exactly downdated Gaussian sketch.
* 120000x5000 CSC family: the same at reduced and at full size.
* ls_sparse_spqr with an ill-conditioned R11 (success = 0: LSQR on the
  perturbed operator, x0 and back-solve with the unperturbed R11).
"""
import numpy as np
import pytest
import scipy.linalg as sl
from scipy.linalg.lapack import dtrcon

from conftest import c2_problem, hqrrp_block_mismatches, rel, sparse_problem

pytestmark = pytest.mark.gpu

DENSE_KW = dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, abs_tol=1e-8)


@pytest.mark.parametrize("n,d,seed", [(8000, 800, 1), (12000, 1000, 2)])
def test_c2_family_matches_oracle(skl, O, n, d, seed):
    """d = 800 takes the greedy dgeqp3 branch of the CPQR, d = 1000 HQRRP."""
    A, b = c2_problem(seed, n=n, d=d)
    xo, io, _ = O.ls_dense(A, b, seed=17, **DENSE_KW)
    assert io.flag == 0
    x, info = skl.dense_solve(A, b, cpqr=True, seed=17, **DENSE_KW)
    assert info.flag == 0 and info.rank == io.rank == d
    assert abs(info.it - io.it) <= 2, (info.it, io.it)
    assert rel(x, xo) < 1e-8
    xm, im = skl.dense_solve(A, b, cpqr=True, seed=17, precision=skl.SKL_MIXED, **DENSE_KW)
    assert im.flag == 0 and im.rank == d
    assert abs(im.it - io.it) <= 2, (im.it, io.it)
    assert rel(xm, xo) < 1e-8
    est = im.lsqr_nrm_ar / (im.lsqr_nrm_a * im.lsqr_nrm_r)
    assert est < 1e-10


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible to torch")
    return torch


def _qr_solution(torch, At, b):
    Q, R = torch.linalg.qr(At.t(), mode="reduced")
    y = Q.t() @ b
    del Q
    return torch.linalg.solve_triangular(R, y[:, None], upper=True)[:, 0]


def test_c2_full_size_properties(skl):
    """40000x4000: FP64 and mixed solves converge (LSQR estimate <= 1e-10),
    agree with a dense QR solution to 1e-8 and with each other."""
    torch = _torch()
    n, d = 40000, 4000
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    At = torch.randn(d, n, generator=g, dtype=torch.float64, device="cuda")
    At *= torch.logspace(0, -6, d, dtype=torch.float64, device="cuda")[:, None]
    At *= torch.exp(2.0 * torch.randn(n, generator=g, dtype=torch.float64, device="cuda"))[None, :]
    xs = torch.randn(d, generator=g, dtype=torch.float64, device="cuda")
    b = At.t() @ xs + 1e-8 * torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
    out = {}
    for prec in (skl.SKL_FP64, skl.SKL_MIXED):
        x = torch.zeros(d, dtype=torch.float64, device="cuda")
        info = skl.dense_solve_dev(n, d, At.data_ptr(), n, b.data_ptr(), x.data_ptr(), cpqr=True,
                                   precision=prec, seed=99, **DENSE_KW)
        torch.cuda.synchronize()
        assert info.flag == 0 and info.rank == d, (prec, info.flag, info.rank)
        assert info.lsqr_nrm_ar / (info.lsqr_nrm_a * info.lsqr_nrm_r) <= 1e-10
        out[prec] = (x, info)
    xq = _qr_solution(torch, At, b)
    for prec, (x, info) in out.items():
        assert ((x - xq).norm() / xq.norm()).item() < 1e-8, prec
    x64, x32 = out[skl.SKL_FP64][0], out[skl.SKL_MIXED][0]
    assert ((x32 - x64).norm() / x64.norm()).item() < 1e-8
    i64, i32 = out[skl.SKL_FP64][1], out[skl.SKL_MIXED][1]
    print(f"40000x4000: it fp64 {i64.it} mixed {i32.it}; x0 residual {i64.x0_residual:.3e} / "
          f"{i32.x0_residual:.3e}; escalated {i32.escalated}")


def test_hqrrp_pivots_at_c2_sketch_shape(skl):
    """The GPU's HQRRP on a 6800x4000 matrix (the 40000x4000 sketch's shape,
    columns logspace(0,-6)): every block's pivot set and panel order follow
    from the exactly downdated Gaussian sketch, rebuilt by numpy from the
    factorization's own reflectors (no oracle involved)."""
    import oracle_py as O
    rng = np.random.default_rng(12)
    m, n = 6800, 4000
    A = rng.standard_normal((m, n)) * np.logspace(0, -6, n)
    W, jpvt, tau = skl.cpqr(A)
    assert sorted(jpvt) == list(range(1, n + 1))
    assert hqrrp_block_mismatches(O, A, W, jpvt, tau, 64, 10) == []


def test_c3_family_matches_oracle(skl, O):
    A, b = sparse_problem(6, n=24000, d=1000, density=0.01, decades=8)
    kw = dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)
    xo, io, _ = O.ls_sparse(A, b, **kw)
    x, info = skl.sparse_solve(A, b, **kw)
    assert info.flag == io.flag == 0 and info.rank == io.rank == 1000
    assert info.success == io.success == 1
    assert abs(info.it - io.it) <= 2, (info.it, io.it)
    assert rel(x, xo) < 1e-8
    xm, im = skl.sparse_solve(A, b, precision=skl.SKL_MIXED, **kw)
    assert im.flag == 0 and abs(im.it - io.it) <= 2
    assert rel(xm, xo) < 1e-8


def test_c3_full_size_properties(skl):
    """120000x5000 CSC, 1% density, columns logspace(0,-8), unit diagonal:
    FP64 and mixed converge and agree with a dense QR solution to 1e-8."""
    torch = _torch()
    A, b = sparse_problem(2024, n=120000, d=5000, density=0.01, decades=8)
    kw = dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)
    xs = {}
    for prec in (skl.SKL_FP64, skl.SKL_MIXED):
        x, info = skl.sparse_solve(A, b, precision=prec, **kw)
        assert info.flag == 0 and info.rank == 5000 and info.success == 1
        assert info.lsqr_nrm_ar / (info.lsqr_nrm_a * info.lsqr_nrm_r) <= 1e-10
        xs[prec] = x
    At = torch.from_numpy(A.toarray().T.copy()).cuda()
    xq = _qr_solution(torch, At, torch.from_numpy(b).cuda()).cpu().numpy()
    del At
    for prec, x in xs.items():
        assert rel(x, xq) < 1e-8, prec


def test_sparse_ill_conditioned_r11_perturbed_operator(skl, O):
    """rcond(R11) about 4e-12, below rcond_threshold 1e-10: success = 0, so
    LSQR runs on A (R11 + perturb I)^-1 while x0 and the back-solve use R11
    itself (src/ls_sparse_spqr.cpp:66-69,136-144,181-188).  The solution then
    is not the least-squares one: its residual sits above LAPACK's minimum,
    for the oracle and the GPU alike, and x follows the oracle.  The condition
    number is the exact inf-norm one here; LAPACK's dtrcon('I') estimate (the
    reference's, src/cholmod_cond_est.cpp:24-34) gives the same value on this
    R11."""
    A, b = sparse_problem(8, n=4000, d=100, density=0.02, decades=11, unit_diag=False)
    b = b + 1e-3 * np.random.default_rng(1).standard_normal(b.shape[0])
    kw = dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)
    xo, io, _ = O.ls_sparse(A, b, **kw)
    x, info = skl.sparse_solve(A, b, **kw)
    assert info.success == io.success == 0
    assert info.rank == io.rank == 100 and info.flag == io.flag == 0
    assert abs(info.rcond - io.rcond) <= 1e-6 * io.rcond
    # the reference's estimate decides the same way
    m = int(np.ceil(1.5 * 100))
    rows, vals = O.sparse_pattern(1, A.shape[0], m, 3)
    SA = O.sketch_sparse(m, rows, vals, 3, A.indptr.astype(np.int64),
                         A.indices.astype(np.int64), A.data, A.shape[0], 100)
    R, _, _, rank, _ = O.sparse_precond(SA, O.sketch_vec(m, rows, vals, 3, b))
    est = dtrcon(np.triu(R[:rank, :rank]), norm="I", uplo="U", diag="N")[0]
    assert est < 1e-10 and abs(est - io.rcond) <= 0.1 * io.rcond
    assert abs(info.it - io.it) <= 2, (info.it, io.it)
    assert rel(x, xo) < 1e-8
    Ad = A.toarray()
    r_min = np.linalg.norm(Ad @ sl.lstsq(Ad, b, lapack_driver="gelss")[0] - b)
    r, ro = np.linalg.norm(Ad @ x - b), np.linalg.norm(Ad @ xo - b)
    assert r > r_min * (1 + 1e-3) and abs(r - ro) <= 1e-6 * ro


def test_precision_ablation_cholqr64(skl, O):
    """SKL_CHOLQR64 (the mixed solver's CholeskyQR with its Gram in FP64, a
    benchmarking ablation): the same x as the oracle on the 40000x4000 and
    120000x5000 families at reduced size, with the oracle's iteration count."""
    A, b = c2_problem(3, n=8000, d=800)
    xo, io, _ = O.ls_dense(A, b, seed=17, **DENSE_KW)
    x, info = skl.dense_solve(A, b, cpqr=True, seed=17, precision=skl.SKL_CHOLQR64, **DENSE_KW)
    assert info.flag == 0 and info.escalated == 0 and abs(info.it - io.it) <= 2
    assert rel(x, xo) < 1e-8
    As, bs = sparse_problem(7, n=24000, d=1000, density=0.01, decades=8)
    kw = dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)
    xo, io, _ = O.ls_sparse(As, bs, **kw)
    x, info = skl.sparse_solve(As, bs, precision=skl.SKL_CHOLQR64, **kw)
    assert info.flag == 0 and info.escalated == 0 and abs(info.it - io.it) <= 2
    assert rel(x, xo) < 1e-8
