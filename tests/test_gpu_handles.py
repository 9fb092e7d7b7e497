"""HBM-resident problems (skl_upload_dense / skl_upload_csc, skl_solve_handle,
skl_release) and the server's handle and device-pointer requests.  A solve on
a handle runs the same kernels on the same A, b and seed as the one-shot
solver, so x and the iteration count agree bit for bit."""
import numpy as np
import pytest

from conftest import c1_problem, sparse_problem

pytestmark = pytest.mark.gpu

DENSE_KW = dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=41)
SPARSE_KW = dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)


@pytest.mark.parametrize("precision", [0, 1])
def test_dense_handle_equals_solver(skl, precision):
    """n = 6000 >= 4096: the handle holds the row-major copy that one-pass
    LSQR reads; every dense method on it against its one-shot solver."""
    A, b = c1_problem(3, n=6000, d=300)
    kw = dict(DENSE_KW, precision=precision)
    with skl.Handle(A) as h:
        n, d, nnz, by = h.query()
        assert (n, d, nnz) == (6000, 300, -1)
        assert by >= 2 * 8 * 6000 * 300  # A and its row-major copy
        for method, ref in [(skl.METHOD_DENSE_CPQR, dict(cpqr=True)),
                            (skl.METHOD_DENSE_NOCPQR, dict(cpqr=False)),
                            (skl.METHOD_DENSE_BLENDENPIK, dict(sampling="uniform"))]:
            if method == skl.METHOD_DENSE_BLENDENPIK and precision == 1:
                continue
            xr, ir = skl.dense_solve(A, b, **ref, **kw)
            for rep in range(2):
                x, info = h.solve(b, method=method, **kw)
                assert info.flag == ir.flag == 0
                assert info.it == ir.it and info.rank == ir.rank, (method, info.it, ir.it)
                assert np.array_equal(x, xr), (method, rep)


def test_sparse_handle_equals_solver(skl):
    A, b = sparse_problem(5, n=8000, d=400, density=0.02)
    with skl.Handle(Acsc=A) as h:
        assert h.query()[2] == A.nnz
        for precision in (0, 1):
            xr, ir = skl.sparse_solve(A, b, precision=precision, **SPARSE_KW)
            x, info = h.solve(b, precision=precision, **SPARSE_KW)
            assert info.it == ir.it and info.rank == ir.rank and info.success == ir.success
            assert np.array_equal(x, xr), precision
        # a second right-hand side on the same resident A
        b2 = np.random.default_rng(9).standard_normal(8000)
        xr2, _ = skl.sparse_solve(A, b2, **SPARSE_KW)
        x2, _ = h.solve(b2, **SPARSE_KW)
        assert np.array_equal(x2, xr2)


def test_handle_device_pointers(skl):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible to torch")
    A, b = c1_problem(4, n=5000, d=200)
    xr, ir = skl.dense_solve(A, b, **DENSE_KW)
    At = torch.from_numpy(np.asfortranarray(A).T.copy()).cuda()  # column-major A as rows of At
    bt = torch.from_numpy(b).cuda()
    xt = torch.zeros(200, dtype=torch.float64, device="cuda")
    with skl.Handle(dev=(5000, 200, At.data_ptr(), 5000)) as h:
        del At  # the handle holds its own copy
        torch.cuda.synchronize()
        info = h.solve_dev(bt.data_ptr(), xt.data_ptr(), **DENSE_KW)
        torch.cuda.synchronize()
    assert info.it == ir.it
    assert np.array_equal(xt.cpu().numpy(), xr)


def test_handle_errors(skl):
    A, b = c1_problem(6, n=2000, d=40)
    As, bs = sparse_problem(6, n=2000, d=40)
    with skl.Handle(A) as h, skl.Handle(Acsc=As) as hs:
        with pytest.raises(skl.SklError) as e:
            h.solve(b, method=skl.METHOD_SPARSE_SPQR)
        assert e.value.code == 5
        with pytest.raises(skl.SklError) as e:
            hs.solve(bs, method=skl.METHOD_DENSE_CPQR)
        assert e.value.code == 5
    with pytest.raises(skl.SklError) as e:
        skl.Handle(A[:30])  # n < d
    assert e.value.code == 1
    bad = As.copy()
    bad.indices[3] = 5000  # row index beyond n
    with pytest.raises(skl.SklError) as e:
        skl.Handle(Acsc=bad)
    assert e.value.code == 5


def test_server_handles_and_device_requests(skl):
    """Handle requests and device-pointer requests in one stream, against
    the one-shot solves."""
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible to torch")
    A1, b1 = c1_problem(7, n=6000, d=250)
    A2, b2 = sparse_problem(8, n=7000, d=300, density=0.015)
    b1s = [b1, np.random.default_rng(1).standard_normal(6000)]
    b2s = [b2, np.random.default_rng(2).standard_normal(7000)]
    ref1 = [skl.dense_solve(A1, bb, **DENSE_KW)[0] for bb in b1s]
    ref2 = [skl.sparse_solve(A2, bb, **SPARSE_KW)[0] for bb in b2s]
    At = torch.from_numpy(np.asfortranarray(A1).T.copy()).cuda()
    cp = torch.from_numpy(A2.indptr.astype(np.int64)).cuda()
    ri = torch.from_numpy(A2.indices.astype(np.int64)).cuda()
    va = torch.from_numpy(A2.data.copy()).cuda()
    bt1 = torch.from_numpy(b1).cuda()
    bt2 = torch.from_numpy(b2).cuda()
    xt1 = torch.zeros(250, dtype=torch.float64, device="cuda")
    xt2 = torch.zeros(300, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with skl.Handle(A1) as h1, skl.Handle(Acsc=A2) as h2, skl.Server(slots=3) as S:
        tk = []
        for rep in range(2):
            for i in range(2):
                tk.append(("h1", i, S.submit_handle(h1, b1s[i], **DENSE_KW)))
                tk.append(("h2", i, S.submit_handle(h2, b2s[i], **SPARSE_KW)))
        td1 = S.submit_dense_dev(6000, 250, At.data_ptr(), 6000, bt1.data_ptr(), xt1.data_ptr(),
                                 **DENSE_KW)
        td2 = S.submit_sparse_dev(7000, 300, A2.nnz, cp.data_ptr(), ri.data_ptr(), va.data_ptr(),
                                  bt2.data_ptr(), xt2.data_ptr(), **SPARSE_KW)
        for which, i, t in tk:
            x, info, q, r = S.wait(t)
            ref = (ref1 if which == "h1" else ref2)[i]
            assert np.array_equal(x, ref), (which, i)
        S.wait(td1)
        S.wait(td2)
    torch.cuda.synchronize()
    assert np.array_equal(xt1.cpu().numpy(), ref1[0])
    assert np.array_equal(xt2.cpu().numpy(), ref2[0])
