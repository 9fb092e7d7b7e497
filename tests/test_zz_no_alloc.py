"""skl_dense_solve_dev_noalloc (the robust dense solver with every device
buffer carved from the library's static scratch array) against the entry
point it varies, skl_dense_solve_dev, through the C ABI."""
import ctypes

import numpy as np
import pytest

from conftest import c1_problem, c2_problem, rel

pytestmark = pytest.mark.gpu

KW = dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, abs_tol=1e-8, seed=21)


def _free_device_bytes():
    """hipMemGetInfo of the runtime libskl.so itself links (/opt/rocm)."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def test_noalloc_matches_dev_entry_point(skl):
    torch = pytest.importorskip("torch")
    assert torch.cuda.is_available()
    assert skl.lib().skl_init(0) == 0
    # two calls in a row, the second with a smaller input; the first takes
    # the four-step transform (n = 20000) and its column chunks, the second
    # the single-kernel one (n = 3000 <= 4096)
    cases = [(c2_problem(3, n=20000, d=400), 400), (c1_problem(4, n=3000, d=150), 150)]
    for (A, b), d in cases:
        n = A.shape[0]
        At = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()  # column-major A
        bt = torch.from_numpy(b).cuda()
        x_dev = torch.zeros(d, dtype=torch.float64, device="cuda")
        x_na = torch.zeros(d, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        i_dev = skl.dense_solve_dev(n, d, At.data_ptr(), n, bt.data_ptr(), x_dev.data_ptr(),
                                    cpqr=True, **KW)
        i_na = skl.dense_solve_dev_noalloc(n, d, At.data_ptr(), n, bt.data_ptr(), x_na.data_ptr(),
                                           **KW)
        torch.cuda.synchronize()
        assert i_na.flag == i_dev.flag == 0
        assert i_na.rank == i_dev.rank == d
        assert abs(i_na.it - i_dev.it) <= 2, (i_na.it, i_dev.it)
        xd, xn = x_dev.cpu().numpy(), x_na.cpu().numpy()
        assert rel(xn, xd) < 1e-8
        # again: the scratch keeps the last call's contents, the result is the
        # same bit for bit, and the call takes no device memory
        x_again = torch.zeros(d, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        free0 = _free_device_bytes()
        skl.dense_solve_dev_noalloc(n, d, At.data_ptr(), n, bt.data_ptr(), x_again.data_ptr(),
                                    **KW)
        free1 = _free_device_bytes()
        assert np.array_equal(x_again.cpu().numpy(), xn)
        assert free1 == free0, (free0, free1)


def test_noalloc_refuses_what_it_cannot_hold(skl):
    assert skl.lib().skl_init(0) == 0
    with pytest.raises(skl.SklError) as e:
        skl.dense_solve_dev_noalloc(1_000_000, 3000, 0, 1_000_000, 0, 0, **KW)
    assert e.value.code == 6 and "MiB" in str(e.value)
    with pytest.raises(skl.SklError) as e:
        skl.dense_solve_dev_noalloc(4000, 100, 0, 4000, 0, 0, precision=skl.SKL_MIXED, **KW)
    assert e.value.code == 6
    with pytest.raises(RuntimeError, match="Matrix is not overdetermined."):
        skl.dense_solve_dev_noalloc(10, 20, 0, 10, 0, 0, **KW)
