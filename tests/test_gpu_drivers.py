"""Command-line drivers (reference test/ls_blendenpik_hashing*.cpp, test/ls_qr.cpp,
test/ls_blendenpik.cpp)
on the reference's fixtures: the printed residual matches the known answer."""
import os
import re
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "ski-lls_amd", "drivers", "bin")


def residual(driver, path, *args):
    out = subprocess.run([os.path.join(BIN, driver), path, *map(str, args)], capture_output=True,
                         text=True, timeout=120, env=dict(os.environ, SKL_SEED="1"))
    assert out.returncode == 0, out.stderr
    m = re.search(r"Residual of \S+ is: (\S+)", out.stdout)
    assert m, out.stdout
    return float(m.group(1))


def test_dense_driver_full_rank(kat):
    for f in ["50by15DenseFullRank.mtx", "120by20DenseFullRank.mtx"]:
        r = residual("ls_blendenpik_hashing", os.path.join(GOLDEN, "data", f))
        ref = kat["full_rank_gelsd_residual"][f]
        assert abs(r - ref) < 1e-5 * ref, (f, r, ref)


def test_blendenpik_driver_full_rank(kat):
    for f in ["50by15DenseFullRank.mtx", "120by20DenseFullRank.mtx"]:
        r = residual("ls_blendenpik", os.path.join(GOLDEN, "data", f))
        ref = kat["full_rank_gelsd_residual"][f]
        assert abs(r - ref) < 1e-5 * ref, (f, r, ref)


def test_sparse_driver_full_rank(kat):
    for f in ["50by20SparseFullRank.mtx", "120by15SparseFullRank.mtx"]:
        r = residual("ls_qr", os.path.join(GOLDEN, "data", f))
        ref = kat["full_rank_gelsd_residual"][f]
        assert abs(r - ref) < 1e-5 * ref, (f, r, ref)


@pytest.mark.parametrize("driver", ["ls_qr", "ls_blendenpik_hashing_sparse_input"])
def test_rank_deficient_driver(kat, driver):
    """The twelve rank-deficient SuiteSparse matrices of test/test_solvers.sh:
    26-31 through ls_sparse_spqr (ls_qr) and through the dense robust solver
    on the densified matrix: the printed residual is the expected one (the
    least-squares minimum, 6 significant digits in the script; the two
    consistent systems have a residual at rounding level)."""
    for name, ref in kat["rank_deficient"].items():
        r = residual(driver, os.path.join(GOLDEN, "data", "rank_deficient", name + ".mtx"))
        assert abs(r - ref) <= 1e-5 * max(1.0, ref) + 1e-10, (driver, name, r, ref)


def test_test_solvers_script(kat):
    """The script's output: every solver line carries a residual, and for the
    rank-deficient set both solvers' residuals equal the expected value the
    script prints before them (test/test_solvers.sh:28-29)."""
    out = subprocess.run([os.path.join(ROOT, "ski-lls_amd", "drivers", "test_solvers.sh")],
                         capture_output=True, text=True, timeout=600,
                         env=dict(os.environ, SKL_SEED="1"))
    assert out.returncode == 0, out.stderr[-2000:]
    lines = [l for l in out.stdout.splitlines() if "gives a solution with residual" in l]
    assert len(lines) == 6 + 12 + 24
    vals = [float(re.search(r"residual (\S+)$", l).group(1)) for l in lines]
    assert all(v >= 0 for v in vals), lines
    blocks = out.stdout.split("TESTING ")
    checked = 0
    for blk in blocks:
        m = re.search(r"We should get (\S+)", blk)
        if not m:
            continue
        ref = float(m.group(1))
        res = [float(x) for x in re.findall(r"gives a solution with residual (\S+)", blk)]
        assert len(res) == 2, blk
        for r in res:
            assert abs(r - ref) <= 1e-5 * max(1.0, ref) + 1e-10, (blk, r, ref)
        checked += 1
    assert checked == 12
