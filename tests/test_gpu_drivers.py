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


def test_rank_deficient_driver(kat):
    for name in ["n3c4-b1", "ch4-4-b1", "rel5", "lp_modszk1"]:
        r = residual("ls_blendenpik_hashing_sparse_input",
                     os.path.join(GOLDEN, "data", "rank_deficient", name + ".mtx"))
        ref = kat["rank_deficient"][name]
        assert abs(r - ref) < 1e-4 * max(1.0, ref) + 1e-5, (name, r, ref)


def test_test_solvers_script():
    out = subprocess.run([os.path.join(ROOT, "ski-lls_amd", "drivers", "test_solvers.sh")],
                         capture_output=True, text=True, timeout=600,
                         env=dict(os.environ, SKL_SEED="1"))
    assert out.returncode == 0, out.stderr[-2000:]
    lines = [l for l in out.stdout.splitlines() if "gives a solution with residual" in l]
    assert len(lines) == 6 + 12 + 24
    bad = [l for l in lines if not re.search(r"residual \S+$", l)]
    assert not bad, (bad, out.stderr[-2000:])
