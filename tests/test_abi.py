"""The C-ABI boundary without a GPU: the library loads, exports every symbol
include/skl.h declares plus the three ski-lls.h entry points, the host-side
pattern generator is bit-exact with the oracle, and the C++ drop-in header
compiles against a caller written like the reference's examples."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "ski-lls_amd", "libskl.so")


def declared_symbols():
    with open(os.path.join(ROOT, "include", "skl.h")) as fh:
        txt = fh.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(skl_[a-z_0-9]+)\s*\(", txt)))


def exported():
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB], text=True)
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_library_exports_every_declared_symbol():
    syms = exported()
    decl = declared_symbols()
    assert len(decl) >= 20
    missing = [s for s in decl if s not in syms]
    assert not missing, missing


def test_shim_exports_reference_entry_points():
    syms = exported()
    demangled = subprocess.check_output(["c++filt"], input="\n".join(syms), text=True)
    for name in ["ls_dense_hashing_blendenpik(Mat<double>&, Vec<double>&, Vec<double>&, long&, "
                 "int&, long&, double, long, double, double, double, long, int, int)",
                 "ls_dense_hashing_blendenpik_noCPQR(Mat<double>&, Vec<double>&, Vec<double>&, "
                 "long&, int&, long&, double, long, double, long, int, int)",
                 "ls_dense_blendenpik(Mat<double>&, Vec<double>&, Vec<double>&, long&, int&, "
                 "long&, double, double, long, int, int)",
                 "ls_sparse_spqr(cs_dl_sparse&, Vec<double>&, Vec<double>&, long&, int&, long&, "
                 "double, long, double, int, double, long, double, double, int)"]:
        assert name in demangled, name


def test_python_binding_lists_all_exports(skl):
    assert sorted(skl.EXPORTS) == declared_symbols()


def test_default_opts_match_reference_macros(skl):
    o = skl.default_opts()
    # include/bench_config.hpp:7-20
    assert o.it_tol == 1e-6 and o.max_it == 10000 and o.gamma == 1.1 and o.k == 3
    assert o.perturb == 1e-10 and o.rcond_threshold == 1e-10 and o.abs_tol == 1e-8
    assert o.ordering == 2 and o.precision == skl.SKL_FP64


def test_error_strings(skl):
    L = skl.lib()
    assert L.skl_error_string(1).decode() == "Matrix is not overdetermined."
    assert L.skl_error_string(2).decode() == "SPQR Factorization fails"


@pytest.mark.parametrize("seed,m,n,k,dense", [(5, 340, 4000, 1, True), (1, 7500, 120000, 3, False),
                                              (77, 50, 300, 3, True), (3, 8, 100, 8, False),
                                              (9, 60, 200003, 2, True), (4, 900, 70001, 1, False)])
def test_host_pattern_bit_exact_with_oracle(skl, O, seed, m, n, k, dense):
    D, rows, vals = skl.hash_pattern(seed, m, n, k, dense)
    if dense:
        Do, ro, vo = O.dense_pattern(seed, n, m, k)
        assert np.array_equal(D, Do)
    else:
        ro, vo = O.sparse_pattern(seed, n, m, k)
    assert np.array_equal(rows, ro) and np.array_equal(vals, vo)


def test_drivers_built_and_report_usage():
    bindir = os.path.join(ROOT, "ski-lls_amd", "drivers", "bin")
    for d in ["ls_blendenpik_hashing", "ls_blendenpik_hashing_sparse_input", "ls_qr"]:
        out = subprocess.run([os.path.join(bindir, d)], capture_output=True, text=True)
        assert out.returncode == 2 and "usage: " + d in out.stderr


def test_cpp_example_compiles(tmp_path):
    """A caller written against the reference's ski-lls.h API builds and links."""
    src = tmp_path / "ex.cpp"
    src.write_text(r'''
#include <iostream>
#include "ski-lls.h"
int main() {
    long m = 4, n = 2;
    double data_A[8] = {0.306051,-1.53078,1.64493,-1.61322,-0.2829,0.474476,-0.586278,-0.610202};
    double data_b[4] = {0.0649338,0.845946,-0.0164085,0.247119};
    Mat_d A(m, n, data_A); Vec_d b(m, data_b); Vec_d x(A.n());
    long it, rank; int flag;
    ls_dense_hashing_blendenpik(A, b, x, rank, flag, it, OVER_SAMPLING_RATIO, NNZ_PER_COLUMN,
                                ABS_TOL, 1e-10, IT_TOL, MAX_IT, 0, 0);
    A.mv('n', 1, x, -1, b);
    std::cout << x << std::endl;
    long col[4] = {0, 1, 3, 5}; long row[5] = {0, 0, 2, 1, 3}; double val[5] = {2, 3, 1, 4, 5};
    cs_dl S = {5, 4, 3, col, row, val, -1};
    Vec_d xs(3);
    ls_sparse_spqr(S, b, xs, rank, flag, it, OVER_SAMPLING_RATIO, NNZ_PER_COLUMN);
    return 0;
}
''')
    # test/ls_blendenpik.cpp reaches ls_dense_blendenpik through blendenpik.hpp
    src2 = tmp_path / "ex2.cpp"
    src2.write_text(r'''
#include "blendenpik.hpp"
int main() {
    double a[8] = {1, 2, 3, 4, 5, 6, 7, 9}, bb[4] = {1, 1, 1, 1};
    Mat_d A(4, 2, a); Vec_d b(4, bb); Vec_d x(2);
    long it, rank; int flag;
    ls_dense_blendenpik(A, b, x, rank, flag, it, 2.0, IT_TOL, MAX_IT, 0, WISDOM);
    return 0;
}
''')
    exe2 = tmp_path / "ex2"
    subprocess.check_call(["g++", "-std=c++11", "-O1", "-I", os.path.join(ROOT, "include"),
                           str(src2), "-o", str(exe2), LIB, "-Wl,-rpath," + os.path.dirname(LIB)])
    assert exe2.exists()
    exe = tmp_path / "ex"
    subprocess.check_call(["g++", "-std=c++11", "-O1", "-I", os.path.join(ROOT, "include"),
                           str(src), "-o", str(exe), LIB, "-Wl,-rpath," + os.path.dirname(LIB)])
    assert exe.exists()


@pytest.mark.parametrize("seed,m,n", [(5, 340, 4000), (1, 6800, 40000), (77, 300, 300)])
def test_host_subsample_bit_exact_with_oracle(skl, O, seed, m, n):
    D, idx = skl.subsample_pattern(seed, m, n)
    Do, io = O.subsample_pattern(seed, n, m)
    assert np.array_equal(D, Do) and np.array_equal(idx, io)


def test_ctypes_structs_match_c_layout(skl, tmp_path):
    """Every field of the ctypes mirrors sits at the offset gcc gives it in
    include/skl.h, and the structs have the same size."""
    structs = {"skl_opts": skl.Opts, "skl_info": skl.Info, "skl_lsqr_step": skl.LsqrStep}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "skl.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = {}
    for line in subprocess.check_output([str(exe)], text=True).splitlines():
        cname, f, v = line.split()
        got[(cname, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
