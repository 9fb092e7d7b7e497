"""Generate the committed golden fixtures under tests/golden/.

Run in the build container (it reads the reference's data files):
    python tests/golden/gen_golden.py /root/reference

Writes
  data/*.mtx, data/rank_deficient/*.mtx   the reference's own test matrices
                                          (data/, data/rank_deficient_sparse_paper/)
  kat.json        known answers: the two API examples
                  (test/ski-llsDenseExample.cpp, test/ski-llsSparseExample.cpp),
                  the rank-deficient residuals of test/test_solvers.sh:28-29,
                  and dgelsd residuals (b = ones) of the full-rank matrices
  patterns.json   hashing patterns and sign vectors drawn with the platform's
                  libc rand() (ctypes), in the reference's draw order
                  (src/dense_helpers.cpp:12-14, src/sparse_helpers.cpp:74-116):
                  the first 64 entries and a SHA-256 of the full arrays
  gaussian.json   HQRRP's Gaussian sketch matrix (srand(12), Marsaglia polar,
                  hqrrp/.../NoFLA_HQRRP_WY_blk_var4_skint.c:585-624), first
                  entries and SHA-256
"""
import ctypes
import hashlib
import json
import math
import os
import shutil
import sys

import numpy as np
import scipy.io
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))

# test/test_solvers.sh:26-31
RANK_DEF = [("lp_ship12l", 18.3360), ("Franz1", 26.5028), ("GL7d26", 50.8747),
            ("cis-n4c6-b2", 6.12592e-14), ("lp_modszk1", 33.2357), ("rel5", 14.0199),
            ("ch5-5-b1", 7.31938), ("n3c5-b2", 8.99553e-15), ("ch4-4-b1", 4.23281),
            ("n3c5-b1", 3.46410), ("n3c4-b1", 1.82574), ("cis-n4c6-b3", 30.9959)]


def libc_pattern(libc, seed, m, n, k, dense):
    libc.srand(ctypes.c_uint(seed))
    D = None
    if dense:
        D = np.array([float((libc.rand() % 2) * 2 - 1) for _ in range(n)])
    arr = list(range(m))
    rows = np.empty(n * k, dtype=np.int64)
    for i in range(n):
        for t in range(k):
            j = libc.rand() % (m - t) + t
            arr[t], arr[j] = arr[j], arr[t]
        rows[k * i:k * i + k] = arr[:k]
    sc = math.sqrt(k)
    vals = np.array([1 / sc * float((libc.rand() % 2) * 2 - 1) for _ in range(n * k)])
    return D, rows, vals


def libc_gaussian(libc, rows, cols):
    libc.srand(ctypes.c_uint(12))
    out = np.empty(rows * cols)
    toggle, b2 = 0, 0.0
    for q in range(rows * cols):
        if toggle:
            out[q] = b2
            toggle = 0
            continue
        while True:
            c1 = -1.0 + 2.0 * (libc.rand() / 2147483647)
            c2 = -1.0 + 2.0 * (libc.rand() / 2147483647)
            a = c1 * c1 + c2 * c2
            if not (a == 0 or a >= 1):
                break
        f = math.sqrt((-2 * math.log(a)) / a)
        out[q] = c1 * f
        b2 = c2 * f
        toggle = 1
    return out.reshape(cols, rows).T  # column-major fill order


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main(ref):
    os.makedirs(os.path.join(HERE, "data", "rank_deficient"), exist_ok=True)
    for f in sorted(os.listdir(os.path.join(ref, "data"))):
        if f.endswith(".mtx"):
            shutil.copy(os.path.join(ref, "data", f), os.path.join(HERE, "data", f))
    rd = os.path.join(ref, "data", "rank_deficient_sparse_paper")
    for f in sorted(os.listdir(rd)):
        if f.endswith(".mtx"):
            shutil.copy(os.path.join(rd, f), os.path.join(HERE, "data", "rank_deficient", f))

    kat = {}
    # test/ski-llsDenseExample.cpp:10-14
    A = np.array([0.306051, -1.53078, 1.64493, -1.61322,
                  -0.2829, 0.474476, -0.586278, -0.610202]).reshape(2, 4).T
    b = np.array([0.0649338, 0.845946, -0.0164085, 0.247119])
    x = scipy.linalg.lstsq(A, b, lapack_driver="gelsd")[0]
    kat["dense_example"] = {"A": A.T.ravel().tolist(), "m": 4, "n": 2, "b": b.tolist(),
                            "x": x.tolist(), "residual": float(np.linalg.norm(A @ x - b)),
                            "comment_x": [-0.2122, 0.720]}
    # test/ski-llsSparseExample.cpp:13-20
    col = [0, 1, 3, 5]
    row = [0, 0, 2, 1, 3]
    val = [2.0, 3.0, 1.0, 4.0, 5.0]
    As = scipy.sparse.csc_matrix((val, row, col), shape=(4, 3))
    x = scipy.linalg.lstsq(As.toarray(), b, lapack_driver="gelsd")[0]
    kat["sparse_example"] = {"colptr": col, "rowind": row, "val": val, "m": 4, "n": 3,
                             "b": b.tolist(), "x": x.tolist(),
                             "residual": float(np.linalg.norm(As.toarray() @ x - b)),
                             "comment_x": [0.0571, -0.0164, 0.1127]}
    kat["rank_deficient"] = {name: r for name, r in RANK_DEF}
    full = {}
    for f in sorted(os.listdir(os.path.join(HERE, "data"))):
        if not f.endswith(".mtx"):
            continue
        M = scipy.io.mmread(os.path.join(HERE, "data", f))
        M = M.toarray() if scipy.sparse.issparse(M) else np.asarray(M)
        if M.shape[0] < M.shape[1]:
            M = M.T
        ones = np.ones(M.shape[0])
        xs = scipy.linalg.lstsq(M, ones, lapack_driver="gelsd")[0]
        full[f] = float(np.linalg.norm(M @ xs - ones))
    kat["full_rank_gelsd_residual"] = full
    with open(os.path.join(HERE, "kat.json"), "w") as fh:
        json.dump(kat, fh, indent=1)

    libc = ctypes.CDLL("libc.so.6")
    pats = {}
    for name, seed, m, n, k, dense in [("c1_dense", 1234, 340, 4000, 1, True),
                                       ("c2_dense", 1234, 6800, 40000, 1, True),
                                       ("c3_sparse", 1, 7500, 120000, 3, False),
                                       ("k3_dense_small", 99, 50, 300, 3, True)]:
        D, rows, vals = libc_pattern(libc, seed, m, n, k, dense)
        e = {"seed": seed, "m": m, "n": n, "k": k, "dense": dense,
             "rows64": rows[:64].tolist(), "vals64": vals[:64].tolist(),
             "rows_sha": sha(rows), "vals_sha": sha(vals)}
        if dense:
            e["D64"] = D[:64].tolist()
            e["D_sha"] = sha(D)
        pats[name] = e
    with open(os.path.join(HERE, "patterns.json"), "w") as fh:
        json.dump(pats, fh, indent=1)

    G = libc_gaussian(libc, 74, 1700)
    with open(os.path.join(HERE, "gaussian.json"), "w") as fh:
        json.dump({"rows": 74, "cols": 1700, "first_col": G[:, 0].tolist(),
                   "col1": G[:, 1].tolist(), "sha": sha(np.asfortranarray(G).T)}, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
