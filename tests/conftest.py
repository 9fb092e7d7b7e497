import importlib.util
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(TESTS, "golden")
sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_skl():
    spec = importlib.util.spec_from_file_location(
        "skl", os.path.join(ROOT, "ski-lls_amd", "skl.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _torch_runtime_first():
    """torch ships its own HIP and HSA runtime libraries beside the ROCm ones
    libskl.so links; two runtimes share the process only when torch's starts
    first (once libskl's holds the device, torch reports no GPU).  Tests that
    take device memory from torch therefore need torch initialised before the
    library's first call."""
    try:
        import torch
        torch.cuda.is_available()
    except Exception:
        pass


@pytest.fixture(scope="session")
def skl():
    _torch_runtime_first()
    return load_skl()


@pytest.fixture(scope="session")
def O():
    import oracle_py
    oracle_py.lib()
    return oracle_py


@pytest.fixture(scope="session")
def kat():
    with open(os.path.join(GOLDEN, "kat.json")) as fh:
        return json.load(fh)


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if not (np.iscomplexobj(a) or np.iscomplexobj(b)):
        a = a.astype(np.float64)
        b = b.astype(np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def read_mtx(path):
    """MatrixMarket reader returning a dense float64 array; fat matrices are
    transposed as the reference's drivers do (test/ls_blendenpik_hashing_sparse_input.cpp:43-48)."""
    import scipy.io
    import scipy.sparse
    M = scipy.io.mmread(path)
    M = M.toarray() if scipy.sparse.issparse(M) else np.asarray(M, dtype=np.float64)
    if M.shape[0] < M.shape[1]:
        M = M.T
    return np.asfortranarray(M)


def c1_problem(seed=0, n=4000, d=200):
    """Dense problem of the 4000x200 configuration: iid N(0,1) entries times
    column scaling logspace(0,-4,d); b = A x* + 1e-8 noise."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, d)) * np.logspace(0, -4, d)
    xs = rng.standard_normal(d)
    b = A @ xs + 1e-8 * rng.standard_normal(n)
    return np.asfortranarray(A), b


def sparse_problem(seed=0, n=3000, d=100, density=0.02, decades=8, unit_diag=True):
    """Sparse problem of the 120000x5000 family at reduced size: uniform
    positions, N(0,1) values, columns scaled by logspace(0,-decades,d), unit
    diagonal on the top d x d block, b = A x* + 1e-8 noise.  unit_diag:
    True adds the diagonal after the column scaling (columns keep norm >= 1),
    "scaled" adds it before (A = (S + I) diag(s): full rank, column norms
    spread over the scaling's range), False leaves it out."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    A = sp.random(n, d, density=density, random_state=rng, format="csc",
                  data_rvs=rng.standard_normal)
    scale = sp.diags(np.logspace(0, -decades, d))
    if unit_diag == "scaled":
        A = (A + sp.eye(n, d, format="csc")) @ scale
    else:
        A = A @ scale
        if unit_diag:
            A = A + sp.eye(n, d, format="csc")
    A = sp.csc_matrix(A)
    A.sort_indices()
    xs = rng.standard_normal(d)
    b = A @ xs + 1e-8 * rng.standard_normal(n)
    return A, b


def hqrrp_block_mismatches(O, A, QR, jpvt, tau, nb, pp):
    """Per-block check of an HQRRP factorization against the mathematics it
    rests on (hqrrp/lapack_compatible_sources/NoFLA_HQRRP_WY_blk_var4_skint.c:
    422-568, Downdate_Y :627-718).  After j blocks the downdated sketch of the
    remaining columns is, in exact arithmetic,
        Y_j = (G Q_1 ... Q_j)[:, j:m] (Q_j^T ... Q_1^T A)[j:m, remaining],
    with G the pinned Gaussian (srand(12)) and Q_i the block reflectors the
    factorization stored.  Block j + 1 must pick the first b greedy (dgeqp3)
    pivots of Y_j, and its panel must order them as the greedy QRP of the
    trailing A restricted to them.  Y_j is rebuilt here from the final
    reflectors by numpy/LAPACK, independently of the oracle's running
    downdate.  Returns the blocks that disagree."""
    import scipy.linalg
    m, n = A.shape
    mY = nb + pp
    Acur = np.array(A, dtype=np.float64)              # Q^T A, original column order
    Gcur = O.gaussian_matrix(O.rng(12), mY, m)         # G Q
    jp = np.asarray(jpvt) - 1
    bad = []
    for j in range(0, min(m, n), nb):
        b = min(nb, n - j, m - j)
        last = (j + nb >= m) or (j + nb >= n)
        cand = np.sort(jp[j:])
        if not last:
            Yh = Gcur[:, j:] @ Acur[j:, cand]
            _, P = scipy.linalg.qr(Yh, pivoting=True, mode="r")
            chosen = cand[P[:b]]
            _, Pp = scipy.linalg.qr(Acur[j:, chosen], pivoting=True, mode="r")
            expect = chosen[Pp]
        else:
            _, Pp = scipy.linalg.qr(Acur[j:, cand], pivoting=True, mode="r")
            expect = cand[Pp[:b]]
        if not np.array_equal(expect, jp[j:j + b]):
            bad.append(j // nb)
        # apply this block's reflectors (compact WY, dlarft forward columnwise)
        # to the rows j: of the columns still to be factored, and to G(:, j:)
        V = np.zeros((m - j, b))
        for i in range(b):
            V[i, i] = 1.0
            V[i + 1:, i] = QR[j + i + 1:, j + i]
        T = np.zeros((b, b))
        for i in range(b):
            T[i, i] = tau[j + i]
            T[:i, i] = -tau[j + i] * (T[:i, :i] @ (V[:, :i].T @ V[:, i]))
        rest = np.sort(jp[j + b:])
        if rest.size:
            sub = Acur[j:, rest]
            Acur[j:, rest] = sub - V @ (T.T @ (V.T @ sub))
        Gcur[:, j:] -= (Gcur[:, j:] @ V) @ T @ V.T
    return bad


def c2_problem(seed=0, n=40000, d=4000, lognormal=2.0, decades=6):
    """Dense problem of the 40000x4000 configuration family: A = diag(w) G
    diag(logspace(0,-decades,d)), G iid N(0,1), w_i lognormal(0, lognormal)
    (coherent rows); b = A x* + 1e-8 noise."""
    rng = np.random.default_rng(seed)
    w = np.exp(lognormal * rng.standard_normal(n))
    A = (rng.standard_normal((n, d)) * np.logspace(0, -decades, d)) * w[:, None]
    xs = rng.standard_normal(d)
    b = A @ xs + 1e-8 * rng.standard_normal(n)
    return np.asfortranarray(A), b
