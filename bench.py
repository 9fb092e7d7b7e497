#!/usr/bin/env python3
"""Seconds per least-squares solve of the sketch-and-precondition solver
(ls_dense_hashing_blendenpik) on one MI355X, inputs resident in HBM.

Workload (default): dense 40000 x 4000, A = diag(w) G diag(logspace(0,-6,4000))
with G iid N(0,1) and w_i lognormal(0,2); b = A x* + 1e-8 N(0,1); hashing
sketch s=1 with m = 6800 rows after the randomized DHT; LSQR tol 1e-10.
One step = one full solve (pattern, sketch, CPQR, x0, LSQR, back-transform).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import importlib.util
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# MI355X published peaks: HBM3E bandwidth, dense FP64 and FP32 matrix rates
HBM_PEAK_GBS = 8000.0
MFMA_F64_TFLOPS = 78.6
MFMA_F32_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_skl():
    spec = importlib.util.spec_from_file_location("skl", os.path.join(ROOT, "ski-lls_amd", "skl.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.lib()
    return mod


WORKLOADS = {
    "dense_40000x4000": dict(kind="dense", n=40000, d=4000, decades=6, lognormal=2.0, gamma=1.7,
                             k=1),
    "dense_4000x200": dict(kind="dense", n=4000, d=200, decades=4, lognormal=0.0, gamma=1.7, k=1),
    "sparse_120000x5000": dict(kind="sparse", n=120000, d=5000, decades=8, density=0.01,
                               gamma=1.5, k=3),
    # request stream: d log-uniform in [20, 2000], n log-uniform in [10 d, 200000],
    # half dense (4000x200 family), half CSC at density U[0.5%, 2%] (sparse family)
    "serving_stream": dict(kind="serving", dmin=20, dmax=2000, nmax=200000, pool=32,
                           max_dense_elems=120_000_000),
}


def make_dense(torch, W, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n, d = W["n"], W["d"]
    cs = torch.logspace(0, -W["decades"], d, dtype=torch.float64, device="cuda")
    # column-major A (n x d) is the row-major (d x n) tensor At
    At = torch.randn(d, n, generator=g, dtype=torch.float64, device="cuda")
    At *= cs[:, None]
    if W["lognormal"] > 0:
        w = torch.exp(W["lognormal"] * torch.randn(n, generator=g, dtype=torch.float64,
                                                   device="cuda"))
        At *= w[None, :]
    xs = torch.randn(d, generator=g, dtype=torch.float64, device="cuda")
    b = At.t() @ xs + 1e-8 * torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
    return At.contiguous(), b.contiguous()


def make_sparse_host(W, seed, decades=None, diag_first=False):
    """CSC A = S diag(s) + I (diag_first: (S + I) diag(s)), s = logspace(0,-decades,d)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    n, d = W["n"], W["d"]
    A = sp.random(n, d, density=W["density"], random_state=rng, format="csc",
                  data_rvs=rng.standard_normal)
    scale = sp.diags(np.logspace(0, -(W["decades"] if decades is None else decades), d))
    if diag_first:
        A = sp.csc_matrix((A + sp.eye(n, d, format="csc")) @ scale)
    else:
        A = sp.csc_matrix(A @ scale + sp.eye(n, d, format="csc"))
    A.sort_indices()
    xs = rng.standard_normal(d)
    b = A @ xs + 1e-8 * rng.standard_normal(n)
    return A, b


SKIP_REFERENCE = False


def reference_solution(torch, At, b):
    """x = R^{-1} Q^T b from a dense Householder QR of A (A = At^T)."""
    if SKIP_REFERENCE:
        return None
    try:
        Q, R = torch.linalg.qr(At.t(), mode="reduced")
        y = Q.t() @ b
        del Q
        x = torch.linalg.solve_triangular(R, y[:, None], upper=True)[:, 0]
        torch.cuda.synchronize()
        return x
    except Exception as e:  # reported as null
        log(f"reference QR solve failed: {e}")
        return None


# ---------------------------------------------------------------------------
# Accuracy: the relative normal-equation residual ||A^T r|| / (||A||_2 ||r||)
# recomputed on the host with r = b - A x and A^T r accumulated in extended
# precision (x86 long double, 64-bit significand), so that the rounding of r
# itself does not set the floor; the same for the dense-QR solution beside it.
# ---------------------------------------------------------------------------
def norm2_power(matvec, rmatvec, d, iters=30):
    v = np.random.default_rng(0).standard_normal(d)
    v /= np.linalg.norm(v)
    s = 0.0
    for _ in range(iters):
        w = rmatvec(matvec(v))
        s = np.linalg.norm(w)
        v = w / s
    return math.sqrt(s)


def ne_residual_dense(A, b, xs):
    """A: n x d float64 (any layout); xs: name -> x (host).  -> name -> value."""
    A_ld = np.asarray(A, dtype=np.longdouble)
    b_ld = np.asarray(b, dtype=np.longdouble)
    nA = norm2_power(lambda v: A @ v, lambda u: A.T @ u, A.shape[1])
    out = {}
    for name, x in xs.items():
        if x is None:
            out[name] = None
            continue
        r = b_ld - A_ld @ np.asarray(x, dtype=np.longdouble)
        g = (A_ld.T @ r).astype(np.float64)
        out[name] = float(np.linalg.norm(g) / (nA * np.linalg.norm(r.astype(np.float64))))
    return out, nA


def ne_residual_sparse(A, b, xs):
    """A: scipy CSC; extended-precision r = b - A x and A^T r per column."""
    data = A.data.astype(np.longdouble)
    idx, ptr = A.indices, A.indptr
    cols = np.repeat(np.arange(A.shape[1]), np.diff(ptr))
    nA = norm2_power(lambda v: A @ v, lambda u: A.T @ u, A.shape[1])
    nz = np.diff(ptr) > 0
    out = {}
    for name, x in xs.items():
        if x is None:
            out[name] = None
            continue
        ax = np.zeros(A.shape[0], dtype=np.longdouble)
        np.add.at(ax, idx, data * np.asarray(x, dtype=np.longdouble)[cols])
        r = np.asarray(b, dtype=np.longdouble) - ax
        g = np.zeros(A.shape[1], dtype=np.longdouble)
        g[nz] = np.add.reduceat(data * r[idx], ptr[:-1][nz])
        out[name] = float(np.linalg.norm(g.astype(np.float64)) /
                          (nA * np.linalg.norm(r.astype(np.float64))))
    return out, nA


# ---------------------------------------------------------------------------
# Per-phase rooflines of one solve: algorithmic bytes (or flops) of the phase
# over its measured seconds, against the MI355X peak that bounds it.
#   sketch   8 n d + 8 n (A and b read once) + 8 m (d + 1) (sketch written)
#   QR       2 m d^2 - 2/3 d^3 flop (+ 2 (nb + pp) m d for the HQRRP Gaussian
#            sketch when d >= 1000), FP64 matrix peak; mixed: the FP32 Gram
#            m d^2 (symmetric half) against the FP32 matrix peak
#   x0       8 n d (the residual's pass over A)
#   LSQR     per iteration 8 n d + 8 r (r + 1) + 64 (n + d) (one-pass kernel)
# Sparse: 12 B per nonzero for one pass over A, the sketch writes 8 m (d + 1).
# ---------------------------------------------------------------------------
def phase_rooflines(n, d, m, phases, it, nnz=None, mixed=False, hqrrp=False):
    t = phases
    out = {}
    pass_a = 12.0 * nnz if nnz is not None else 8.0 * n * d
    if t.get("t_sketch", 0) > 0:
        by = pass_a + 8.0 * n + 8.0 * m * (d + 1)
        gbs = by / t["t_sketch"] / 1e9
        out["sketch"] = {"bytes": by, "GB/s": gbs, "frac_hbm": gbs / HBM_PEAK_GBS}
    if t.get("t_qr", 0) > 0:
        if mixed:
            fl = m * d * d + d ** 3 / 3.0
            tf = fl / t["t_qr"] / 1e12
            out["qr"] = {"flop": fl, "TFLOP/s": tf, "frac_mfma_f32": tf / MFMA_F32_TFLOPS}
        else:
            fl = 2.0 * m * d * d - 2.0 / 3.0 * d ** 3 + (2.0 * 74 * m * d if hqrrp else 0.0)
            tf = fl / t["t_qr"] / 1e12
            out["qr"] = {"flop": fl, "TFLOP/s": tf, "frac_mfma_f64": tf / MFMA_F64_TFLOPS}
    if t.get("t_x0", 0) > 0:
        gbs = pass_a / t["t_x0"] / 1e9
        out["x0"] = {"bytes": pass_a, "GB/s": gbs, "frac_hbm": gbs / HBM_PEAK_GBS}
    if t.get("t_lsqr", 0) > 0 and it > 0:
        by = it * (pass_a + 8.0 * d * (d + 1) + 64.0 * (n + d))
        gbs = by / t["t_lsqr"] / 1e9
        out["lsqr"] = {"bytes_per_iteration": by / it, "GB/s": gbs, "frac_hbm": gbs / HBM_PEAK_GBS}
    return out


# ---------------------------------------------------------------------------
# CPU baseline: the BLAS/LAPACK restatement oracle/cpu_ref.py on this host,
# at one thread (the reference's configuration) and at all cores.
# ---------------------------------------------------------------------------
def _cpu_ref():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import cpu_ref
    return cpu_ref


def _median3(f):
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        f()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[1]


def cpu_baseline_dense(A, b, W, gpu_iters, budget_s):
    """The 4000 x 200 configuration solved in full (median of 3) at one
    thread and at all cores; the 40000 x 4000 one in full at all cores when
    its estimate fits the budget, and at one thread from a bounded sample:
    the transform + hash of 400 columns (x d/400), dgeqp3 of m x 1000 (x by
    the 2 m c^2 - 2/3 c^3 flop count) and 3 LSQR iterations (x the GPU's
    iteration count)."""
    from threadpoolctl import threadpool_limits
    cr = _cpu_ref()
    cores = cr.host_cores()
    n, d = A.shape
    m = int(math.ceil(d * W["gamma"]))
    kw = dict(gamma=W["gamma"], k=W["k"], abs_tol=1e-8, rcond=1e-12, it_tol=1e-10, seed=1234)
    t_start = time.perf_counter()
    rng = np.random.default_rng(7)
    A1 = np.asfortranarray(rng.standard_normal((4000, 200)) * np.logspace(0, -4, 200))
    b1 = A1 @ rng.standard_normal(200) + 1e-8 * rng.standard_normal(4000)
    c1 = {}
    for label, th in [("1_thread", 1), ("all_cores", cores)]:
        with threadpool_limits(th):
            c1[label] = _median3(lambda: cr.dense_solve(A1, b1, workers=th, **kw))
    # 40000 x 4000: one thread from samples
    flops = lambda c: 2.0 * m * c * c - 2.0 / 3.0 * c ** 3
    cs, cq, nit = min(d, 400), min(d, 1000), 3
    with threadpool_limits(1):
        ts = sorted(cr.dense_solve(np.asfortranarray(A[:, :cs]), b, lsqr_it_cap=0,
                                   **kw)[1]["t_sketch"] for _ in range(3))[1]
        Sq = np.asfortranarray(rng.standard_normal((m, cq)))
        tq = _median3(lambda: __import__("scipy.linalg").linalg.qr(Sq, pivoting=True,
                                                                  mode="economic"))
        Ad = np.asfortranarray(A)
        R = np.triu(rng.standard_normal((d, d))) + 10 * np.eye(d)
        import scipy.linalg as sl
        mv = lambda v: Ad @ sl.solve_triangular(R, v)
        rmv = lambda u: sl.solve_triangular(R, Ad.T @ u, trans="T")
        tl = _median3(lambda: cr.lsqr(mv, rmv, b.copy(), d, 0.0, nit)) / nit
    one = {"sketch": ts * d / cs, "cpqr": tq * flops(d) / flops(cq), "lsqr_per_it": tl}
    one_total = one["sketch"] + one["cpqr"] + tl * (gpu_iters + 1)
    # all cores: the full solve when it fits
    allc = None
    with threadpool_limits(cores):
        # a rough estimate (memory-bound GEMVs scale with a few threads only)
        est = one_total / min(8.0, max(1.0, cores / 2.0))
        if est <= budget_s:
            x, info = cr.dense_solve(A, b, workers=cores, **kw)
            allc = {"sec_per_solve": info["t_total"], "lsqr_iterations": info["it"],
                    "phases_s": {k: info[k] for k in ["t_sketch", "t_qr", "t_x0", "t_lsqr"]},
                    "measured": "full solve, once"}
    spent = time.perf_counter() - t_start
    value, cores_used = (allc["sec_per_solve"], cores) if allc else (one_total, 1)
    return {
        "value": value, "unit": "sec/solve", "cores": cores_used, "kind": "port",
        "sample": ("oracle/cpu_ref.py (numpy/scipy BLAS + LAPACK dgeqp3 in place of HQRRP, "
                   "which the reference's hqrrp/README.md:93-95 puts about 3x faster on one "
                   "core): " + (f"the full {n}x{d} solve at all cores" if allc else
                                "one thread from samples") + f"; {spent:.0f}s of CPU work"),
        "cpu_model": cr.cpu_model(), "nproc": cores,
        "one_thread_sampled": {"sec_per_solve": one_total, "phases_s": one,
                               "sample": f"transform+hash of {cs}/{d} columns (x{d / cs:.0f}), "
                                         f"dgeqp3 of {m}x{cq} (x{flops(d) / flops(cq):.1f} by "
                                         f"flops), {nit} LSQR iterations (x{gpu_iters + 1}), "
                                         "each the median of 3"},
        "all_cores_full": allc,
        "dense_4000x200_full_solve_median_of_3_s": c1,
    }


def cpu_baseline_sparse(A, b, W, gpu_iters):
    """The sparse solve restated on the host (dgeqrf in place of SPQR) at one
    thread from samples: the hashing sketch in full, dgeqrf of m x 1000 (x by
    flops) and 3 LSQR iterations (x the GPU's iteration count)."""
    from threadpoolctl import threadpool_limits
    import scipy.linalg as sl
    cr = _cpu_ref()
    n, d = A.shape
    m = int(math.ceil(d * W["gamma"]))
    rng = np.random.default_rng(8)
    flops = lambda c: 2.0 * m * c * c - 2.0 / 3.0 * c ** 3
    cq, nit = min(d, 1000), 3
    t0 = time.perf_counter()
    with threadpool_limits(1):
        rows, vals = cr.O.sparse_pattern(1, n, m, W["k"])
        S = cr.hashing_matrix(m, n, rows, vals, W["k"])
        ts = _median3(lambda: (S @ A).toarray())
        Sq = np.asfortranarray(rng.standard_normal((m, cq)))
        tq = _median3(lambda: sl.qr(Sq, mode="economic"))
        R = np.triu(rng.standard_normal((d, d))) + 10 * np.eye(d)
        Ar = A.tocsr()
        mv = lambda v: Ar @ sl.solve_triangular(R, v)
        rmv = lambda u: sl.solve_triangular(R, Ar.T @ u, trans="T")
        tl = _median3(lambda: cr.lsqr(mv, rmv, b.copy(), d, 0.0, nit)) / nit
    total = ts + tq * flops(d) / flops(cq) + tl * (gpu_iters + 1)
    return {"value": total, "unit": "sec/solve", "cores": 1, "kind": "port",
            "sample": (f"oracle/cpu_ref.py at one thread: the hashing sketch in full, dgeqrf of "
                       f"{m}x{cq} (x{flops(d) / flops(cq):.1f} by flops; SuiteSparseQR is "
                       f"absent), {nit} LSQR iterations (x{gpu_iters + 1}), each the median of "
                       f"3; {time.perf_counter() - t0:.0f}s of CPU work"),
            "cpu_model": cr.cpu_model(), "nproc": cr.host_cores(),
            "phases_s": {"sketch": ts, "qr": tq * flops(d) / flops(cq), "lsqr_per_it": tl}}


def profile_pass(skl, torch, solve, nsolves=3):
    """Kernel timing of the LSQR operator products over FP64 solves run after
    the timed ones (the event-timed launches are kept out of the timed
    region); one unrecorded solve first builds the timed variant's graphs."""
    skl.profile(True)
    solve(skl.SKL_FP64)
    skl.profile(True, reset=True)
    for _ in range(nsolves):
        solve(skl.SKL_FP64)
    torch.cuda.synchronize()
    skl.profile(False, reset=False)
    return skl.profile_read()


def precision_ablation(args, skl, torch, solve, x, res):
    """SKL_CHOLQR64: the mixed path's algorithm (unpivoted CholeskyQR of the
    column-equilibrated FP64 sketch, the same start refinement and fallbacks)
    with the Gram matrix in FP64, so that mixed vs this isolates what FP32
    arithmetic buys; FP64 vs this is the algorithm swap (pivoted HQRRP or the
    Householder QR against CholeskyQR)."""
    for _ in range(max(1, args.warmup)):
        solve(skl.SKL_CHOLQR64)
    torch.cuda.synchronize()
    steps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(steps):
        info = solve(skl.SKL_CHOLQR64)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    xr = res["fp64"]["x"]
    out = {"sec_per_solve": sec, "it": info.it, "flag": info.flag, "escalated": info.escalated,
           "phases_s": {k: getattr(info, k) for k in
                        ["t_pattern", "t_sketch", "t_qr", "t_x0", "t_lsqr", "t_total"]},
           "x_rel_vs_fp64": ((x - xr).norm() / xr.norm()).item()}
    mixed = res["mixed"]
    out["mixed_speedup_vs_fp64_gram"] = sec / mixed["sec_per_solve"]
    out["factor_phase_speedup_vs_fp64_gram"] = (info.t_qr / mixed["phases"]["t_qr"]
                                                if mixed["phases"]["t_qr"] > 0 else None)
    out["fp64_speedup_from_algorithm"] = res["fp64"]["sec_per_solve"] / sec
    log(f"precision ablation (CholeskyQR, FP64 Gram): {sec * 1e3:.2f} ms/solve, it={info.it}, "
        f"qr {info.t_qr * 1e3:.2f} ms; mixed is {out['mixed_speedup_vs_fp64_gram']:.3f}x faster")
    return out


def run_dense(args, skl, torch, W):
    n, d = W["n"], W["d"]
    log(f"generating {n}x{d} dense problem on the GPU")
    At, b = make_dense(torch, W, seed=2024)
    x = torch.zeros(d, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    kw = dict(gamma=W["gamma"], k=W["k"], it_tol=1e-10, rcond=1e-12, abs_tol=1e-8, seed=1234)

    def solve(precision):
        return skl.dense_solve_dev(n, d, At.data_ptr(), n, b.data_ptr(), x.data_ptr(), cpqr=True,
                                   precision=precision, **kw)

    res = {}
    for prec, name in [(skl.SKL_FP64, "fp64"), (skl.SKL_MIXED, "mixed")]:
        for _ in range(args.warmup):
            solve(prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = []
        lat = []
        for _ in range(args.steps):
            ts = time.perf_counter()
            infos.append(solve(prec))
            lat.append(time.perf_counter() - ts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        info = infos[-1]
        r = b - At.t() @ x
        res[name] = {"sec_per_solve": (t1 - t0) / args.steps, "lat": lat, "it": info.it,
                     "flag": info.flag, "rank": info.rank, "escalated": info.escalated,
                     "precond_eig": [info.precond_lmin, info.precond_lmax],
                     "lsqr_est": info.lsqr_nrm_ar / (info.lsqr_nrm_a * info.lsqr_nrm_r),
                     "r_norm": r.norm().item(),
                     "phases": {k: getattr(info, k) for k in
                                ["t_pattern", "t_sketch", "t_qr", "t_x0", "t_lsqr", "t_total"]},
                     "x": x.clone()}
        log(f"{name}: {res[name]['sec_per_solve'] * 1e3:.2f} ms/solve, it={info.it}, "
            f"flag={info.flag}, phases={res[name]['phases']}")
        log(f"{name} per step (wall ms, in-solver ms): "
            f"{[(round(a * 1e3, 1), round(b.t_total * 1e3, 1)) for a, b in zip(lat, infos)]}")
    res["cholqr64"] = precision_ablation(args, skl, torch, solve, x, res)
    res["prof"] = profile_pass(skl, torch, solve)
    # reference solution by Householder QR (rocSOLVER through torch).  For this
    # row-weighted problem the recomputed ||A^T r|| / (||A|| ||r||) sits near
    # 1e-5 even for LAPACK's own least-squares solution (rounding in r and the
    # row scaling set its floor), so accuracy is judged against the QR solution.
    xq = reference_solution(torch, At, b)
    res["xq"] = xq
    for name in ["fp64", "mixed"]:
        R = res[name]
        R["x_rel_vs_qr"] = (None if xq is None else
                            ((R["x"] - xq).norm() / xq.norm()).item())
    res["x_rel_mixed_vs_fp64"] = ((res["mixed"]["x"] - res["fp64"]["x"]).norm() /
                                  res["fp64"]["x"].norm()).item()
    if not args.no_other_solvers:
        res["other"] = other_dense_solvers(args, skl, torch, W, At, b, res["fp64"]["x"], xq)
    return At, b, res


def other_dense_solvers(args, skl, torch, W, At, b, x_ref, xq):
    """The two other dense entry points on the same matrix (no sketch-and-solve
    start, LSQR from zero): ls_dense_hashing_blendenpik_noCPQR and
    ls_dense_blendenpik, each in FP64 and mixed precision, one warm-up and two
    timed solves."""
    n, d = W["n"], W["d"]
    x = torch.zeros(d, dtype=torch.float64, device="cuda")
    out = {}
    for label, sampling in [("ls_dense_hashing_blendenpik_noCPQR", "hashing"),
                            ("ls_dense_blendenpik", "uniform")]:
        out[label] = {}
        for prec, pname in [(skl.SKL_FP64, "fp64"), (skl.SKL_MIXED, "mixed")]:
            def solve():
                return skl.dense_solve_dev(n, d, At.data_ptr(), n, b.data_ptr(), x.data_ptr(),
                                           cpqr=False, sampling=sampling, precision=prec,
                                           gamma=W["gamma"], k=W["k"], it_tol=1e-10, seed=1234)
            solve()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(2):
                info = solve()
            torch.cuda.synchronize()
            sec = (time.perf_counter() - t0) / 2
            ref = xq if xq is not None else x_ref
            out[label][pname] = {"sec_per_solve": sec, "it": info.it, "flag": info.flag,
                                 "escalated": info.escalated,
                                 "x_rel_vs_qr": ((x - ref).norm() / ref.norm()).item()}
            log(f"{label} {pname}: {sec * 1e3:.2f} ms/solve, it={info.it}, flag={info.flag}")
    return out


def run_sparse(args, skl, torch, W):
    n, d = W["n"], W["d"]
    log(f"generating {n}x{d} sparse problem on the host")
    A, bh = make_sparse_host(W, seed=2024)
    nnz = A.nnz
    cp = torch.from_numpy(A.indptr.astype(np.int64)).cuda()
    ri = torch.from_numpy(A.indices.astype(np.int64)).cuda()
    va = torch.from_numpy(A.data.astype(np.float64)).cuda()
    b = torch.from_numpy(bh).cuda()
    x = torch.zeros(d, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    kw = dict(gamma=W["gamma"], k=W["k"], it_tol=1e-10, abs_tol=1e-8, seed=1)

    def solve(precision):
        return skl.sparse_solve_dev(n, d, nnz, cp.data_ptr(), ri.data_ptr(), va.data_ptr(),
                                    b.data_ptr(), x.data_ptr(), precision=precision, **kw)

    res = {}
    for prec, name in [(skl.SKL_FP64, "fp64"), (skl.SKL_MIXED, "mixed")]:
        for _ in range(args.warmup):
            solve(prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos, lat = [], []
        for _ in range(args.steps):
            ts = time.perf_counter()
            infos.append(solve(prec))
            lat.append(time.perf_counter() - ts)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        info = infos[-1]
        res[name] = {"sec_per_solve": (t1 - t0) / args.steps, "lat": lat, "it": info.it,
                     "flag": info.flag, "rank": info.rank, "escalated": info.escalated,
                     "precond_eig": [info.precond_lmin, info.precond_lmax],
                     "lsqr_est": info.lsqr_nrm_ar / (info.lsqr_nrm_a * info.lsqr_nrm_r),
                     "phases": {k: getattr(info, k) for k in
                                ["t_pattern", "t_sketch", "t_qr", "t_x0", "t_lsqr", "t_total"]},
                     "x": x.clone()}
        log(f"{name}: {res[name]['sec_per_solve'] * 1e3:.2f} ms/solve, it={info.it}, "
            f"flag={info.flag}, rank={info.rank}, phases={res[name]['phases']}")
        log(f"{name} per step (wall ms, in-solver ms): "
            f"{[(round(a * 1e3, 1), round(b.t_total * 1e3, 1)) for a, b in zip(lat, infos)]}")
    res["cholqr64"] = precision_ablation(args, skl, torch, solve, x, res)
    res["prof"] = profile_pass(skl, torch, solve)
    # reference solution: dense Householder QR of A on the GPU (4.8 GB in FP64)
    xq = None
    try:
        if SKIP_REFERENCE:
            raise RuntimeError("--no-reference")
        Ad = torch.sparse_csc_tensor(cp, ri, va, size=(n, d), dtype=torch.float64).to_dense()
        xq = reference_solution(torch, Ad.t().contiguous(), b)
        del Ad
    except Exception as e:
        log(f"dense reference skipped: {e}")
    res["xq"] = xq
    for name in ["fp64", "mixed"]:
        R = res[name]
        R["x_rel_vs_qr"] = (None if xq is None else
                            ((R["x"] - xq).norm() / xq.norm()).item())
    res["x_rel_mixed_vs_fp64"] = ((res["mixed"]["x"] - res["fp64"]["x"]).norm() /
                                  res["fp64"]["x"].norm()).item()
    res["host"] = (A, bh)
    if not args.no_escalation_case:
        res["escalation"] = escalation_case(skl, torch, W, kw)
    return res


def escalation_case(skl, torch, W, kw):
    """Mixed-precision solve of the same generator with the unit diagonal added
    before the column scaling logspace(0,-10,d): column norms spread over 1e10,
    which the FP32 factor cannot rank, so the FP64 factorization must run."""
    n, d = W["n"], W["d"]
    log(f"generating the {n}x{d} escalation case (logspace(0,-10), diagonal scaled)")
    A, bh = make_sparse_host(W, seed=2025, decades=10, diag_first=True)
    cp = torch.from_numpy(A.indptr.astype(np.int64)).cuda()
    ri = torch.from_numpy(A.indices.astype(np.int64)).cuda()
    va = torch.from_numpy(A.data.astype(np.float64)).cuda()
    b = torch.from_numpy(bh).cuda()
    out = {}
    xs = {}
    for prec, name in [(skl.SKL_FP64, "fp64"), (skl.SKL_MIXED, "mixed")]:
        x = torch.zeros(d, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = skl.sparse_solve_dev(n, d, A.nnz, cp.data_ptr(), ri.data_ptr(), va.data_ptr(),
                                    b.data_ptr(), x.data_ptr(), precision=prec, **kw)
        torch.cuda.synchronize()
        xs[name] = x
        out[name] = {"sec": time.perf_counter() - t0, "it": info.it, "flag": info.flag,
                     "escalated": info.escalated,
                     "lsqr_estimate": info.lsqr_nrm_ar / (info.lsqr_nrm_a * info.lsqr_nrm_r)}
        log(f"escalation case {name}: {out[name]}")
    out["x_rel_mixed_vs_fp64"] = ((xs["mixed"] - xs["fp64"]).norm() / xs["fp64"].norm()).item()
    out["generator"] = f"(S + I) diag(logspace(0,-10,{d})), S as the workload"
    return out


def make_pool(W, seed):
    """Distinct problems the request stream draws from (inputs generated once,
    host resident, dense in column-major order)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    pool = []
    for i in range(W["pool"]):
        d = int(round(math.exp(rng.uniform(math.log(W["dmin"]), math.log(W["dmax"])))))
        n = int(round(math.exp(rng.uniform(math.log(10 * d), math.log(W["nmax"])))))
        if i % 2 == 0:
            n = min(n, max(10 * d, W["max_dense_elems"] // d))
            A = np.asfortranarray(rng.standard_normal((n, d)) * np.logspace(0, -4, d)[None, :])
            xs = rng.standard_normal(d)
            b = A @ xs + 1e-8 * rng.standard_normal(n)
            pool.append(dict(kind="dense", n=n, d=d, A=A, b=b,
                             opts=dict(gamma=1.7, k=1, it_tol=1e-10, rcond=1e-12, seed=1000 + i)))
        else:
            dens = rng.uniform(0.005, 0.02)
            A = sp.random(n, d, density=dens, random_state=rng, format="csc",
                          data_rvs=rng.standard_normal)
            A = sp.csc_matrix(A @ sp.diags(np.logspace(0, -8, d)) + sp.eye(n, d, format="csc"))
            A.sort_indices()
            xs = rng.standard_normal(d)
            b = A @ xs + 1e-8 * rng.standard_normal(n)
            pool.append(dict(kind="sparse", n=n, d=d, A=A, b=b,
                             opts=dict(gamma=1.5, k=3, it_tol=1e-10, seed=1)))
    return pool


def run_serving(args, skl, W, requests, rates, resident=False, pool=None):
    """The serving loop (continuous batching over worker slots, FP32
    preconditioner, FP64 LSQR) against sequential plain-FP64 solves of the same
    request stream, both measured: Poisson arrivals at each rate (a multiple of
    one sequential FP64 worker's capacity).  Inputs: host arrays (PCIe
    included), or with `resident` every pool problem uploaded once as a handle
    (skl_upload_dense / skl_upload_csc) that both runs solve on, so only b and
    x cross PCIe.  The sequential run is one thread that takes the requests in
    arrival order and solves each when it has arrived and the previous one is
    done; latency is completion minus arrival in both runs."""
    if pool is None:
        log("generating the request pool")
        pool = make_pool(W, seed=77)
    rng = np.random.default_rng(78)
    which = rng.integers(0, len(pool), size=requests)
    handles = []
    if resident:
        for P in pool:
            P["h"] = skl.Handle(P["A"]) if P["kind"] == "dense" else skl.Handle(Acsc=P["A"])
            handles.append(P["h"])

    def solve(P, precision):
        if resident:
            return P["h"].solve(P["b"], precision=precision, **P["opts"])
        if P["kind"] == "dense":
            return skl.dense_solve(P["A"], P["b"], precision=precision, **P["opts"])
        return skl.sparse_solve(P["A"], P["b"], precision=precision, **P["opts"])

    def submit(S, P, precision):
        if resident:
            return S.submit_handle(P["h"], P["b"], precision=precision, **P["opts"])
        if P["kind"] == "dense":
            return S.submit_dense(P["A"], P["b"], precision=precision, **P["opts"])
        return S.submit_sparse(P["A"], P["b"], precision=precision, **P["opts"])

    log("sequential FP64 solves of the pool (service times, reference x)")
    t_seq, x_seq = [], []
    for P in pool:
        solve(P, skl.SKL_FP64)
        t0 = time.perf_counter()
        x, info = solve(P, skl.SKL_FP64)
        t_seq.append(time.perf_counter() - t0)
        x_seq.append(x)
    svc = np.array([t_seq[w] for w in which])
    capacity = 1.0 / float(svc.mean())
    pct = lambda a, q: float(np.percentile(np.asarray(a), q))
    out = {"requests": requests, "slots": args.slots,
           "sequential_fp64_capacity_solves_per_s": capacity, "rates": []}
    with skl.Server(slots=args.slots) as S:
        for P in pool[:4]:  # warm the workers' code paths
            S.wait(submit(S, P, skl.SKL_MIXED))
        for load in rates:
            lam = load * capacity
            arr = np.cumsum(rng.exponential(1.0 / lam, size=requests))
            arr -= arr[0]
            log(f"rate {lam:.1f}/s ({load} x sequential capacity): sequential FP64")
            lat_seq = []
            t0 = time.perf_counter()
            for i in range(requests):
                now = time.perf_counter() - t0
                if arr[i] > now:
                    time.sleep(arr[i] - now)
                solve(pool[which[i]], skl.SKL_FP64)
                lat_seq.append(time.perf_counter() - t0 - arr[i])
            span_seq = time.perf_counter() - t0
            log("  server, mixed precision")
            tickets = []
            t0 = time.perf_counter()
            for i in range(requests):
                now = time.perf_counter() - t0
                if arr[i] > now:
                    time.sleep(arr[i] - now)
                tickets.append(submit(S, pool[which[i]], skl.SKL_MIXED))
            lat, worst = [], 0.0
            for i, tk in enumerate(tickets):
                x, info, q, r = S.wait(tk)
                lat.append(q + r)
                xr = x_seq[which[i]]
                worst = max(worst, float(np.linalg.norm(x - xr) / np.linalg.norm(xr)))
                if info.flag != 0:
                    raise RuntimeError(f"request {i}: LSQR flag {info.flag}")
            span = time.perf_counter() - t0
            out["rates"].append({
                "load_vs_sequential_capacity": load, "arrival_rate_per_s": lam,
                "server_mixed": {"solves_per_s": requests / span, "p50_latency_s": pct(lat, 50),
                                 "p99_latency_s": pct(lat, 99)},
                "sequential_fp64": {"solves_per_s": requests / span_seq,
                                    "p50_latency_s": pct(lat_seq, 50),
                                    "p99_latency_s": pct(lat_seq, 99)},
                "max_x_rel_server_vs_sequential_fp64": worst})
            r = out["rates"][-1]
            log(f"  server p50 {r['server_mixed']['p50_latency_s'] * 1e3:.1f} ms p99 "
                f"{r['server_mixed']['p99_latency_s'] * 1e3:.1f} ms; sequential p50 "
                f"{r['sequential_fp64']['p50_latency_s'] * 1e3:.1f} ms p99 "
                f"{r['sequential_fp64']['p99_latency_s'] * 1e3:.1f} ms")
        # saturation: the whole stream at once
        t1 = time.perf_counter()
        burst = [submit(S, pool[which[i]], skl.SKL_MIXED) for i in range(requests)]
        for tk in burst:
            S.wait(tk)
        out["saturated_server_mixed_solves_per_s"] = requests / (time.perf_counter() - t1)
    for h in handles:
        h.release()
    for P in pool:
        P.pop("h", None)
    out["inputs"] = ("resident: A uploaded once per problem (skl_upload_*), b and x host"
                     if resident else "host arrays, A copied to the device per request")
    out["pool"] = {"problems": len(pool),
                   "dense": sum(P["kind"] == "dense" for P in pool),
                   "sparse": sum(P["kind"] == "sparse" for P in pool),
                   "largest_dense_entries": max(P["n"] * P["d"] for P in pool
                                                if P["kind"] == "dense"),
                   "largest_sparse_nnz": max(P["A"].nnz for P in pool if P["kind"] == "sparse")}
    return out


def dominant_roofline(res, workload):
    """The LSQR operator product with the largest average launch time, timed
    live with HIP events (profile_pass), against the HBM peak."""
    prof = res["prof"]
    dom = max(prof, key=lambda s: prof[s]["avg_ms"])
    P = prof[dom]
    achieved = P["bytes"] / (P["avg_ms"] * 1e-3) / 1e9 if P["avg_ms"] > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get(workload, {}).get(str(dom))
    return {"bound": "hbm", "kernel": P["name"], "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "avg_launch_ms": P["avg_ms"], "launches_timed": P["launches"],
            "bytes_per_launch": P["bytes"],
            "other_kernels": {v["name"]: {"avg_ms": v["avg_ms"],
                                          "GB/s": (v["bytes"] / (v["avg_ms"] * 1e-3) / 1e9
                                                   if v["avg_ms"] > 0 else 0.0)}
                              for s, v in prof.items() if s != dom and v["launches"] > 0}}


def pct_lat(lat, q):
    lat = sorted(lat)
    return lat[min(len(lat) - 1, max(0, int(math.ceil(q / 100.0 * len(lat))) - 1))]


def solver_summary(res, n, d, m, nnz=None, hqrrp=False):
    f64, mixed = res["fp64"], res["mixed"]
    return {
        "fp64": {"sec_per_solve": f64["sec_per_solve"], "lsqr_iterations": f64["it"],
                 "lsqr_estimate": f64["lsqr_est"], "x_rel_vs_qr": f64["x_rel_vs_qr"],
                 "p50_latency_s": pct_lat(f64["lat"], 50), "p99_latency_s": pct_lat(f64["lat"], 99),
                 "phases_s": f64["phases"],
                 "phase_rooflines": phase_rooflines(n, d, m, f64["phases"], f64["it"], nnz,
                                                    hqrrp=hqrrp)},
        "mixed": {"sec_per_solve": mixed["sec_per_solve"], "lsqr_iterations": mixed["it"],
                  "escalated": mixed["escalated"], "lsqr_estimate": mixed["lsqr_est"],
                  "precond_eig_estimate": mixed["precond_eig"],
                  "x_rel_vs_qr": mixed["x_rel_vs_qr"],
                  "speedup_vs_fp64": f64["sec_per_solve"] / mixed["sec_per_solve"],
                  "x_rel_vs_fp64": res["x_rel_mixed_vs_fp64"], "phases_s": mixed["phases"],
                  "phase_rooflines": phase_rooflines(n, d, m, mixed["phases"], mixed["it"], nnz,
                                                     mixed=True)},
        # the same CholeskyQR with an FP64 Gram: mixed vs this is the precision
        # effect alone, fp64 vs this the algorithm swap
        "precision_ablation_cholqr_fp64_gram": {k: v for k, v in res["cholqr64"].items()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="dense_40000x4000", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0,
                    help="seconds the all-cores CPU solve may take before it is sampled")
    ap.add_argument("--no-escalation-case", action="store_true",
                    help="sparse workload: skip the logspace(0,-10) FP64-fallback check")
    ap.add_argument("--no-other-solvers", action="store_true",
                    help="dense workloads: skip the noCPQR and Blendenpik entry points")
    ap.add_argument("--no-reference", action="store_true",
                    help="skip the dense Householder QR solution that x is compared with "
                         "(x_rel_vs_qr is then null)")
    ap.add_argument("--no-sparse", action="store_true",
                    help="default workload: skip the 120000x5000 sparse configuration")
    ap.add_argument("--no-serving", action="store_true",
                    help="default workload: skip the short serving stream")
    ap.add_argument("--no-residual", action="store_true",
                    help="skip the extended-precision normal-equation residuals")
    ap.add_argument("--requests", type=int, default=2000, help="serving: requests in the stream")
    ap.add_argument("--autotune", action="store_true",
                    help="serving: tune the LSQR launch shapes of the pool's buckets first "
                         "(both runs then use them)")
    ap.add_argument("--inputs", default="both", choices=["host", "resident", "both"],
                    help="serving: host arrays per request, problems resident on the device "
                         "(handles), or both runs")
    ap.add_argument("--slots", type=int, default=4, help="serving: worker slots")
    ap.add_argument("--rates", default="0.5,0.8,1.2",
                    help="serving: arrival rates as multiples of one sequential FP64 worker's "
                         "capacity")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the measurement, copy 1 GiB device-to-device three times "
                         "(reference kernel for scaling HBM byte counters)")
    args = ap.parse_args()
    global SKIP_REFERENCE
    SKIP_REFERENCE = args.no_reference
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the solver has no multi-GPU path (the sketch "
              "would shard by columns and LSQR by rows); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(3)
    skl = load_skl()
    W = WORKLOADS[args.workload]
    metric = "sec/solve to rel. normal-eq residual 1e-10, mixed vs FP64 on GPU; p99 latency"
    if W["kind"] == "serving":
        rates = [float(r) for r in args.rates.split(",")]
        log("generating the request pool")
        pool = make_pool(W, seed=77)
        tuned = None
        if args.autotune:
            # LSQR launch shapes for the pool's buckets, before the stream
            log("autotuning the LSQR launch shapes of the pool's buckets")
            skl.autotune_clear()
            t0 = time.perf_counter()
            per = []
            for P in pool:
                nnz = P["A"].nnz if P["kind"] == "sparse" else None
                best, dflt = skl.autotune_lsqr(P["n"], P["d"], nnz)
                per.append({"kind": P["kind"], "n": P["n"], "d": P["d"], "nnz": nnz,
                            "ms_per_it_tuned": best, "ms_per_it_default": dflt})
            tuned = {"buckets": skl.autotune_entries(), "tuning_s": time.perf_counter() - t0,
                     "per_problem": per,
                     "geomean_speedup_per_iteration": float(np.exp(np.mean(
                         [np.log(q["ms_per_it_default"] / q["ms_per_it_tuned"]) for q in per])))}
            log(f"  {tuned['buckets']} buckets in {tuned['tuning_s']:.1f} s, LSQR iteration "
                f"{tuned['geomean_speedup_per_iteration']:.3f}x faster (geometric mean)")
        if args.inputs in ("host", "both"):
            sv = run_serving(args, skl, W, args.requests, rates, pool=pool)
            if args.inputs == "both":
                log("the same stream with every problem resident on the device (handles)")
                sv["resident_inputs"] = run_serving(args, skl, W, args.requests, rates,
                                                    resident=True, pool=pool)
        else:
            sv = run_serving(args, skl, W, args.requests, rates, resident=True, pool=pool)
        if tuned is not None:
            sv["autotune"] = tuned
        mid = sv["rates"][len(rates) // 2]
        print(json.dumps({
            "metric": metric, "value": mid["server_mixed"]["p99_latency_s"],
            "unit": "sec (p99 latency)", "n_gpus": 1, "steps": sv["requests"], "warmup": 4,
            "ms_per_step": 1e3 / mid["server_mixed"]["solves_per_s"], "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": args.workload, "requests": sv["requests"],
                       "slots": sv["slots"], "generator": "d logU[20,2000], n logU[10d,200000], "
                       "half dense (logspace(0,-4) columns), half CSC density U[0.5%,2%] "
                       "(logspace(0,-8) columns + unit diagonal), Poisson arrivals"},
            "serving": sv, "roofline": None, "cpu_baseline": None}), flush=True)
        skl.lib().skl_finalize()
        return
    sparse = W["kind"] == "sparse"
    out = {"metric": metric, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic"}
    if not sparse:
        n, d = W["n"], W["d"]
        m = int(math.ceil(d * W["gamma"]))
        At, b, res = run_dense(args, skl, torch, W)
        summ = solver_summary(res, n, d, m, hqrrp=d >= 1000)
        f64 = res["fp64"]
        out.update({
            "value": f64["sec_per_solve"], "unit": "sec/solve",
            "ms_per_step": f64["sec_per_solve"] * 1e3,
            "config": {"workload": args.workload, "solver": "ls_dense_hashing_blendenpik",
                       "n": n, "d": d, "m": m, "k": W["k"], "it_tol": 1e-10, "rcond": 1e-12,
                       "generator": f"diag(w) G diag(logspace(0,-{W['decades']},{d})), "
                                    f"w lognormal(0,{W['lognormal']})"},
            "lsqr_iterations": f64["it"], "lsqr_estimate": f64["lsqr_est"],
            "x_rel_vs_qr": f64["x_rel_vs_qr"],
            "p50_latency_s": summ["fp64"]["p50_latency_s"],
            "p99_latency_s": summ["fp64"]["p99_latency_s"],
            "phases_s": f64["phases"], "phase_rooflines": summ["fp64"]["phase_rooflines"],
            "mixed": summ["mixed"],
            "precision_ablation_cholqr_fp64_gram": summ["precision_ablation_cholqr_fp64_gram"],
            **({"other_entry_points": res["other"]} if "other" in res else {}),
            "roofline": dominant_roofline(res, args.workload),
        })
        A_h = At.cpu().numpy().T  # column-major n x d view
        b_h = b.cpu().numpy()
        if not args.no_residual:
            log("normal-equation residuals in extended precision")
            xs = {"fp64": f64["x"].cpu().numpy(), "mixed": res["mixed"]["x"].cpu().numpy(),
                  "dense_qr": None if res["xq"] is None else res["xq"].cpu().numpy()}
            ne, nA = ne_residual_dense(A_h, b_h, xs)
            out["ne_residual"] = {"definition": "||A^T r|| / (||A||_2 ||r||), r = b - A x and "
                                                "A^T r in long double", **ne}
        it64 = f64["it"]
        del At, b, res, f64
        torch.cuda.empty_cache()
        cpu = None
        if not args.no_cpu_baseline:
            log("cpu baseline (oracle/cpu_ref.py)")
            cpu = cpu_baseline_dense(A_h, b_h, W, it64, args.cpu_budget)
            log(f"cpu baseline: {cpu['value']:.2f} s/solve at {cpu['cores']} cores ({cpu['sample']})")
        out["cpu_baseline"] = cpu
        del A_h, b_h
        if not args.no_sparse:
            out["sparse_120000x5000"] = sparse_section(args, skl, torch,
                                                       WORKLOADS["sparse_120000x5000"])
        if not args.no_serving:
            Ws = dict(WORKLOADS["serving_stream"], pool=12, max_dense_elems=20_000_000)
            log("short serving stream")
            out["serving_stream_short"] = run_serving(args, skl, Ws, 100, [0.8])
    else:
        sec = sparse_section(args, skl, torch, W)
        out.update({"value": sec["fp64"]["sec_per_solve"], "unit": "sec/solve",
                    "ms_per_step": sec["fp64"]["sec_per_solve"] * 1e3, "config": sec["config"],
                    "roofline": sec.pop("roofline"), "cpu_baseline": sec.pop("cpu_baseline"),
                    "sparse": sec})
    print(json.dumps(out), flush=True)
    if args.calibrate_copy:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 3 x 1 GiB device copies done")
    skl.lib().skl_finalize()


def sparse_section(args, skl, torch, W):
    """The 120000 x 5000 CSC configuration: FP64 and mixed solves, the
    escalation case, residuals, rooflines and the CPU baseline."""
    n, d = W["n"], W["d"]
    m = int(math.ceil(d * W["gamma"]))
    res = run_sparse(args, skl, torch, W)
    A, bh = res["host"]
    summ = solver_summary(res, n, d, m, nnz=A.nnz)
    sec = {"config": {"workload": "sparse_120000x5000", "solver": "ls_sparse_spqr", "n": n,
                      "d": d, "m": m, "k": W["k"], "nnz": int(A.nnz), "it_tol": 1e-10,
                      "generator": f"CSC density {W['density']}, N(0,1) values, columns x "
                                   f"logspace(0,-{W['decades']},{d}), + unit diagonal"},
           **summ, "roofline": dominant_roofline(res, "sparse_120000x5000")}
    if "escalation" in res:
        sec["mixed"]["escalation_case"] = res["escalation"]
    res.pop("prof", None)
    if not args.no_residual:
        log("sparse normal-equation residuals in extended precision")
        xs = {"fp64": res["fp64"]["x"].cpu().numpy(), "mixed": res["mixed"]["x"].cpu().numpy(),
              "dense_qr": None if res["xq"] is None else res["xq"].cpu().numpy()}
        ne, _ = ne_residual_sparse(A, bh, xs)
        sec["ne_residual"] = {"definition": "||A^T r|| / (||A||_2 ||r||), r = b - A x and "
                                            "A^T r in long double", **ne}
    it = res["fp64"]["it"]
    del res
    torch.cuda.empty_cache()
    sec["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        log("sparse cpu baseline (oracle/cpu_ref.py)")
        sec["cpu_baseline"] = cpu_baseline_sparse(A, bh, W, it)
        log(f"sparse cpu baseline: {sec['cpu_baseline']['value']:.2f} s/solve")
    return sec


if __name__ == "__main__":
    main()
