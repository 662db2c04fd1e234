#!/usr/bin/env python3
"""Headline benchmark: FP64-accurate Ax = b solves/s at n = 32768 on one MI355X.

Workload (mixed-precision LU-IR, one right-hand side): A = U diag(sigma) V^T with U, V
Haar-orthogonal (QR of seeded Gaussians with sign-fixed R), sigma geometric from 1 to
1/kappa, b ~ N(0,1); FP64 storage in HBM. A step is one complete solve to FP64 normwise
backward error <= n * 2^-53: fp16 LU (fp16 MFMA trailing updates with fp32 accumulation,
fp32 panels with tournament pivoting in 16-column blocks, ELX_MIXED_PANEL) and FP64
refinement (IR, switching to GMRES-IR when it stalls), with the
residual and the error measured on the device in FP64. The plain FP64 GPU LU + solve of the
same system (a fresh copy of A per solve, as El::LinearSolve copies A) is timed beside it,
and the reference's CPU path (the oracle's restatement of El::LU + lu::SolveAfter over the
host BLAS) on a bounded sample.

Prints one JSON line. `value` = mixed-precision solves/s.
"""
import argparse
import contextlib
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {"hbm_GBs": 8000.0, "fp64_TFs": 78.6, "fp32_TFs": 157.3, "fp16_TFs": 2500.0,
         "bf16_TFs": 2500.0}
CLASSES = ["lu_panel", "dgemm_f64", "lpgemm", "trsm_small", "row_moves", "trsv", "residual_f64",
           "pack"]


def make_factors(n, seed):
    """U, V Haar-orthogonal (QR of seeded Gaussians with sign-fixed R) for randsvd systems."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)

    def haar():
        G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
        Q, R = torch.linalg.qr(G)
        del G
        return Q * torch.sign(torch.diagonal(R))[None, :]

    return haar(), haar(), g


def randsvd_from(U, V, kappa):
    """A = U diag(sigma) V^T with sigma geometric from 1 to 1/kappa, column-major FP64."""
    import torch
    from elemental_amd import device as dev
    n = U.shape[0]
    sig = kappa ** (-torch.arange(n, dtype=torch.float64, device="cuda") / max(n - 1, 1))
    A = dev.colmajor_empty(n, n)
    A.copy_((U * sig[None, :]) @ V.T.contiguous())
    torch.cuda.synchronize()
    return A


def make_problem(n, kappa, seed, matrix="randsvd"):
    import torch
    from elemental_amd import device as dev
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    if matrix == "gaussian":
        # N(0,1) entries: the same kernels and traffic as randsvd without the QR
        A = dev.colmajor_empty(n, n)
        A.copy_(torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g).T)
        b = dev.colmajor_empty(n, 1)
        b.copy_(torch.randn(n, 1, dtype=torch.float64, device="cuda", generator=g))
        torch.cuda.synchronize()
        return A, b

    def haar():
        G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
        Q, R = torch.linalg.qr(G)
        del G
        return Q * torch.sign(torch.diagonal(R))[None, :]

    U = haar()
    V = haar()
    sig = kappa ** (-torch.arange(n, dtype=torch.float64, device="cuda") / max(n - 1, 1))
    A = dev.colmajor_empty(n, n)
    A.copy_((U * sig[None, :]) @ V.T.contiguous())
    del U, V
    b = dev.colmajor_empty(n, 1)
    b.copy_(torch.randn(n, 1, dtype=torch.float64, device="cuda", generator=g))
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return A, b


def backward_error(A, x, b):
    """max over columns of ||b - A x||inf / (||A||inf ||x||inf + ||b||inf), in FP64."""
    r = b - A @ x
    an = A.abs().sum(dim=1).max()
    eta = r.abs().amax(dim=0) / (an * x.abs().amax(dim=0) + b.abs().amax(dim=0))
    return float(eta.max())


def prof_read(lib):
    out = {}
    for c, name in enumerate(CLASSES):
        cnt = ctypes.c_int64()
        ms = ctypes.c_double()
        work = ctypes.c_double()
        lib.elx_profile_query(c, ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(work))
        if cnt.value:
            out[name] = {"launches": cnt.value, "ms": ms.value, "work": work.value}
    return out


def roofline_of(cls, rec, factor):
    t = rec["ms"] * 1e-3
    if cls in ("lpgemm",):
        peak = PEAKS["bf16_TFs" if factor == "bf16" else "fp16_TFs"]  # fp32: fp16 MFMA products
        ach = rec["work"] / t / 1e12
        return {"bound": "mfma", "achieved": round(ach, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 5)}
    if cls in ("dgemm_f64",):
        peak = PEAKS["fp64_TFs"]
        ach = rec["work"] / t / 1e12
        return {"bound": "mfma", "achieved": round(ach, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(ach / peak, 5)}
    ach = rec["work"] / t / 1e9
    return {"bound": "hbm", "achieved": round(ach, 2), "peak": PEAKS["hbm_GBs"], "unit": "GB/s",
            "frac": round(ach / PEAKS["hbm_GBs"], 5)}


# kernel class -> (kernel-name prefixes whose dispatches are the class's launches, prefixes of
# helper kernels launched inside the same class launch) in a rocprofv3 counter CSV
CLASS_KERNELS = {
    "lu_panel": (("void (anonymous namespace)::calu_panel_kernel<",
                  "void (anonymous namespace)::lu_panel_kernel<float>"), ()),
    "lpgemm": (("void (anonymous namespace)::lpgemm_kernel",), ()),
    "row_moves": (("(anonymous namespace)::row_moves_rowmajor_kernel",), ()),
    "trsm_small": (("void (anonymous namespace)::trsm_small_kernel<float>",
                    "(anonymous namespace)::trmm_inv_kernel"), ()),
    "trsv": (("(anonymous namespace)::trsv_kernel",), ()),
    "residual_f64": (("(anonymous namespace)::gemv_partial_kernel",),
                     ("(anonymous namespace)::residual_finish_kernel",)),
    "pack": (("void (anonymous namespace)::pack_kernel",), ()),
}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "lu_ir_n32768_hbm_traffic_pmc.md")


def counters_summary(fetch_csv, write_csv):
    """Markdown summary of two rocprofv3 --pmc CSVs (FETCH_SIZE, WRITE_SIZE) over a
    `bench.py --counters` run: per kernel class, KiB per dispatch as read, then HBM bytes per
    dispatch after scaling each counter by what it reads for the run's closing 1 GiB copy
    (the last dispatch, a device-to-device hipMemcpy)."""
    import csv
    def load(path):
        with open(path, newline="") as f:
            return [(r["Kernel_Name"], float(r["Counter_Value"])) for r in csv.DictReader(f)]
    fetch, write = load(fetch_csv), load(write_csv)
    gib = float(1 << 30)
    f_scale = gib / (fetch[-1][1] * 1024.0)
    w_scale = gib / (write[-1][1] * 1024.0)
    out = ["| kernel class | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | "
           "HBM bytes/dispatch (scaled) |", "|---|---|---|---|---|"]
    machine = []
    for cls, (main, helpers) in CLASS_KERNELS.items():
        launches = sum(1 for k, _ in fetch if k.startswith(main))
        if not launches or launches != sum(1 for k, _ in write if k.startswith(main)):
            continue
        fm = sum(v for k, v in fetch if k.startswith(main + helpers)) / launches
        wm = sum(v for k, v in write if k.startswith(main + helpers)) / launches
        tb = 1024.0 * (fm * f_scale + wm * w_scale)
        out.append(f"| {cls} | {launches} | {fm:.1f} | {wm:.1f} | {tb:.4g} |")
        machine.append(f"traffic_bytes_per_launch[{cls}] = {int(round(tb))}")
    out.append("")
    out.append(f"Calibration (1 GiB copy): FETCH_SIZE read {fetch[-1][1]:.1f} KiB -> scale "
               f"{f_scale:.4f}; WRITE_SIZE read {write[-1][1]:.1f} KiB -> scale {w_scale:.4f}.")
    out.append("")
    out.extend(machine)
    return "\n".join(out)


def traffic_from_profile(cls):
    """HBM bytes per launch of `cls` from the committed counter summary (n = 32768, fp16
    factor), or None when the summary or the class is absent."""
    import re
    try:
        with open(TRAFFIC_PROFILE) as f:
            text = f.read()
    except OSError:
        return None
    m = re.search(r"^traffic_bytes_per_launch\[" + re.escape(cls) + r"\] = (\d+)$", text, re.M)
    return int(m.group(1)) if m else None


def pick_host_blas(orc, cores):
    """The fastest host BLAS on this machine (a short dgemm calibration among MKL, OpenBLAS
    and scipy's bundled OpenBLAS); returns (backend path, dgemm GFLOP/s) or (None, None)."""
    import glob
    import numpy as np
    cands = ["mkl", "openblas"] + sorted(
        glob.glob("/usr/local/lib/python3*/dist-packages/scipy.libs/libscipy_openblas*.so"))
    best, best_t = None, None
    m = 2048
    Ac = np.asfortranarray(np.random.default_rng(1).standard_normal((m, m)))
    for cand in cands:
        try:
            orc.set_backend(cand)
        except RuntimeError:
            continue
        orc.set_threads(cores)
        orc.gemm("N", "N", 1.0, Ac, Ac, 0.0, Ac)
        t0 = time.perf_counter()
        orc.gemm("N", "N", 1.0, Ac, Ac, 0.0, Ac)
        t = time.perf_counter() - t0
        if best_t is None or t < best_t:
            best, best_t = cand, t
    if best is None:
        return None, None
    return orc.set_backend(best), 2 * m ** 3 / best_t / 1e9


def cpu_baseline(n_metric, A_sample, b_sample, n_small, reps=3):
    """The reference's CPU path on the GPU box's host cores: the oracle's restated El::LU +
    lu::SolveAfter (nb = 128, the reference's call sequence) over the fastest host BLAS, on a
    randsvd sample of the same family at a bounded size. Medians of `reps` runs: (a) all cores
    at the sample size, (b) one thread at n_small (the sample's leading block), (c) the same
    BLAS's LAPACK dgetrf + a solve on all cores at the sample size. Each is scaled by the n^3
    ratio to solves/s at n_metric."""
    import numpy as np
    from oracle import oracle as orc
    cores = len(os.sched_getaffinity(0))
    backend, gemm_gflops = pick_host_blas(orc, cores)
    if backend is None:
        return None
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    physical = None
    try:
        ids = set()
        phys = core = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                phys = line.split(":")[1].strip()
            elif line.startswith("core id"):
                core = line.split(":")[1].strip()
                ids.add((phys, core))
        physical = len(ids) or None
    except OSError:
        pass

    def timed(fn):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), [round(t, 3) for t in ts]

    def el_lu_solve(A, b):
        F, d, info = orc.lu(A)
        return orc.lu_solve_after(F, d, b), info

    ns = A_sample.shape[0]
    orc.set_threads(cores)
    t_all, runs_all = timed(lambda: el_lu_solve(A_sample, b_sample))
    x, info = el_lu_solve(A_sample, b_sample)
    eta_all = orc.backward_error(A_sample, x, b_sample)
    As = np.asfortranarray(A_sample[:n_small, :n_small])
    bs = np.asfortranarray(b_sample[:n_small])
    orc.set_threads(1)
    t_one, runs_one = timed(lambda: el_lu_solve(As, bs))
    orc.set_threads(cores)

    def host_lapack():
        F, ip, inf = orc.host_dgetrf(A_sample)
        if inf == -100:
            return None
        return orc.lu_solve_after(F, ip, b_sample)
    lapack = None
    if host_lapack() is not None:
        t_lap, runs_lap = timed(host_lapack)
        lapack = {"value": round((ns / n_metric) ** 3 / t_lap, 6), "unit": "solves/s",
                  "seconds_at_sample_median": round(t_lap, 3), "runs": runs_lap,
                  "routine": "host LAPACK dgetrf + the restated lu::SolveAfter"}
    orc.set_backend("naive")
    return {"value": round((ns / n_metric) ** 3 / t_all, 6), "unit": "solves/s",
            "cores": cores, "kind": "port",
            "sample": f"oracle El::LU + lu::SolveAfter (nb=128) over {os.path.basename(backend)} "
                      f"with {cores} threads on a randsvd system (same kappa) at n={ns}, median of "
                      f"{reps} runs, scaled by (n_sample/n)^3 to n={n_metric}",
            "seconds_at_sample_median": round(t_all, 3), "runs": runs_all,
            "backward_error_at_sample": eta_all, "info": int(info),
            "one_thread": {"value": round((n_small / n_metric) ** 3 / t_one, 8),
                           "unit": "solves/s", "cores": 1, "n_sample": n_small,
                           "seconds_at_sample_median": round(t_one, 3), "runs": runs_one},
            "host_lapack_dgetrf": lapack, "blas": backend, "host_dgemm_gflops": round(gemm_gflops, 1),
            "cpu_model": cpu_model, "logical_cpus": os.cpu_count(), "physical_cores": physical}


def config1_check():
    """Config 1 (n = 1024, A and b from Elemental's Gaussian stream, seed (21<<16)|0): the GPU
    El::LU + lu::SolveAfter against the oracle's restatement -- pivot agreement with near-ties
    counted, double-double backward errors of both, and the reference's own LU test criterion
    (tests/lapack_like/LU.cpp:43-53)."""
    import numpy as np
    from elemental_amd import device as dev
    from oracle import oracle as orc
    orc.set_backend("naive")
    n = 1024
    orc.rng_seed()
    A = orc.gaussian(n, n)
    b = orc.gaussian(n, 1)
    F_ref, d_ref, info_ref = orc.lu(A)
    x_ref = orc.lu_solve_after(F_ref, d_ref, b)
    dA = dev.to_device(A)
    d, info = dev.getrf(dA)
    dB = dev.to_device(b)
    dev.getrs(dA, d, dB)
    x = dev.to_host(dB)
    d = d.cpu().numpy()
    agree = orc.pivot_agreement(A, d_ref, d)
    eta = orc.backward_error(A, x, b)
    rel = np.abs(A @ x - b).max() / (2.0 ** -53 * n * max(np.abs(A).sum(0).max(), np.abs(x).sum(0).max()))
    return {"n": n, "pivot_mismatches": agree["mismatches"], "near_ties": agree["near_ties"],
            "first_mismatch": agree["first"], "backward_error_gpu": eta,
            "backward_error_oracle": orc.backward_error(A, x_ref, b),
            "tolerance": n * 2.0 ** -53, "el_lu_test_rel_error": rel,
            "factor_error_gpu": orc.lu_factor_error(A, dev.to_host(dA), d),
            "max_abs_diff_x_vs_oracle": float(np.abs(x - x_ref).max())}


def measure_peaks():
    """Re-measured peaks on this box: large library GEMMs (torch.matmul, FP64/FP32/FP16/BF16,
    8192^3, best of 3) and a 4 GiB device-to-device copy (read + write bytes)."""
    import torch
    out = {}
    for dt, key in ((torch.float64, "fp64_TFs"), (torch.float32, "fp32_TFs"),
                    (torch.float16, "fp16_TFs"), (torch.bfloat16, "bf16_TFs")):
        m = 8192
        a = torch.randn(m, m, device="cuda", dtype=torch.float32).to(dt)
        bm = torch.randn(m, m, device="cuda", dtype=torch.float32).to(dt)
        c = a @ bm
        best = None
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            c = a @ bm
            e1.record()
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) * 1e-3
            best = t if best is None else min(best, t)
        out[key] = {"datasheet": PEAKS[key], "measured": round(2 * m ** 3 / best / 1e12, 1)}
        del a, bm, c
    src = torch.ones(1 << 29, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = None
    for _ in range(3):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) * 1e-3
        best = t if best is None else min(best, t)
    out["hbm_GBs"] = {"datasheet": PEAKS["hbm_GBs"], "measured": round(2 * src.numel() * 8 / best / 1e9, 1)}
    del src, dst
    torch.cuda.empty_cache()
    out["method"] = "torch.matmul (rocBLAS/hipBLASLt) 8192^3 and a 4 GiB copy, best of 3"
    return out


def config3_rows(args, U, V, A, b):
    """The rest of config 3 at n = 32768, one timed solve each (outside the headline's timed
    region): bf16 factors at the headline kappa, 64 right-hand sides at the headline kappa
    (block refinement), and fp16 factors at kappa 1e4 and 1e6 (AUTO escalation)."""
    import torch
    from elemental_amd import device as dev
    n = A.shape[0]
    rows = []
    g = torch.Generator(device="cuda")
    g.manual_seed(args.seed + 7)
    B64 = dev.colmajor_empty(n, 64)
    B64.copy_(torch.randn(n, 64, dtype=torch.float64, device="cuda", generator=g))

    def run(Am, rhs, factor, kappa):
        X = dev.colmajor_empty(n, rhs.shape[1])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = dev.gesv_mixed(Am, rhs, X, factor=factor, refine="auto")
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        row = {"kappa": kappa, "factor": factor, "nrhs": rhs.shape[1], "ms": round(1e3 * t, 1),
               "converged": st["converged"], "backward_error": backward_error(Am, X, rhs),
               "tolerance": n * 2.0 ** -53, "fell_back": st["fell_back"],
               "outer_iters": st["outer_iters"], "inner_iters": st["inner_iters"]}
        print(f"[bench] config3 {row}", file=sys.stderr, flush=True)
        return row

    rows.append(run(A, b, "bf16", args.kappa))
    rows.append(run(A, B64, args.factor, args.kappa))
    del B64
    for kappa in (1e4, 1e6):
        Ak = randsvd_from(U, V, kappa)
        rows.append(run(Ak, b, args.factor, kappa))
        del Ak
        torch.cuda.empty_cache()
    return rows


def hpd_solve(args):
    """HPD system n = 16384 with 16 right-hand sides: A = G G^T / n + I (FP64), B ~ N(0,1).
    Plain FP64 El::HPDSolve path (copy A, Cholesky, solve) against the low-precision
    Cholesky + FP64 refinement (fp16 factor, AUTO), both to backward error <= n u."""
    import torch
    from elemental_amd import device as dev
    n, nrhs = args.hpd_n, 16
    g = torch.Generator(device="cuda")
    g.manual_seed(args.seed + 2)
    G = torch.randn(n, n, dtype=torch.float64, device="cuda", generator=g)
    A = dev.colmajor_empty(n, n)
    A.copy_(G @ G.T.contiguous() / n + torch.eye(n, dtype=torch.float64, device="cuda"))
    del G
    B = dev.colmajor_empty(n, nrhs)
    B.copy_(torch.randn(n, nrhs, dtype=torch.float64, device="cuda", generator=g))
    torch.cuda.empty_cache()
    out = {"n": n, "nrhs": nrhs, "matrix": "G G^T/n + I", "flops_factor": n ** 3 / 3.0}
    Ac = dev.colmajor_empty(n, n)
    Xc = dev.colmajor_empty(n, nrhs)
    t_by = {}
    for gemm in ("own", "library"):
        with gemm_env("ELX_DGEMM", gemm):
            times = []
            for rep in range(args.hpd_steps + 1):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                Ac.copy_(A)
                Xc.copy_(B)
                dev.posv(Ac, Xc)
                torch.cuda.synchronize()
                if rep:
                    times.append(time.perf_counter() - t0)
            t_by[gemm] = sum(times) / len(times)
    out["plain_fp64_ms"] = round(1e3 * min(t_by.values()), 2)
    out["plain_fp64_ms_own_gemm"] = round(1e3 * t_by["own"], 2)
    out["plain_fp64_ms_rocblas_gemm"] = round(1e3 * t_by["library"], 2)
    out["plain_fp64_backward_error"] = backward_error(A, Xc, B)
    del Ac
    torch.cuda.empty_cache()
    X = dev.colmajor_empty(n, nrhs)
    times, st = [], None
    for rep in range(args.hpd_steps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = dev.posv_mixed(A, B, X, factor=args.factor, refine="auto")
        torch.cuda.synchronize()
        if rep:
            times.append(time.perf_counter() - t0)
    out["mixed_ms"] = round(1e3 * sum(times) / len(times), 2)
    out["mixed_backward_error"] = backward_error(A, X, B)
    out["mixed_stats"] = {k: st[k] for k in ("converged", "outer_iters", "inner_iters", "fell_back",
                                              "factor_seconds", "refine_seconds")}
    out["speedup"] = round(out["plain_fp64_ms"] / out["mixed_ms"], 3)
    out["tolerance"] = n * 2.0 ** -53
    print(f"[bench] hpd solve {out}", file=sys.stderr, flush=True)
    del A, B, X, Xc
    torch.cuda.empty_cache()
    return out


def batched_lu_ir(args):
    """Nonsymmetric systems (2,000, the full configuration; --lu-batch-count), n log-uniform in
    [64, 2048] (rounded to a multiple of 64),
    kappa log-uniform in [1e1, 1e8], A = U diag(sigma) V^T (U, V Haar from batched QR, sigma
    geometric), 4 right-hand sides each: fp16 LU + FP64 GMRES-IR with the FP64 LU fallback,
    every system to backward error <= n u."""
    import numpy as np
    import torch
    from elemental_amd import device as dev
    rng = np.random.default_rng(args.seed + 5)
    count = args.lu_batch_count
    ns = np.clip((np.exp(rng.uniform(np.log(64), np.log(2048), count)) / 64).round() * 64, 64, 2048).astype(int)
    kappas = 10.0 ** rng.uniform(1, 8, count)
    g = torch.Generator(device="cuda")
    g.manual_seed(args.seed + 5)
    As, Bs, Xs = [None] * count, [None] * count, [None] * count
    t0 = time.perf_counter()
    for n in sorted(set(ns.tolist())):
        idx = [i for i in range(count) if ns[i] == n]
        def haar(k):
            Q, R = torch.linalg.qr(torch.randn(k, n, n, dtype=torch.float64, device="cuda", generator=g))
            return Q * torch.sign(torch.diagonal(R, dim1=1, dim2=2))[:, None, :]
        U, V = haar(len(idx)), haar(len(idx))
        for j, i in enumerate(idx):
            sig = kappas[i] ** (-torch.arange(n, dtype=torch.float64, device="cuda") / (n - 1))
            A = dev.colmajor_empty(n, n)
            A.copy_((U[j] * sig[None, :]) @ V[j].T.contiguous())
            As[i] = A
            Bs[i] = dev.colmajor_empty(n, 4)
            Bs[i].copy_(torch.randn(n, 4, dtype=torch.float64, device="cuda", generator=g))
            Xs[i] = dev.colmajor_empty(n, 4)
        del U, V
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    status, stats = dev.gesv_mixed_batched(As, Bs, Xs, factor=args.factor, refine="auto")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    worst = 0.0
    for i in range(count):
        worst = max(worst, backward_error(As[i], Xs[i], Bs[i]) / (ns[i] * 2.0 ** -53))
    fell = sum(st["fell_back"] for st in stats)
    out = {"count": count, "n_range": [64, 2048], "kappa_range": [1e1, 1e8], "nrhs": 4,
           "factor": args.factor, "status": status, "seconds": round(wall, 3),
           "systems_per_s": round(count / wall, 2),
           "credited_flops": float(np.sum(2.0 / 3.0 * ns.astype(np.float64) ** 3 + 8.0 * ns.astype(np.float64) ** 2)),
           "fell_back": int(fell), "converged": int(sum(st["converged"] for st in stats)),
           "max_backward_error_over_n_u": worst, "generation_seconds": round(gen_s, 2),
           "schedule": "device-side: one persistent workgroup per system (fp16-MFMA LU, FP64 "
                       "IR/FGMRES), FP64 LU tier for the systems whose refinement stalls; the host "
                       "reads the per-system states once at the end"}
    out["credited_tflops"] = round(out["credited_flops"] / wall / 1e12, 4)
    print(f"[bench] batched lu-ir {out}", file=sys.stderr, flush=True)
    del As, Bs, Xs
    torch.cuda.empty_cache()
    return out


@contextlib.contextmanager
def gemm_env(name, value):
    """Runs the block with environment variable `name` set to `value` ("own" removes it: the
    library's own kernels are the default)."""
    old = os.environ.get(name)
    if value == "own":
        os.environ.pop(name, None)
    else:
        os.environ[name] = value
    try:
        yield
    finally:
        if old is None:
            os.environ.pop(name, None)
        else:
            os.environ[name] = old


def batched_cholesky(args, lib):
    """20,000 HPD matrices, n_i uniform in [16, 512], A_i = G_i G_i^T / n_i + I, 1..8 right-hand
    sides each: timed batched factorization (and solve) in FP64 and FP32."""
    import numpy as np
    import torch
    from elemental_amd import device as dev
    rng = np.random.default_rng(args.seed + 4)
    count = args.batch_count
    ns = rng.integers(16, 513, size=count)
    nrhs = rng.integers(1, 9, size=count)
    g = torch.Generator(device="cuda")
    g.manual_seed(args.seed + 4)
    out = {"count": int(count), "n_range": [16, 512], "nrhs_range": [1, 8],
           "factor_flops": float(np.sum(ns.astype(np.float64) ** 3) / 3.0)}
    # matrices grouped by order, one (k, n, n) tensor per order (symmetric, so the row-major
    # storage is the column-major matrix)
    orders = {}
    for i, n in enumerate(ns):
        orders.setdefault(int(n), []).append(i)
    for dtype, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
        store, rhs_store, mats, rhs = {}, {}, [None] * count, [None] * count
        for n, idx in orders.items():
            G = torch.randn(len(idx), n, n, dtype=torch.float64, device="cuda", generator=g)
            A = torch.bmm(G, G.transpose(1, 2).contiguous()) / n
            A += torch.eye(n, dtype=torch.float64, device="cuda")
            store[n] = A.to(dtype)
            del G, A
            for k, i in enumerate(idx):
                mats[i] = store[n][k]
                r = int(nrhs[i])
                rhs[i] = dev.colmajor_empty(n, r, dtype=dtype)
                rhs[i].copy_(torch.randn(n, r, dtype=torch.float64, device="cuda", generator=g))
        backup = {n: t.clone() for n, t in store.items()}
        rhs_backup = [r.clone() for r in rhs]
        fac = dev.Batch([m.t() for m in mats], dtype)   # .t(): column-major view, same data
        rb = dev.Batch(rhs, dtype)
        times, stimes = [], []
        for rep in range(args.batch_steps + 1):
            for n in store:
                store[n].copy_(backup[n])
            for r, r0 in zip(rhs, rhs_backup):
                r.copy_(r0)
            torch.cuda.synchronize()
            e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            e0.record()
            info = dev.potrf_vbatched(fac)
            e1.record()
            dev.potrs_vbatched(fac, rb)
            e2.record()
            torch.cuda.synchronize()
            if rep > 0:
                times.append(e0.elapsed_time(e1))
                stimes.append(e1.elapsed_time(e2))
        bad = int((info != 0).sum().item())
        # backward error of a sample of solves, in FP64 on the device
        worst = 0.0
        for i in list(range(0, count, max(1, count // 64))):
            A0 = backup[int(ns[i])][orders[int(ns[i])].index(i)].to(torch.float64)
            x = rhs[i].to(torch.float64)
            b = rhs_backup[i].to(torch.float64)
            r = b - A0 @ x
            eta = float(r.abs().max() / (A0.abs().sum(1).max() * x.abs().max() + b.abs().max()))
            worst = max(worst, eta / (int(ns[i]) * (2.0 ** -53 if tag == "f64" else 2.0 ** -24)))
        ms = sum(times) / len(times)
        tf = out["factor_flops"] / (ms * 1e-3) / 1e12
        peak = PEAKS["fp64_TFs"] if tag == "f64" else PEAKS["fp32_TFs"]
        out[tag] = {"factor_ms": round(ms, 3), "factor_tflops": round(tf, 3),
                    "frac_of_peak": round(tf / peak, 4), "peak_tflops": peak,
                    "solve_ms": round(sum(stimes) / len(stimes), 3), "non_hpd": bad,
                    "solve_backward_error_max_over_n_u": worst}
        print(f"[bench] batched cholesky {tag}: {out[tag]}", file=sys.stderr, flush=True)
        del store, backup, mats, rhs, rhs_backup, fac, rb
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--kappa", type=float, default=1e2)
    ap.add_argument("--factor", default="fp16", choices=["fp16", "bf16", "fp32"])
    ap.add_argument("--refine", default="auto", choices=["ir", "gmres", "auto"])
    ap.add_argument("--seed", type=int, default=20240)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-n", type=int, default=12288)
    ap.add_argument("--cpu-n-one-thread", type=int, default=4096)
    ap.add_argument("--skip-config3", action="store_true")
    ap.add_argument("--skip-peaks", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--no-profile-pass", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-hpd", action="store_true")
    ap.add_argument("--matrix", default="randsvd", choices=["randsvd", "gaussian"])
    ap.add_argument("--counters", action="store_true",
                    help="one solve and a 1 GiB copy, for hardware-counter runs; no JSON bench line")
    ap.add_argument("--counters-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print the markdown summary of two rocprofv3 --pmc CSVs and exit")
    ap.add_argument("--skip-lu-batch", action="store_true")
    ap.add_argument("--lu-batch-count", type=int, default=2000)
    ap.add_argument("--hpd-n", type=int, default=16384)
    ap.add_argument("--hpd-steps", type=int, default=1)
    ap.add_argument("--batch-count", type=int, default=20000)
    ap.add_argument("--batch-steps", type=int, default=2)
    args = ap.parse_args()
    if args.counters_summary:
        print(counters_summary(*args.counters_summary))
        return
    if args.gpus != 1:
        print(f"bench.py measures one MI355X; --gpus {args.gpus} is not supported "
              "(this build has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    import torch
    from elemental_amd import device as dev
    from elemental_amd._lib import load
    lib = load()
    lib.elx_profile_enable.argtypes = [ctypes.c_int]
    lib.elx_profile_query.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    torch.cuda.set_device(0)
    n = args.n
    print(f"[bench] generating n={n} kappa={args.kappa:g}", file=sys.stderr, flush=True)
    if args.counters:
        # one solve plus a 1 GiB copy (the counters' calibration) for rocprofv3 --pmc runs
        A, b = make_problem(n, args.kappa, args.seed, "gaussian")
        X = dev.colmajor_empty(n, 1)
        st = dev.gesv_mixed(A, b, X, factor=args.factor, refine=args.refine)
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"counters_run": True, "solve": st, "copy_bytes": src.numel() * 8}), flush=True)
        return
    peaks = None if args.skip_peaks else measure_peaks()
    if peaks:
        print(f"[bench] peaks {peaks}", file=sys.stderr, flush=True)
    U = V = None
    if args.matrix == "randsvd":
        U, V, g = make_factors(n, args.seed)
        A = randsvd_from(U, V, args.kappa)
        b = dev.colmajor_empty(n, 1)
        b.copy_(torch.randn(n, 1, dtype=torch.float64, device="cuda", generator=g))
        torch.cuda.synchronize()
    else:
        A, b = make_problem(n, args.kappa, args.seed, args.matrix)
    X = dev.colmajor_empty(n, 1)

    def mixed_solve():
        return dev.gesv_mixed(A, b, X, factor=args.factor, refine=args.refine)

    for _ in range(args.warmup):
        st = mixed_solve()
        print(f"[bench] warmup mixed {st}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(args.steps):
        stats.append(mixed_solve())
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t0
    ms_step = 1e3 * t_mixed / args.steps
    value = args.steps / t_mixed
    ok = all(s["converged"] == 1 for s in stats)
    eta_dev = max(s["backward_error"] for s in stats)
    eta_chk = backward_error(A, X, b)
    print(f"[bench] mixed {ms_step:.1f} ms/solve, eta(device)={eta_dev:.3e} eta(check)={eta_chk:.3e} "
          f"{stats[-1]}", file=sys.stderr, flush=True)

    plain = None
    if not args.skip_plain:
        Ac = dev.colmajor_empty(n, n)
        xb = dev.colmajor_empty(n, 1)
        # the plain solve on this library's FP64 GEMM (the default) and on rocBLAS's
        # (ELX_DGEMM=library); the speedup is quoted against the faster of the two
        t_by = {}
        for gemm in ("own", "library"):
            with gemm_env("ELX_DGEMM", gemm):
                times = []
                for rep in range(args.plain_steps + 1):
                    torch.cuda.synchronize()
                    t1 = time.perf_counter()
                    Ac.copy_(A)
                    xb.copy_(b)
                    d, info = dev.getrf(Ac)
                    dev.getrs(Ac, d, xb)
                    torch.cuda.synchronize()
                    if rep > 0:
                        times.append(time.perf_counter() - t1)
                t_by[gemm] = sum(times) / len(times)
        fastest = min(t_by, key=t_by.get)
        t_plain = t_by[fastest]
        plain = {"solves_per_s": round(1.0 / t_plain, 4), "ms_per_solve": round(1e3 * t_plain, 2),
                 "faster_gemm": fastest,
                 "ms_per_solve_own_gemm": round(1e3 * t_by["own"], 2),
                 "ms_per_solve_rocblas_gemm": round(1e3 * t_by["library"], 2),
                 "backward_error": backward_error(A, xb, b), "info": int(info.item())}
        if not args.no_profile_pass:
            # kernel classes of one profiled plain FP64 solve
            lib.elx_profile_reset()
            lib.elx_profile_enable(1)
            Ac.copy_(A)
            xb.copy_(b)
            d, info = dev.getrf(Ac)
            dev.getrs(Ac, d, xb)
            torch.cuda.synchronize()
            plain["kernels"] = prof_read(lib)
            lib.elx_profile_enable(0)
        del Ac
        torch.cuda.empty_cache()
        print(f"[bench] plain fp64 {plain}", file=sys.stderr, flush=True)
        # for reference: the mixed solve with rocBLAS's gemm_ex for the rank-NB updates
        # (ELX_LOWP_GEMM=library) in place of this library's kernel, one warmup and one step
        with gemm_env("ELX_LOWP_GEMM", "library"):
            mixed_solve()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            mixed_solve()
            torch.cuda.synchronize()
            plain["mixed_ms_with_rocblas_rank_nb_update"] = round(1e3 * (time.perf_counter() - t1), 2)

    kernels = {}
    roof = None
    if not args.no_profile_pass:
        lib.elx_profile_reset()
        lib.elx_profile_enable(1)
        mixed_solve()
        torch.cuda.synchronize()
        kernels = prof_read(lib)
        lib.elx_profile_enable(0)
        dom = max(kernels, key=lambda k: kernels[k]["ms"])
        roof = roofline_of(dom, kernels[dom], args.factor)
        roof["kernel"] = dom
        roof["traffic"] = (traffic_from_profile(dom)
                           if (n, args.factor) == (32768, "fp16") else None)
        roof["avg_launch_us"] = round(1e3 * kernels[dom]["ms"] / kernels[dom]["launches"], 2)
        for k, v in kernels.items():
            v["share"] = round(v["ms"] / sum(x["ms"] for x in kernels.values()), 4)
        print(f"[bench] kernel classes {json.dumps(kernels)}", file=sys.stderr, flush=True)

    config3 = None
    if not args.skip_config3 and U is not None:
        config3 = config3_rows(args, U, V, A, b)
    del U, V
    torch.cuda.empty_cache()

    hpd = None
    if not args.skip_hpd:
        hpd = hpd_solve(args)

    batched = None
    if not args.skip_batched:
        batched = batched_cholesky(args, lib)

    lu_batch = None
    if not args.skip_lu_batch:
        lu_batch = batched_lu_ir(args)

    cpu = None
    config1 = None
    if not args.skip_cpu:
        del A
        torch.cuda.empty_cache()
        # a randsvd sample of the same family at a bounded size, built on the device
        Us, Vs, gs = make_factors(args.cpu_n, args.seed + 9)
        As = randsvd_from(Us, Vs, args.kappa)
        bs = torch.randn(args.cpu_n, 1, dtype=torch.float64, device="cuda", generator=gs)
        A_h, b_h = dev.to_host(As), dev.to_host(bs)
        del Us, Vs, As, bs
        torch.cuda.empty_cache()
        cpu = cpu_baseline(n, A_h, b_h, args.cpu_n_one_thread)
        del A_h, b_h
        print(f"[bench] cpu baseline {cpu}", file=sys.stderr, flush=True)
        config1 = config1_check()
        print(f"[bench] config1 {config1}", file=sys.stderr, flush=True)

    credited = 2.0 / 3.0 * n ** 3 + 2.0 * n ** 2
    whole = {"credited_flops": credited,
             "mixed_tflops": round(credited * value / 1e12, 2),
             "mixed_frac_of_fp16_peak": round(credited * value / 1e12 / PEAKS["fp16_TFs"], 4),
             "note": "HPL-MxP convention: (2/3)n^3 + 2n^2 credited whatever the precision"}
    if plain:
        whole["plain_fp64_tflops"] = round(credited * plain["solves_per_s"] / 1e12, 2)
        whole["plain_fp64_frac_of_fp64_peak"] = round(
            credited * plain["solves_per_s"] / 1e12 / PEAKS["fp64_TFs"], 4)

    line = {
        "metric": "FP64-accurate Ax=b solves/s: FP16-LU-IR vs FP64 LU at n=32768, 1 MI355X",
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.factor + "+f64",
        "data": "synthetic",
        "config": {"workload": f"LU-IR solve, n={n}, randsvd kappa={args.kappa:g}, 1 RHS, "
                               f"{args.factor} factor, FP64 refinement ({args.refine})",
                   "n": n, "kappa": args.kappa, "nrhs": 1, "factor": args.factor,
                   "refine": args.refine, "seed": args.seed,
                   "panel": os.environ.get("ELX_MIXED_PANEL", "calu16")},
        "converged": ok,
        "backward_error": eta_dev,
        "backward_error_check": eta_chk,
        "tolerance": n * 2.0 ** -53,
        "refine_stats": {k: stats[-1][k] for k in ("outer_iters", "inner_iters", "factor_seconds",
                                                    "refine_seconds")},
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (round(value / plain["solves_per_s"], 3) if plain else None),
        "roofline": roof,
        "kernels": kernels,
        "cpu_baseline": cpu,
        "whole_solve": whole,
        "peaks": peaks,
        "config1_check": config1,
        "config3": config3,
        "hpd_solve": hpd,
        "batched_cholesky": batched,
        "batched_lu_ir": lu_batch,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
