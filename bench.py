"""Benchmark: eigensolves/s (+ p50/p99 latency, residuals) of the implicitly
restarted Lanczos solver on one MI355X, with the SpMV's GB/s, the roofline of
the dominant kernel and the CPU oracle timed on a bounded sample beside it.

Default workload (one "step" = one full eigensolve): normalized Laplacian of a
symmetric Chung-Lu power-law graph (4,194,304 vertices, exponent 2.1, mean
degree 16, max expected degree 2e4, weights U(0,1], giant component kept),
FP64 CSR; nev = 20, ncv = 60, largest algebraic, tol 1e-10.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload chunglu|stencil27|laplace2d]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
WORKLOADS = {
    # name: (generator, kwargs, nev, ncv, label)
    "chunglu": ("chung_lu_laplacian", {"n": 1 << 22, "seed": 0}, 20, 60,
                "chung-lu normalized Laplacian, 2^22 vertices (giant component), mean degree 16, exponent 2.1"),
    "stencil27": ("stencil27", {"m": 128, "seed": 0}, 32, 96, "27-point stencil 128^3, seeded coefficients"),
    "laplace2d": ("laplacian_2d", {"m": 256}, 10, 30, "2-D 5-point Laplacian 256x256 (Dirichlet)"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else float("nan")


def residuals(op, values, vectors, n):
    """max_i ||A v_i - l_i v_i|| / ||A||, with ||A|| estimated by max |l_i| (a lower
    bound, so the ratio is conservative), computed with the GPU SpMV."""
    import torch

    V = torch.tensor(np.ascontiguousarray(vectors[:, : len(values)].T), device="cuda")
    AV = torch.empty_like(V)
    for i in range(len(values)):
        op.spmv(V[i].data_ptr(), AV[i].data_ptr())
    torch.cuda.synchronize()
    lam = torch.tensor(values, device="cuda")[:, None]
    r = torch.linalg.norm(AV - lam * V, dim=1).cpu().numpy()
    nrm = float(np.max(np.abs(values)))
    ortho = float((V @ V.T - torch.eye(len(values), device="cuda", dtype=V.dtype)).abs().max())
    return float(r.max() / nrm), ortho


def spmv_timing(op, n, reps=50):
    """Average duration of the SpMV kernel over `reps` launches on torch's
    current stream, measured with HIP events on that stream."""
    import torch

    x = torch.rand(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    s = torch.cuda.current_stream()
    for _ in range(5):
        op.spmv(x.data_ptr(), y.data_ptr(), s.cuda_stream)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(s)
    for _ in range(reps):
        op.spmv(x.data_ptr(), y.data_ptr(), s.cuda_stream)
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) * 1e-3 / reps


def cpu_baseline(rp, col, val, nev, ncv, tol, gpu_spmv_per_solve):
    """The CPU oracle (oracle/hlanc_oracle.c, a scalar-per-thread restatement of
    ds_solver_mkl) on a bounded sample: its first restart cycle on the same matrix
    (ncv Lanczos steps + one restart), extrapolated to a full solve by the number
    of Lanczos steps the GPU solve needed."""
    import oracle as O

    threads = min(16, os.cpu_count() or 1)
    A = O.Csr(rp, col, val)
    t = time.perf_counter()
    r = O.ds_solver(A, nev, ncv, tol=tol, maxiter=1, nthreads=threads, max_restarts_then_stop=1)
    dt = time.perf_counter() - t
    steps = max(int(r.spmv), 1)
    per_solve = dt * gpu_spmv_per_solve / steps
    return {"value": 1.0 / per_solve, "unit": "eigensolves/s", "cores": threads, "kind": "port",
            "sample": f"oracle first restart cycle ({steps} Lanczos steps, {dt:.1f} s on {threads} threads), "
                      f"scaled to the {gpu_spmv_per_solve} steps of the GPU solve"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="chunglu", choices=sorted(WORKLOADS))
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--maxiter", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy 5 times (scale for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (no collectives); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_calibration:
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 5 x 1 GiB")
        return

    import hlanc_amd as H
    from hlanc_amd import workloads as W

    gen, kw, nev, ncv, label = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    rp, col, val = getattr(W, gen)(**kw)
    n = len(rp) - 1
    log(f"workload {args.workload}: n={n} nnz={len(val)} generated in {time.perf_counter() - t0:.1f} s")
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))
    torch.cuda.init()

    for i in range(args.warmup):
        r = H.ds_solver_gpu(n, nev, ncv, op, tol=args.tol, maxiter=args.maxiter)
        log(f"warmup {i}: ret {r.ret} restarts {r.info.restarts} spmv {r.info.spmv} {r.info.seconds:.3f} s")

    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        r = H.ds_solver_gpu(n, nev, ncv, op, tol=args.tol, maxiter=args.maxiter)
        lat.append(time.perf_counter() - t)
        log(f"solve {i}: ret {r.ret} restarts {r.info.restarts} spmv {r.info.spmv} {lat[-1]:.3f} s")
        if r.ret != 0:
            raise SystemExit(f"solve {i} did not converge (ret {r.ret})")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    value = args.steps / wall

    resid, ortho = residuals(op, r.values, r.vectors, n)
    log(f"max rel residual {resid:.2e}, orthogonality {ortho:.2e}")

    # per-kernel-class profile of one more solve (HIP events around each launch)
    import ctypes

    o = H.options(args.tol, args.maxiter)
    o.flags = H.HL_FLAG_PROFILE
    values = np.zeros(ncv)
    vecs = np.zeros((ncv, n)).T
    info = H.Info()
    rc = H.lib().hl_ds_solve_csr(op.h, nev, ncv, values.ctypes.data_as(H._dp), vecs.ctypes.data_as(H._dp), n,
                                 ctypes.byref(o), ctypes.byref(info))
    pr = H.Profile()
    H.lib().hl_profile_last(ctypes.byref(pr))
    names = H.PROFILE_CLASSES
    classes = {names[k]: {"launches": int(pr.launches[k]), "seconds": pr.seconds[k], "bytes": pr.bytes[k],
                          "GBps": (pr.bytes[k] / pr.seconds[k] / 1e9) if pr.seconds[k] > 0 else None}
               for k in range(len(names)) if pr.launches[k] > 0}
    dom = max(range(len(names)), key=lambda k: pr.seconds[k])
    log(f"profile (ret {rc}): " + ", ".join(f"{k} {v['launches']}x {v['seconds']:.3f}s {v['GBps'] or 0:.0f} GB/s"
                                           for k, v in classes.items()))
    achieved = pr.bytes[dom] / pr.seconds[dom] / 1e9

    t_spmv = spmv_timing(op, n)
    spmv_bytes = op.spmv_bytes()
    spmv_gbps = spmv_bytes / t_spmv / 1e9
    log(f"SpMV {t_spmv * 1e6:.1f} us/launch, {spmv_gbps:.0f} GB/s ({spmv_bytes / 1e6:.1f} MB algorithmic)")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(rp, col, val, nev, ncv, args.tol, int(r.info.spmv))
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "eigensolves/s at rel.resid<=1e-10 (+ p50/p99 latency); SpMV GB/s vs HBM peak",
        "value": value,
        "unit": "eigensolves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": label, "n": n, "nnz": int(len(val)), "nev": nev, "ncv": ncv, "tol": args.tol,
                   "which": "largest algebraic", "basis": "fp64"},
        "latency_s": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
        "max_rel_resid": resid,
        "ritz_orthogonality": ortho,
        "restarts": int(r.info.restarts),
        "spmv_per_solve": int(r.info.spmv),
        "cgs_passes_per_solve": int(r.info.cgs_passes),
        "spmv": {"GBps": spmv_gbps, "us_per_launch": t_spmv * 1e6, "bytes_per_launch": spmv_bytes,
                 "frac_of_peak": spmv_gbps / HBM_PEAK_GBPS},
        "kernel_classes": classes,
        "roofline": {"kernel": names[dom], "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
