"""Benchmark: eigensolves/s (+ p50/p99 latency, residuals) of the implicitly
restarted Lanczos solver on one MI355X, with the SpMV's GB/s, the roofline of
the dominant kernel and the CPU oracle timed on a bounded sample beside it.

Default workload (one "step" = one full eigensolve): normalized Laplacian of a
symmetric Chung-Lu power-law graph (4,194,304 vertices, exponent 2.1, mean
degree 16, max expected degree 2e4, weights U(0,1], giant component kept),
FP64 CSR; nev = 20, ncv = 60, largest algebraic, tol 1e-10.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload chunglu|stencil27|laplace2d]

Two more workloads measure the batched and served forms (not part of the
default run):

* ``--workload batch``: 512 independent symmetric matrices, n log-uniform in
  [1e3, 2e5] (a quarter 2-D 5-point stencils, a quarter 3-D 27-point stencils,
  half Erdos-Renyi graphs of degree 8..32 with weights U(-1,1) and a diagonal
  shift), nev = 8, ncv = 24.  One step solves all 512: once one solve at a time
  (the plain ds_solver_gpu path) and once as one batch (hl_ds_solve_batch); the
  value is the batched eigensolves/s.
* ``--workload serve``: Poisson arrivals at each of ``--rates`` requests/s of solves
  drawn from a pool of matrices with n log-uniform in [4e3, 2e6], nev uniform
  in {4, 8, 16, 32}, ncv = 3 nev; continuous batching under an HBM budget,
  against first-come-first-served one at a time on the same arrival stream;
  reports eigensolves/s and p50/p99 latency from arrival to result.
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


BASIS_LABEL = {"fp64": "fp64", "fp32": "mixed: fp32 basis and values to 1e-6, then fp64 to tol",
               "proj32": "fp64 basis + fp32 copy for the projection passes (DGKS correction every step)"}


def basis_prec(H, basis):
    return {"fp64": H.HL_PREC_F64, "fp32": H.HL_PREC_F32, "proj32": H.HL_PREC_F64_PROJ32}[basis]


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
    Vh = np.ascontiguousarray(vectors[:, : len(values)])
    ortho = float(np.abs(Vh.T @ Vh - np.eye(len(values))).max())  # on the host: a skinny product
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


def check_pairs(ops, results, nevs):
    """max over all returned pairs of ||A v - l v|| / max|l| (GPU SpMV)."""
    worst = 0.0
    for op, r, nev in zip(ops, results, nevs):
        if r.vectors is None:
            continue
        res, _ = residuals(op, r.values[:nev], r.vectors, op.nrows)
        worst = max(worst, res)
    return worst


def batch_cpu_baseline(host, sample, results, ns, nnzs, nev, ncv, tol, maxiter, budget_s=20.0):
    """The CPU oracle's full solves of a sample of the mix (smallest first, until
    about budget_s seconds are spent), extrapolated to the whole mix with a per-solve
    cost model: Lanczos steps x (nnz + 3 ncv n / 2), the SpMV and CGS work of a
    step, using each matrix's step count from the GPU run."""
    import oracle as O

    threads = min(16, os.cpu_count() or 1)
    cost = [r.info.spmv * (nnzs[i] + 1.5 * ncv * ns[i]) for i, r in enumerate(results)]
    t_cpu, c_done, done = 0.0, 0.0, 0
    for i in sorted(sample, key=lambda k: ns[k]):
        rp, col, val = host[i]
        t = time.perf_counter()
        O.ds_solver(O.Csr(rp, col, val), nev, ncv, tol=tol, maxiter=maxiter, nthreads=threads)
        t_cpu += time.perf_counter() - t
        c_done += cost[i]
        done += 1
        if t_cpu > budget_s:
            break
    total = t_cpu / c_done * sum(cost)
    return {"value": len(results) / total, "unit": "eigensolves/s", "cores": threads, "kind": "port",
            "sample": f"oracle full solves of {done} matrices of the mix at evenly spaced size ranks, "
                      f"smallest first ({t_cpu:.1f} s on {threads} threads), scaled to all "
                      f"{len(results)} by Lanczos steps x (nnz + 1.5 ncv n)"}


def sell_where_even(H, ops, nnzs, limit=1.1):
    """The sliced-ELL SpMV form on every operator whose 64-row chunks hold at
    most `limit` x nnz entries (even row lengths); returns how many took it."""
    taken = 0
    for op, z in zip(ops, nnzs):
        try:
            op.set_sell(True)
        except H.HLancError:  # refused: rows too uneven
            continue
        if op.sell()[1] > limit * z:
            op.set_sell(False)
        else:
            taken += 1
    return taken


def run_batch(args, H, torch):
    from hlanc_amd import workloads as W

    nev, ncv, count = 8, 24, args.count
    t0 = time.perf_counter()
    mix = W.batch_mix(count, seed=0)
    mixed = args.basis == "fp32"
    ops = [H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=m[1][0], col=m[1][1], val=m[1][2]), fp32_values=mixed)
           for m in mix]
    ns = [op.nrows for op in ops]
    nnzs = [len(m[1][1]) for m in mix]
    nnz = sum(nnzs)
    # SpMV form per operator (the batched and the plain path share the operators)
    n_sell = 0 if args.no_autotune or mixed else sell_where_even(H, ops, nnzs)
    # host copies of 16 matrices at evenly spaced size ranks, for the CPU oracle's sample
    sample = sorted(range(count), key=lambda i: ns[i])[:: max(1, count // 16)]
    host = {i: mix[i][1] for i in sample}
    del mix
    log(f"workload batch: {count} matrices, sum n {sum(ns)}, sum nnz {nnz}, built in "
        f"{time.perf_counter() - t0:.1f} s; {n_sell} operators in the sliced-ELL form")
    torch.cuda.init()
    kw = dict(tol=args.tol, maxiter=args.maxiter, use_mfma=args.mfma)
    prec = basis_prec(H, args.basis)

    vec_out = [H.pinned_empty(op.nrows, ncv) for op in ops]  # page-locked outputs, allocated once

    def batched():
        return H.ds_solve_batch(ops, nev, ncv, max_active=args.max_active, vectors_out=vec_out, basis_prec=prec,
                                **kw)

    def plain():  # the plain FP64 ds_solver_gpu path, one solve after another
        out = []
        for op, v in zip(ops, vec_out):
            out.append(H.ds_solver_gpu(op.nrows, nev, ncv, op, vectors_out=v, basis_prec=H.HL_PREC_F64, **kw))
        return out

    for i in range(args.warmup):
        batched()
    walls_b, walls_p = [], []
    for i in range(args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        rb = batched()
        walls_b.append(time.perf_counter() - t)
        log(f"batched step {i}: {walls_b[-1]:.3f} s")
    if args.no_plain:
        rp, walls_p = None, [float("nan")]
    else:
        t = time.perf_counter()
        rp = plain()
        walls_p.append(time.perf_counter() - t)
        log(f"plain one-at-a-time: {walls_p[-1]:.3f} s")
    classes = None
    if args.profile:
        import ctypes

        H.ds_solve_batch(ops, nev, ncv, max_active=args.max_active, flags=H.HL_FLAG_PROFILE, basis_prec=prec, **kw)
        pr = H.Profile()
        H.lib().hl_profile_last(ctypes.byref(pr))
        classes = {H.PROFILE_CLASSES[k]: {"launches": int(pr.launches[k]), "seconds": pr.seconds[k],
                                          "GBps": pr.bytes[k] / pr.seconds[k] / 1e9 if pr.seconds[k] else None}
                   for k in range(len(H.PROFILE_CLASSES)) if pr.launches[k]}
        log("profiled batch: " + ", ".join(f"{k} {v['launches']}x {v['seconds']:.3f}s" for k, v in classes.items()))
    fails = sum(r.status != 0 for r in rb)
    worst = check_pairs(ops, rb, [nev] * count)  # before the plain run's vectors overwrite the buffers
    same = None  # bitwise identity with the plain solves holds for the same basis mode (FP64)
    if rp is not None and args.basis == "fp64":
        same = all(np.array_equal(a.values, b.values) and a.status == b.ret for a, b in zip(rb, rp))
    err64 = None
    if rp is not None:  # each batched value against the plain FP64 solve's
        err64 = max(float(np.max(np.abs(a.values - b.values) / np.abs(b.values))) for a, b in zip(rb, rp))
    lat = [r.t_done for r in rb]
    value = count / float(np.mean(walls_b))
    plain_value = count / walls_p[0]
    log(f"batched {value:.1f} solves/s vs plain {plain_value:.1f} solves/s ({value / plain_value:.2f}x); "
        f"failures {fails}; bitwise same as plain: {same}; max rel value error vs plain FP64 {err64}; "
        f"max rel residual {worst:.2e}")
    restarts = [r.info.restarts for r in rb]
    cpu = None
    if not args.no_cpu:
        cpu = batch_cpu_baseline(host, sample, rb, ns, nnzs, nev, ncv, args.tol, args.maxiter)
        log(f"cpu baseline: {cpu}")
    return {
        "metric": "eigensolves/s at rel.resid<=1e-10 (+ p50/p99 latency)",
        "value": value, "unit": "eigensolves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * float(np.mean(walls_b)), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if mixed else "f64", "data": "synthetic (seeded generator)",
        "config": {"workload": f"{count} uneven matrices, n log-uniform [1e3, 2e5], stencils + Erdos-Renyi",
                   "sum_n": int(sum(ns)), "sum_nnz": int(nnz), "nev": nev, "ncv": ncv, "tol": args.tol,
                   "max_active": args.max_active, "basis": BASIS_LABEL[args.basis],
                   "sliced_ell_operators": n_sell},
        "plain_value": plain_value, "plain_basis": "fp64", "speedup_vs_plain": value / plain_value,
        "max_rel_value_err_vs_fp64": err64,
        "latency_s": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
        "bitwise_same_as_plain": same, "failures": int(fails), "max_rel_resid": worst,
        "restarts": {"min": int(min(restarts)), "median": float(np.median(restarts)), "max": int(max(restarts))},
        "kernel_classes": classes,
        "cpu_baseline": cpu,
    }


def run_serve(args, H, torch):
    from hlanc_amd import workloads as W

    t0 = time.perf_counter()
    pool = W.serve_mix(args.pool, seed=0)
    mixed = args.basis == "fp32"
    prec = basis_prec(H, args.basis)
    ops = [H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=m[1][0], col=m[1][1], val=m[1][2]), fp32_values=mixed)
           for m in pool]
    nnzs = [len(m[1][1]) for m in pool]
    nnz = sum(nnzs)
    del pool
    n_sell = 0 if args.no_autotune or mixed else sell_where_even(H, ops, nnzs)
    log(f"workload serve: pool of {len(ops)} matrices, sum n {sum(op.nrows for op in ops)}, sum nnz {nnz}, "
        f"built in {time.perf_counter() - t0:.1f} s; {n_sell} operators in the sliced-ELL form")
    torch.cuda.init()
    rng = np.random.default_rng(args.seed)
    R = args.requests
    gaps = rng.exponential(1.0, R)  # one request sequence, scaled to each offered rate
    pick = rng.integers(0, len(ops), R)
    nevs = rng.choice([4, 8, 16, 32], R)
    req_ops = [ops[k] for k in pick]
    H.ds_solve_batch(req_ops[:2], list(nevs[:2]), list(3 * nevs[:2]), tol=args.tol, maxiter=args.maxiter,
                     basis_prec=prec)

    def serve(arrivals, max_active, budget):
        t = time.perf_counter()
        res = H.ds_solve_batch(req_ops, list(nevs), list(3 * nevs), max_active=max_active, hbm_budget=budget,
                               tol=args.tol, maxiter=args.maxiter, arrivals=arrivals, vectors=True,
                               basis_prec=prec)
        return res, time.perf_counter() - t

    rates = [float(x) for x in str(args.rates).split(",")]
    per_rate, worst, same, fails = [], 0.0, True, 0
    for rate in rates:
        arrivals = np.cumsum(gaps) / rate
        rs, wall_s = serve(arrivals, args.max_active, args.hbm_budget)
        worst = max(worst, check_pairs(req_ops, rs, list(nevs)))
        for r in rs:
            r.vectors = None  # host memory: the stream's vectors are checked, then dropped
        rf, wall_f = serve(arrivals, 1, 0.0)  # first come, first served, one at a time
        for r in rf:
            r.vectors = None
        lat_s = [r.latency for r in rs]
        lat_f = [r.latency for r in rf]
        same &= all(np.array_equal(a.values, b.values) and a.status == b.status for a, b in zip(rs, rf))
        fails += sum(r.status != 0 for r in rs)
        rec = {"offered_req_per_s": rate, "value": R / wall_s,
               "latency_s": {"p50": percentile(lat_s, 50), "p99": percentile(lat_s, 99)},
               "fcfs": {"value": R / wall_f, "latency_s": {"p50": percentile(lat_f, 50),
                                                          "p99": percentile(lat_f, 99)}}}
        per_rate.append(rec)
        log(f"rate {rate}: served {rec['value']:.2f} solves/s p50 {rec['latency_s']['p50']:.3f} s p99 "
            f"{rec['latency_s']['p99']:.3f} s; FCFS {rec['fcfs']['value']:.2f} solves/s p50 "
            f"{rec['fcfs']['latency_s']['p50']:.3f} s p99 {rec['fcfs']['latency_s']['p99']:.3f} s")
    log(f"same results served and FCFS: {same}; failures {fails}; max rel residual {worst:.2e}")
    top = per_rate[-1]
    return {
        "metric": "eigensolves/s at rel.resid<=1e-10 (+ p50/p99 latency)",
        "value": top["value"], "unit": "eigensolves/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * R / top["value"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if mixed else "f64", "data": "synthetic (seeded generator)",
        "config": {"workload": f"Poisson stream, {R} requests, pool of {len(ops)} matrices n log-uniform "
                               "[4e3, 2e6], nev in {4,8,16,32}, ncv = 3 nev",
                   "offered_req_per_s": rates, "tol": args.tol, "max_active": args.max_active,
                   "hbm_budget_bytes": args.hbm_budget, "basis": BASIS_LABEL[args.basis],
                   "sliced_ell_operators": n_sell},
        "latency_s": top["latency_s"], "fcfs": top["fcfs"], "rates": per_rate,
        "same_results_as_fcfs": bool(same), "failures": int(fails), "max_rel_resid": worst,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="chunglu", choices=sorted(WORKLOADS) + ["batch", "serve"])
    ap.add_argument("--count", type=int, default=512, help="batch: number of matrices")
    ap.add_argument("--max-active", type=int, default=0, help="batch/serve: solves in flight (0: no cap)")
    ap.add_argument("--hbm-budget", type=float, default=64e9, help="serve: workspace budget in bytes")
    ap.add_argument("--rates", default="2,200", help="serve: offered requests/s, comma-separated")
    ap.add_argument("--requests", type=int, default=100, help="serve: requests in the stream")
    ap.add_argument("--pool", type=int, default=32, help="serve: distinct matrices")
    ap.add_argument("--seed", type=int, default=0, help="serve: arrival stream seed")
    ap.add_argument("--profile", action="store_true", help="batch: one more profiled batched pass")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default SpMV work-item shape")
    ap.add_argument("--no-plain", action="store_true", help="batch: skip the one-at-a-time comparison")
    ap.add_argument("--mfma", type=int, choices=(0, 1), default=None,
                    help="restart/Ritz V.Q on the MFMA (1) or FMA (0) kernel; default: the library's")
    ap.add_argument("--basis", default="proj32", choices=("fp64", "proj32", "fp32"),
                    help="proj32: FP32 basis copy for the projection passes; fp32: FP32-basis phase, then FP64")
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

    if args.workload in ("batch", "serve"):
        if args.maxiter == 1000:
            args.maxiter = 3000
        out = (run_batch if args.workload == "batch" else run_serve)(args, H, torch)
        print(json.dumps(out), flush=True)
        return

    gen, kw, nev, ncv, label = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    rp, col, val = getattr(W, gen)(**kw)
    n = len(rp) - 1
    log(f"workload {args.workload}: n={n} nnz={len(val)} generated in {time.perf_counter() - t0:.1f} s")
    mixed = args.basis == "fp32"
    prec = basis_prec(H, args.basis)
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val), fp32_values=mixed)
    torch.cuda.init()
    items = None
    if not args.no_autotune:
        nc, rc_, t_best = op.autotune(reps=10)
        sw, npairs = op.slices()
        sell_on, sell_entries = op.sell()
        items = {"nnz_cap": nc, "rows_cap": rc_, "slice_columns": sw, "virtual_rows": npairs,
                 "sliced_ell": sell_on, "sliced_ell_entries": sell_entries}
        form = (f"sliced-ELL ({sell_entries} entries)" if sell_on else
                f"items of <= {nc} nnz / <= {rc_} rows, column slices of {sw} ({npairs} virtual rows)")
        log(f"SpMV autotune: {form}, {t_best * 1e6:.1f} us per launch")

    # timed solves write their vectors to device memory (the result stays in HBM);
    # one more solve into page-locked host memory gives the rate with the PCIe copy
    vec_dev = torch.empty((ncv, n), dtype=torch.float64, device="cuda").T
    vec_out = H.pinned_empty(n, ncv)
    for i in range(args.warmup):
        r = H.ds_solver_gpu(n, nev, ncv, op, tol=args.tol, maxiter=args.maxiter, vectors_out=vec_dev,
                            use_mfma=args.mfma, basis_prec=prec)
        log(f"warmup {i}: ret {r.ret} restarts {r.info.restarts} spmv {r.info.spmv} {r.info.seconds:.3f} s")

    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        r = H.ds_solver_gpu(n, nev, ncv, op, tol=args.tol, maxiter=args.maxiter, vectors_out=vec_dev,
                            use_mfma=args.mfma, basis_prec=prec)
        lat.append(time.perf_counter() - t)
        log(f"solve {i}: ret {r.ret} restarts {r.info.restarts} spmv {r.info.spmv} "
            f"(fp32 phase {r.info.spmv_f32}) {lat[-1]:.3f} s")
        if r.ret != 0:
            raise SystemExit(f"solve {i} did not converge (ret {r.ret})")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    value = args.steps / wall

    t = time.perf_counter()
    rh = H.ds_solver_gpu(n, nev, ncv, op, tol=args.tol, maxiter=args.maxiter, vectors_out=vec_out,
                         use_mfma=args.mfma, basis_prec=prec)
    t_pcie = time.perf_counter() - t
    log(f"solve into page-locked host memory: {t_pcie:.3f} s ({8.0 * n * ncv / 1e9:.2f} GB of vectors over PCIe)")
    resid, ortho = residuals(op, r.values, r.vectors.cpu().numpy(), n)
    if not np.array_equal(rh.vectors, r.vectors.cpu().numpy()) or not np.array_equal(rh.values, r.values):
        raise SystemExit("host-output solve differs from the device-output solve")
    log(f"max rel residual {resid:.2e}, orthogonality {ortho:.2e}")
    err64 = None
    if args.basis != "fp64":  # the result's values against a plain FP64 solve's (untimed)
        r64 = H.ds_solver_gpu(n, nev, ncv, op, tol=args.tol, maxiter=args.maxiter, use_mfma=args.mfma,
                              basis_prec=H.HL_PREC_F64)
        err64 = float(np.max(np.abs(r.values - r64.values) / np.abs(r64.values)))
        log(f"max rel value error vs FP64 solve {err64:.2e} (FP64 solve {r64.info.seconds:.3f} s, "
            f"{r64.info.spmv} SpMV)")

    # per-kernel-class profile of one more solve (HIP events around each launch)
    import ctypes

    o = H.options(args.tol, args.maxiter, use_mfma=args.mfma, basis_prec=prec)
    o.flags = H.HL_FLAG_PROFILE
    values = np.zeros(ncv)
    info = H.Info()
    rc = H.lib().hl_ds_solve_csr(op.h, nev, ncv, values.ctypes.data_as(H._dp), vec_out.ctypes.data_as(H._dp), n,
                                 ctypes.byref(o), ctypes.byref(info))
    pr = H.Profile()
    H.lib().hl_profile_last(ctypes.byref(pr))
    names = H.PROFILE_CLASSES
    classes = {names[k]: {"launches": int(pr.launches[k]), "seconds": pr.seconds[k], "bytes": pr.bytes[k],
                          "GBps": (pr.bytes[k] / pr.seconds[k] / 1e9) if pr.seconds[k] > 0 else None}
               for k in range(len(names)) if pr.launches[k] > 0}
    dom = max((0, 1, 2, 4), key=lambda k: pr.seconds[k])  # kernel classes with algorithmic bytes
    log(f"profile (ret {rc}): " + ", ".join(f"{k} {v['launches']}x {v['seconds']:.3f}s {v['GBps'] or 0:.0f} GB/s"
                                           for k, v in classes.items()))
    achieved = pr.bytes[dom] / pr.seconds[dom] / 1e9
    traffic = None  # measured HBM bytes per dispatch of the dominant kernel (rocprofv3 PMC runs)
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath):
        key = args.workload + ("" if args.basis == "fp64" else "_" + args.basis)
        rec = json.load(open(tpath)).get(key, {}).get(names[dom])
        if rec:
            traffic = rec["fetch_bytes"] + rec["write_bytes"]

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
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": label, "n": n, "nnz": int(len(val)), "nev": nev, "ncv": ncv, "tol": args.tol,
                   "which": "largest algebraic", "basis": BASIS_LABEL[args.basis], "spmv_items": items,
                   "vectors_out": "device memory (HBM)"},
        "latency_s": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
        "solve_s_with_pcie_vectors": t_pcie,
        "max_rel_resid": resid,
        "ritz_orthogonality": ortho,
        "restarts": int(r.info.restarts),
        "spmv_per_solve": int(r.info.spmv),
        "spmv_fp32_phase": int(r.info.spmv_f32),
        "max_rel_value_err_vs_fp64": err64,
        "cgs_passes_per_solve": int(r.info.cgs_passes),
        "spmv": {"GBps": spmv_gbps, "us_per_launch": t_spmv * 1e6, "bytes_per_launch": spmv_bytes,
                 "frac_of_peak": spmv_gbps / HBM_PEAK_GBPS},
        "kernel_classes": classes,
        "roofline": {"kernel": names[dom], "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
