"""Benchmark: eigensolves/s (+ p50/p99 latency, residuals) of the implicitly
restarted Lanczos solver on one MI355X, with the SpMV's GB/s, the roofline of
the dominant kernel and the CPU oracle timed on a bounded sample beside it.

Default workload (one "step" = one full eigensolve): normalized Laplacian of a
symmetric Chung-Lu power-law graph (4,194,304 vertices, exponent 2.1, mean
degree 16, max expected degree 2e4, weights U(0,1], giant component kept),
FP64 CSR; nev = 20, ncv = 60, largest algebraic, tol 1e-10; the FP64 basis with
its FP32 copy for the projection passes, Rayleigh-quotient values.  Every
returned pair is checked by the library (true FP64 residual); a run whose worst
residual exceeds 1e-10 prints its line and exits with status 3.

The default run also reports, as sub-objects of the same line:

* ``batch``: 512 independent symmetric matrices, n log-uniform in [1e3, 2e5]
  (a quarter 2-D 5-point stencils, a quarter 3-D 27-point stencils, half
  Erdos-Renyi graphs of degree 8..32 with weights U(-1,1) and a diagonal
  shift), nev = 8, ncv = 24: all solved as one continuously batched call in
  the FP32-copy mode, against the plain FP64 ds_solver_gpu path one solve at a
  time on the same GPU and matrices; eigensolves/s, p50/p99 completion.
* ``serve``: a Poisson stream of requests submitted in real time to the
  serving loop (hl_serve_*) at 50 and 200 requests/s, drawn from a pool of 32
  matrices with n log-uniform in [4e3, 2e6] (27-point stencils, Chung-Lu
  Laplacians, Erdos-Renyi graphs), nev uniform in {4, 8, 16, 32}, ncv = 3 nev;
  each request's operator uploaded after admission; eigensolves/s and p50/p99
  latency from submission to result.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload chunglu|stencil27|laplace2d|batch|serve]
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
RESID_BAR = 1e-10
WORKLOADS = {
    # name: (generator, kwargs, nev, ncv, label)
    "chunglu": ("chung_lu_laplacian", {"n": 1 << 22, "seed": 0}, 20, 60,
                "chung-lu normalized Laplacian, 2^22 vertices (giant component), mean degree 16, exponent 2.1"),
    "stencil27": ("stencil27", {"m": 128, "seed": 0}, 32, 96, "27-point stencil 128^3, seeded coefficients"),
    "laplace2d": ("laplacian_2d", {"m": 256}, 10, 30, "2-D 5-point Laplacian 256x256 (Dirichlet)"),
}

BASIS_LABEL = {"fp64": "fp64 basis, the reference's DGKS rule",
               "fp32": "two phases: fp32 basis and values to 1e-6, then fp64 to tol",
               "proj32": "fp64 basis + fp32 copy for the projection passes (DGKS correction every step)"}
DTYPE = {"fp64": "f64", "fp32": "f32 phase + f64", "proj32": "f64 (fp32 basis copy for projections)"}


def basis_prec(H, basis):
    return {"fp64": H.HL_PREC_F64, "fp32": H.HL_PREC_F32, "proj32": H.HL_PREC_F64_PROJ32}[basis]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else float("nan")


def host_threads():
    """Threads the CPU baseline uses: the box's CPU share (OMP_NUM_THREADS, else
    the affinity mask), with what the host reports beside it."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", aff) or aff)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, min(share, aff)), {"nproc_affinity": aff, "os_cpu_count": os.cpu_count(),
                                     "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model}


def device_check(op, values, vecs_dev, nev):
    """Cross-check of the library's residual check: max_i ||A v_i - l_i v_i|| /
    max|l| with the GPU SpMV, and max |V^T V - I|."""
    import torch

    V = vecs_dev[:, :nev].T.contiguous()
    AV = torch.empty_like(V)
    for i in range(nev):
        op.spmv(V[i].data_ptr(), AV[i].data_ptr())
    torch.cuda.synchronize()
    lam = torch.tensor(values, device=V.device, dtype=torch.float64)[:, None]
    res = torch.linalg.norm(AV - lam * V, dim=1).max().item() / float(np.max(np.abs(values)))
    ortho = (V @ V.T - torch.eye(nev, device=V.device, dtype=torch.float64)).abs().max().item()
    return res, ortho


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
    """The CPU oracle (oracle/hlanc_oracle.c: ds_solver_mkl restated with OpenMP
    BLAS-2) on a bounded sample: its first restart cycle on the same matrix
    (ncv Lanczos steps + one restart), scaled to a full solve by the number of
    Lanczos steps the GPU solve needed."""
    import oracle as O

    threads, hostinfo = host_threads()
    A = O.Csr(rp, col, val)
    t = time.perf_counter()
    r = O.ds_solver(A, nev, ncv, tol=tol, maxiter=1, nthreads=threads, max_restarts_then_stop=1)
    dt = time.perf_counter() - t
    steps = max(int(r.spmv), 1)
    per_solve = dt * gpu_spmv_per_solve / steps
    return {"value": 1.0 / per_solve, "unit": "eigensolves/s", "cores": threads, "kind": "port",
            "sample": f"oracle first restart cycle ({steps} Lanczos steps, {dt:.1f} s on {threads} threads, the "
                      f"box's CPU share), scaled to the {gpu_spmv_per_solve} steps of the GPU solve",
            "host": hostinfo}


def batch_cpu_baseline(host, sample, results, ns, nnzs, nev, ncv, tol, maxiter, budget_s=20.0):
    """The CPU oracle's full solves of a sample of the mix (smallest first, until
    about budget_s seconds are spent), scaled to the whole mix by a per-solve
    cost model: Lanczos steps x (nnz + 3 ncv n / 2), using each matrix's step
    count from the GPU run."""
    import oracle as O

    threads, hostinfo = host_threads()
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
                      f"{len(results)} by Lanczos steps x (nnz + 1.5 ncv n)", "host": hostinfo}


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


def gen_threads():
    return max(1, min(16, host_threads()[0]))


# ------------------------------------------------------------------ batch

def run_batch(H, torch, count=512, steps=1, warmup=1, basis="proj32", tol=1e-10, maxiter=3000, max_active=0,
              mfma=None, autotune=True, plain=True, cpu=True, profile=False):
    from hlanc_amd import workloads as W

    nev, ncv = 8, 24
    t0 = time.perf_counter()
    mix = W.batch_mix(count, seed=0, threads=gen_threads())
    mixed = basis == "fp32"
    ops = [H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=m[1][0], col=m[1][1], val=m[1][2]), fp32_values=mixed)
           for m in mix]
    ns = [op.nrows for op in ops]
    nnzs = [len(m[1][1]) for m in mix]
    nnz = sum(nnzs)
    # SpMV form per operator (the batched and the plain path share the operators)
    n_sell = sell_where_even(H, ops, nnzs) if autotune and not mixed else 0
    sample = sorted(range(count), key=lambda i: ns[i])[:: max(1, count // 16)]
    host = {i: mix[i][1] for i in sample}
    del mix
    log(f"workload batch: {count} matrices, sum n {sum(ns)}, sum nnz {nnz}, built in "
        f"{time.perf_counter() - t0:.1f} s; {n_sell} operators in the sliced-ELL form")
    torch.cuda.init()
    kw = dict(tol=tol, maxiter=maxiter, use_mfma=mfma)
    prec = basis_prec(H, basis)
    refine = H.HL_REFINE_RQ
    vec_out = [H.pinned_empty(op.nrows, ncv) for op in ops]  # page-locked outputs, allocated once

    def batched():
        return H.ds_solve_batch(ops, nev, ncv, max_active=max_active, vectors_out=vec_out, basis_prec=prec,
                                refine=refine, **kw)

    for _ in range(warmup):
        batched()
    walls_b = []
    for i in range(steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        rb = batched()
        walls_b.append(time.perf_counter() - t)
        log(f"batched step {i}: {walls_b[-1]:.3f} s")
    fails = sum(r.status != 0 for r in rb)
    worst = max((r.info.max_rel_resid for r in rb if r.status == 0), default=float("nan"))
    unchecked = sum(r.status == 0 and r.info.verified == 0 for r in rb)
    vals_b = [r.values.copy() for r in rb]
    lat = [r.t_done for r in rb]
    rp_res, plain_value, err64 = None, None, None
    if plain:  # the plain FP64 ds_solver_gpu path, one solve after another, same page-locked outputs
        t = time.perf_counter()
        rp_res = [H.ds_solver_gpu(op.nrows, nev, ncv, op, vectors_out=v, basis_prec=H.HL_PREC_F64, **kw)
                  for op, v in zip(ops, vec_out)]
        wall_p = time.perf_counter() - t
        plain_value = count / wall_p
        log(f"plain one-at-a-time FP64: {wall_p:.3f} s")
        err64 = max(float(np.max(np.abs(a - b.values) / np.abs(b.values))) for a, b in zip(vals_b, rp_res))
    classes = None
    if profile:
        import ctypes

        H.ds_solve_batch(ops, nev, ncv, max_active=max_active, flags=H.HL_FLAG_PROFILE, basis_prec=prec, **kw)
        pr = H.Profile()
        H.lib().hl_profile_last(ctypes.byref(pr))
        classes = {H.PROFILE_CLASSES[k]: {"launches": int(pr.launches[k]), "seconds": pr.seconds[k],
                                          "GBps": pr.bytes[k] / pr.seconds[k] / 1e9 if pr.seconds[k] else None}
                   for k in range(len(H.PROFILE_CLASSES)) if pr.launches[k]}
        log("profiled batch: " + ", ".join(f"{k} {v['launches']}x {v['seconds']:.3f}s" for k, v in classes.items()))
    value = count / float(np.mean(walls_b))
    log(f"batched {value:.1f} solves/s" + (f" vs plain FP64 {plain_value:.1f} solves/s ({value / plain_value:.2f}x)"
                                          if plain_value else "") +
        f"; failures {fails}; max rel value error vs plain FP64 {err64}; max rel residual (library) {worst:.2e}")
    restarts = [r.info.restarts for r in rb]
    cpu_rec = batch_cpu_baseline(host, sample, rb, ns, nnzs, nev, ncv, tol, maxiter) if cpu else None
    if cpu_rec:
        log(f"batch cpu baseline: {cpu_rec}")
    out = {
        "metric": "eigensolves/s at rel.resid<=1e-10 (+ p50/p99 latency)",
        "value": value, "unit": "eigensolves/s", "ms_per_step": 1e3 * float(np.mean(walls_b)), "steps": steps,
        "warmup": warmup, "dtype": DTYPE[basis],
        "config": {"workload": f"{count} uneven matrices, n log-uniform [1e3, 2e5], stencils + Erdos-Renyi",
                   "sum_n": int(sum(ns)), "sum_nnz": int(nnz), "nev": nev, "ncv": ncv, "tol": tol,
                   "max_active": max_active, "basis": BASIS_LABEL[basis], "values": "Rayleigh quotients",
                   "sliced_ell_operators": n_sell, "vectors_out": "page-locked host memory"},
        "plain_value": plain_value, "plain_basis": "fp64 (the reference's DGKS rule), one solve at a time",
        "speedup_vs_plain": (value / plain_value) if plain_value else None,
        "max_rel_value_err_vs_fp64": err64,
        "latency_s": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
        "failures": int(fails), "max_rel_resid": worst, "unchecked_pairs": int(unchecked),
        "resid_bar_met": bool(fails == 0 and unchecked == 0 and worst <= RESID_BAR),
        "restarts": {"min": int(min(restarts)), "median": float(np.median(restarts)), "max": int(max(restarts))},
        "kernel_classes": classes,
        "cpu_baseline": cpu_rec,
    }
    for op in ops:
        op.close()
    return out


# ------------------------------------------------------------------ serve

def run_serve(H, torch, rates=(50.0, 200.0), requests=48, pool_size=32, seed=0, basis="proj32", tol=1e-10,
              maxiter=3000, hbm_budget=64e9, max_active=0, fcfs=False):
    """Requests submitted in real time to the serving loop at each offered rate
    (one Poisson sequence scaled to the rate); each request uploads its matrix
    after admission and returns its values and per-pair residuals (its Ritz
    vectors are formed and checked on the device, not copied out)."""
    from hlanc_amd import workloads as W

    t0 = time.perf_counter()
    pool = W.serve_mix(pool_size, seed=0, threads=gen_threads())
    arrays = [m[1] for m in pool]
    nnz = sum(len(a[1]) for a in arrays)
    log(f"workload serve: pool of {len(arrays)} matrices, sum n {sum(len(a[0]) - 1 for a in arrays)}, sum nnz {nnz}, "
        f"built in {time.perf_counter() - t0:.1f} s")
    torch.cuda.init()
    prec = basis_prec(H, basis)
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0, requests)
    pick = rng.integers(0, len(arrays), requests)
    nevs = rng.choice([4, 8, 16, 32], requests)
    kw = dict(tol=tol, maxiter=maxiter, basis_prec=prec, refine=H.HL_REFINE_RQ, vectors=False)

    def stream(rate, cap):
        arrivals = np.cumsum(gaps) / rate
        with H.Server(max_active=cap, hbm_budget=hbm_budget, sell_even=True) as srv:
            tickets = []
            start = time.perf_counter()
            for k in range(requests):
                wait = start + arrivals[k] - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
                tickets.append(srv.submit(arrays[pick[k]], int(nevs[k]), 3 * int(nevs[k]), **kw))
            outs = [srv.wait(t) for t in tickets]
            st = srv.stats()
        t_first = min(o.t_submit for o in outs)
        wall = max(o.t_done for o in outs) - t_first
        return outs, wall, st

    with H.Server() as srv:  # warm-up: code objects loaded, allocator warm
        srv.wait(srv.submit(arrays[0], 4, 12, **kw))
    per_rate, worst, fails, unchecked = [], 0.0, 0, 0
    values_by_rate = []
    for rate in rates:
        outs, wall, st = stream(rate, max_active)
        lat = [o.latency for o in outs]
        fails += sum(o.status != 0 for o in outs)
        unchecked += sum(o.status == 0 and o.info.verified == 0 for o in outs)
        worst = max([worst] + [o.info.max_rel_resid for o in outs if o.status == 0])
        values_by_rate.append([o.values for o in outs])
        rec = {"offered_req_per_s": rate, "value": requests / wall,
               "latency_s": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
               "queue_wait_s": {"p50": percentile([o.t_admit - o.t_submit for o in outs], 50),
                                "p99": percentile([o.t_admit - o.t_submit for o in outs], 99)},
               "peak_footprint_gb": max(o.footprint for o in outs) / 1e9}
        if fcfs:  # one request in flight at a time, same stream
            fo, fwall, _ = stream(rate, 1)
            flat = [o.latency for o in fo]
            rec["fcfs"] = {"value": requests / fwall, "latency_s": {"p50": percentile(flat, 50),
                                                                    "p99": percentile(flat, 99)}}
            same = all(np.array_equal(a.values, b.values) for a, b in zip(outs, fo))
            rec["same_results_as_fcfs"] = bool(same)
        per_rate.append(rec)
        log(f"serve at {rate:g} req/s: {rec['value']:.2f} solves/s, latency p50 {rec['latency_s']['p50']:.3f} s "
            f"p99 {rec['latency_s']['p99']:.3f} s" + (f"; FCFS {rec['fcfs']['value']:.2f} solves/s p50 "
                                                     f"{rec['fcfs']['latency_s']['p50']:.3f} s" if fcfs else ""))
    # each result is the same whatever the arrival rate (so whatever ran beside it)
    same_across = all(np.array_equal(a, b) for vr in values_by_rate[1:] for a, b in zip(values_by_rate[0], vr))
    top = per_rate[-1]
    return {
        "metric": "eigensolves/s at rel.resid<=1e-10 (+ p50/p99 latency)",
        "value": top["value"], "unit": "eigensolves/s", "dtype": DTYPE[basis],
        "config": {"workload": f"Poisson stream, {requests} requests per rate, submitted in real time; pool of "
                               f"{len(arrays)} matrices n log-uniform [4e3, 2e6], nev in {{4,8,16,32}}, ncv = 3 nev",
                   "offered_req_per_s": list(rates), "tol": tol, "max_active": max_active,
                   "hbm_budget_bytes": hbm_budget, "basis": BASIS_LABEL[basis], "values": "Rayleigh quotients",
                   "results": "values + per-pair residuals (Ritz vectors formed and checked on the device)",
                   "operators": "uploaded per request after admission (sliced-ELL where rows are even)"},
        "latency_s": top["latency_s"], "rates": per_rate, "same_results_across_rates": bool(same_across),
        "failures": int(fails), "max_rel_resid": worst, "unchecked_pairs": int(unchecked),
        "resid_bar_met": bool(fails == 0 and unchecked == 0 and worst <= RESID_BAR),
    }


# ------------------------------------------------------------ single matrix

def run_single(args, H, torch):
    from hlanc_amd import workloads as W

    gen, kw, nev, ncv, label = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    rp, col, val = getattr(W, gen)(**kw)
    n, nnz = len(rp) - 1, len(val)
    log(f"workload {args.workload}: n={n} nnz={nnz} generated in {time.perf_counter() - t0:.1f} s")
    mixed = args.basis == "fp32"
    prec = basis_prec(H, args.basis)
    refine = H.HL_REFINE_NONE if args.basis == "fp64" else H.HL_REFINE_RQ
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val), fp32_values=mixed)
    torch.cuda.init()
    items = None
    if not args.no_autotune:
        nc, rc_, t_best = op.autotune(reps=10)
        sw, npairs = op.slices()
        sell_on, sell_entries = op.sell()
        stg_on, stg_pieces, stg_entries, stg_bins = op.staged()
        items = {"nnz_cap": nc, "rows_cap": rc_, "slice_columns": sw, "virtual_rows": npairs,
                 "sliced_ell": sell_on, "sliced_ell_entries": sell_entries,
                 "staged_slices": stg_on, "staged_pieces": stg_pieces, "staged_entries": stg_entries,
                 "staged_bins": stg_bins}
        form = (f"sliced-ELL ({sell_entries} entries)" if sell_on else
                f"staged slices of 16384 columns ({stg_pieces} pieces, {stg_entries} entries, {stg_bins} row bins)"
                if stg_on else
                f"items of <= {nc} nnz / <= {rc_} rows, column slices of {sw} ({npairs} virtual rows)")
        log(f"SpMV autotune: {form}, {t_best * 1e6:.1f} us per launch")

    # timed solves write their vectors to device memory (the result stays in HBM);
    # one more solve into page-locked host memory gives the rate with the PCIe copy
    vec_dev = torch.empty((ncv, n), dtype=torch.float64, device="cuda").T
    solve = lambda out, p=prec, rf=refine: H.ds_solver_gpu(n, nev, ncv, op, tol=args.tol, maxiter=args.maxiter,  # noqa
                                                           vectors_out=out, use_mfma=args.mfma, basis_prec=p,
                                                           refine=rf)
    for i in range(args.warmup):
        r = solve(vec_dev)
        log(f"warmup {i}: ret {r.ret} restarts {r.info.restarts} spmv {r.info.spmv} {r.info.seconds:.3f} s")
    lat, worst = [], 0.0
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        r = solve(vec_dev)
        lat.append(time.perf_counter() - t)
        worst = max(worst, r.info.max_rel_resid if r.ret == 0 else float("inf"))
        log(f"solve {i}: ret {r.ret} restarts {r.info.restarts} spmv {r.info.spmv} "
            f"(fp32 phase {r.info.spmv_f32}) {lat[-1]:.3f} s, max rel residual {r.info.max_rel_resid:.2e}")
        if r.ret != 0:
            raise SystemExit(f"solve {i} did not converge (ret {r.ret})")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    value = args.steps / wall

    vec_host = H.pinned_empty(n, ncv)
    t = time.perf_counter()
    rh = solve(vec_host)
    t_pcie = time.perf_counter() - t
    log(f"solve into page-locked host memory: {t_pcie:.3f} s ({8.0 * n * ncv / 1e9:.2f} GB of vectors over PCIe)")
    if not np.array_equal(rh.vectors, r.vectors.cpu().numpy()) or not np.array_equal(rh.values, r.values):
        raise SystemExit("host-output solve differs from the device-output solve")
    del vec_host, rh
    resid_x, ortho = device_check(op, r.values, vec_dev, nev)
    log(f"library max rel residual {r.info.max_rel_resid:.2e} (||A||_est {r.info.anorm_est:.6g}); cross-check "
        f"{resid_x:.2e}; orthogonality {ortho:.2e}; Rayleigh-quotient shift {r.info.rq_shift:.2e}")
    fp64 = None
    if args.basis != "fp64":  # the same solve in the reference's FP64 mode, vectors to HBM as well (untimed)
        t = time.perf_counter()
        r64 = solve(vec_dev, H.HL_PREC_F64, H.HL_REFINE_NONE)
        t64 = time.perf_counter() - t
        err64 = float(np.max(np.abs(r.values - r64.values) / np.abs(r64.values)))
        fp64 = {"seconds": t64, "spmv": int(r64.info.spmv), "restarts": int(r64.info.restarts),
                "max_rel_resid": r64.info.max_rel_resid, "max_rel_value_err": err64,
                "speedup_of_headline": t64 / float(np.mean(lat))}
        log(f"FP64 mode: {t64:.3f} s ({r64.info.spmv} SpMV), max rel value error of the headline vs it {err64:.2e}")

    # per-kernel-class profile of one more solve (HIP events around each launch on the engine's stream)
    import ctypes

    o = H.options(args.tol, args.maxiter, use_mfma=args.mfma, basis_prec=prec, refine=refine)
    o.flags = H.HL_FLAG_PROFILE
    values = np.zeros(ncv)
    info = H.Info()
    vptr, ldv = H._vectors_args(vec_dev)
    rc = H.lib().hl_ds_solve_csr(op.h, nev, ncv, values.ctypes.data_as(H._dp), vptr, ldv, ctypes.byref(o),
                                 ctypes.byref(info))
    if rc < 0:
        raise SystemExit(f"profiled solve failed: {rc} {H.lib().hl_last_error().decode()}")
    del vec_dev
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
    op.close()
    del rp, col, val
    return {
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
        "dtype": DTYPE[args.basis],
        "data": "synthetic (seeded generator)",
        "config": {"workload": label, "n": n, "nnz": int(nnz), "nev": nev, "ncv": ncv,
                   "tol": args.tol, "which": "largest algebraic", "basis": BASIS_LABEL[args.basis],
                   "values": "Ritz values" if refine == H.HL_REFINE_NONE else "Rayleigh quotients",
                   "spmv_items": items, "vectors_out": "device memory (HBM)"},
        "latency_s": {"p50": percentile(lat, 50), "p99": percentile(lat, 99)},
        "solve_s_with_pcie_vectors": t_pcie,
        "max_rel_resid": worst,
        "resid_check": "library: ||A v - l v|| / (||A||_est ||v||) per returned pair, FP64 SpMV",
        "resid_cross_check": resid_x,
        "anorm_est": r.info.anorm_est,
        "rq_shift": r.info.rq_shift,
        "resid_bar_met": bool(worst <= RESID_BAR),
        "ritz_orthogonality": ortho,
        "restarts": int(r.info.restarts),
        "spmv_per_solve": int(r.info.spmv),
        "spmv_fp32_phase": int(r.info.spmv_f32),
        "cgs_passes_per_solve": int(r.info.cgs_passes),
        "fp64_mode": fp64,
        "spmv": {"GBps": spmv_gbps, "us_per_launch": t_spmv * 1e6, "bytes_per_launch": spmv_bytes,
                 "frac_of_peak": spmv_gbps / HBM_PEAK_GBPS},
        "kernel_classes": classes,
        "roofline": {"kernel": names[dom], "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic},
        "cpu_baseline": cpu,
    }


def launch_probes(H):
    """Costs that decide how a Lanczos step is launched: a graph node (empty
    kernel), a grid-wide barrier of a cooperative launch at several grid sizes,
    and the 2-D Laplacian 256^2 solve (nev 10, ncv 30, FP64 basis, 9 launches
    per Lanczos step) in microseconds per step."""
    from hlanc_amd import workloads as W

    out = {"graph_node_us": {}, "grid_barrier_us": {}}
    for blocks in (1, 256, 1024):
        out["graph_node_us"][blocks] = H.bench_graph_nodes(nodes=20, blocks=blocks, reps=100) * 1e6
        log(f"graph node, {blocks} blocks: {out['graph_node_us'][blocks]:.2f} us")
    for blocks in (8, 64, 256, 512, 1024, 2048):
        out["grid_barrier_us"][blocks] = H.bench_grid_barrier(blocks, barriers=1000) * 1e6
        log(f"grid barrier, {blocks} blocks: {out['grid_barrier_us'][blocks]:.2f} us")
    rp, col, val = W.laplacian_2d(256)
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))
    op.autotune(reps=10)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        r = H.ds_solver_gpu(op.nrows, 10, 30, op, tol=1e-10, maxiter=1000, seed=0, basis_prec=H.HL_PREC_F64)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    out["laplace2d_fp64"] = {"seconds": t, "spmv": int(r.info.spmv), "restarts": int(r.info.restarts),
                             "us_per_step": t / r.info.spmv * 1e6}
    log(f"laplace2d fp64: {t * 1e3:.2f} ms, {t / r.info.spmv * 1e6:.1f} us per Lanczos step")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="chunglu", choices=sorted(WORKLOADS) + ["batch", "serve"])
    ap.add_argument("--no-sub", action="store_true", help="headline only: skip the batch and serve sub-objects")
    ap.add_argument("--count", type=int, default=512, help="batch: number of matrices")
    ap.add_argument("--batch-engines", type=int, default=0,
                    help="batch: engines sharing the admission queue (0: the library's default)")
    ap.add_argument("--max-active", type=int, default=0, help="batch/serve: solves in flight (0: no cap)")
    ap.add_argument("--hbm-budget", type=float, default=64e9, help="serve: footprint budget in bytes")
    ap.add_argument("--rates", default=None, help="serve: offered requests/s, comma-separated")
    ap.add_argument("--requests", type=int, default=None, help="serve: requests per rate")
    ap.add_argument("--pool", type=int, default=32, help="serve: distinct matrices")
    ap.add_argument("--seed", type=int, default=0, help="serve: arrival stream seed")
    ap.add_argument("--fcfs", action="store_true", help="serve: also run the stream one request at a time")
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
    ap.add_argument("--kernel-profile", type=int, default=0, metavar="RESTARTS",
                    help="only build the workload, autotune its SpMV and run one solve capped at RESTARTS "
                         "restarts (the command the rocprofv3 summaries in profiles/ run)")
    ap.add_argument("--launch-probes", action="store_true",
                    help="only measure launch costs: graph nodes, grid-wide barriers, and the 2-D Laplacian "
                         "256^2 solve per Lanczos step (one JSON line)")
    ap.add_argument("--spmv-form", default=None, metavar="FORM",
                    help="kernel profile: a fixed SpMV form instead of the autotuner's: "
                         "'slices:WIDTH:NNZCAP:ROWSCAP', 'items:NNZCAP:ROWSCAP', 'sell' or 'staged'")
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

    if args.launch_probes:
        print(json.dumps(launch_probes(H)), flush=True)
        return

    if args.kernel_profile > 0:
        from hlanc_amd import workloads as W

        gen, kw, nev, ncv, _ = WORKLOADS[args.workload]
        rp, col, val = getattr(W, gen)(**kw)
        op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))
        if args.spmv_form:
            f = args.spmv_form.split(":")
            if f[0] == "sell":
                op.set_sell(True)
            elif f[0] == "staged":
                op.set_staged(True)
            else:
                if f[0] == "slices":
                    op.set_slices(int(f[1]))
                    f = f[1:]
                op.set_items(int(f[1]), int(f[2]))
        elif not args.no_autotune:
            op.autotune(reps=10)
        vec = torch.empty((ncv, len(rp) - 1), dtype=torch.float64, device="cuda").T
        r = H.ds_solver_gpu(len(rp) - 1, nev, ncv, op, tol=args.tol, maxiter=args.kernel_profile, vectors_out=vec,
                            basis_prec=basis_prec(H, args.basis), use_mfma=args.mfma)
        torch.cuda.synchronize()
        log(f"kernel profile: {args.workload} ret {r.ret} restarts {r.info.restarts} spmv {r.info.spmv} "
            f"slices {op.slices()} sell {op.sell()} staged {op.staged()}")
        return

    maxiter_many = 3000 if args.maxiter == 1000 else args.maxiter
    if args.workload == "batch":
        if args.batch_engines > 0:
            H.tune_set("batch_engines", args.batch_engines)
        out = run_batch(H, torch, count=args.count, steps=args.steps, warmup=args.warmup, basis=args.basis,
                        tol=args.tol, maxiter=maxiter_many, max_active=args.max_active, mfma=args.mfma,
                        autotune=not args.no_autotune, plain=not args.no_plain, cpu=not args.no_cpu,
                        profile=args.profile)
        out.update({"n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                    "data": "synthetic (seeded generator)"})
    elif args.workload == "serve":
        rates = tuple(float(x) for x in (args.rates or "50,200,500,1000,2000").split(","))
        out = run_serve(H, torch, rates=rates, requests=args.requests or 100, pool_size=args.pool, seed=args.seed,
                        basis=args.basis, tol=args.tol, maxiter=maxiter_many, hbm_budget=args.hbm_budget,
                        max_active=args.max_active, fcfs=args.fcfs)
        out.update({"n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 / out["value"],
                    "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                    "data": "synthetic (seeded generator)"})
    else:
        out = run_single(args, H, torch)
        if args.workload == "chunglu" and not args.no_sub:
            torch.cuda.empty_cache()
            out["batch"] = run_batch(H, torch, count=args.count, steps=1, warmup=1, tol=args.tol,
                                     maxiter=maxiter_many, autotune=not args.no_autotune, cpu=False)
            rates = tuple(float(x) for x in (args.rates or "50,200").split(","))
            out["serve"] = run_serve(H, torch, rates=rates, requests=args.requests or 48, pool_size=args.pool,
                                     tol=args.tol, maxiter=maxiter_many, hbm_budget=args.hbm_budget)
    print(json.dumps(out), flush=True)
    bars = [out.get("resid_bar_met", True)] + [out[k]["resid_bar_met"] for k in ("batch", "serve") if k in out]
    if not all(bars):
        log("a returned pair missed the 1e-10 residual bar")
        sys.exit(3)


if __name__ == "__main__":
    main()
