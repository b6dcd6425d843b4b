"""Batched solves (hl_ds_solve_batch): every solve of a batch computes bitwise
what ds_solver_gpu computes for it alone, whatever the batch around it, its
admission order, the concurrency cap, the HBM budget or its arrival time; and
the batch's results match the CPU oracle."""
import numpy as np
import pytest
import scipy.sparse as sp

import hlanc_amd as H
import oracle as O
from hlanc_amd import workloads as W

pytestmark = pytest.mark.gpu


def op_for(rp, col, val):
    return H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))


def arrow(n, heavy, seed):
    """Diagonal plus a dense first row/column and one more row of `heavy`
    entries: long rows split into several chunk items."""
    rng = np.random.default_rng(seed)
    cols = np.arange(1, n)
    far = np.setdiff1d(rng.choice(n, heavy + 1, replace=False), [0, n // 2])[:heavy]
    rows = np.concatenate([np.zeros(n - 1, int), cols, [n // 2] * heavy, far])
    cc = np.concatenate([cols, np.zeros(n - 1, int), far, [n // 2] * heavy])
    vv = np.concatenate([np.full(2 * (n - 1), 0.01), np.full(2 * heavy, 0.02)])
    A = sp.coo_matrix((vv, (rows, cc)), shape=(n, n)).tocsr() + sp.diags(rng.uniform(0, 1, n))
    A.sum_duplicates()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data


def mixed_set():
    """Operators of very different shapes: stencils, power-law graphs with
    heavy rows, an arrow matrix, a small 2-D Laplacian."""
    mats = [W.stencil27(8, seed=1), W.chung_lu_laplacian(1 << 13, max_degree=3000, seed=2),
            W.stencil27(11, seed=3), arrow(20000, 9000, 4), W.laplacian_2d(24),
            W.chung_lu_laplacian(1 << 14, max_degree=12000, seed=5), W.stencil27(6, seed=6)]
    return mats, [op_for(*m) for m in mats]


NEV = [6, 5, 8, 6, 4, 6, 3]
NCV = [18, 16, 24, 20, 12, 20, 10]


@pytest.fixture(scope="module")
def mixed():
    mats, ops = mixed_set()
    alone = [H.ds_solver_gpu(op.nrows, NEV[i], NCV[i], op, tol=1e-10, maxiter=3000, seed=100 + i)
             for i, op in enumerate(ops)]
    return mats, ops, alone


def assert_same(batch, alone, vectors=True):
    for b, a in zip(batch, alone):
        assert b.status == a.ret
        np.testing.assert_array_equal(b.values, a.values)
        if vectors:
            np.testing.assert_array_equal(b.vectors, a.vectors)
        assert b.info.restarts == a.info.restarts and b.info.spmv == a.info.spmv
        assert b.info.cgs_passes == a.info.cgs_passes


def run_batch(ops, **kw):
    return H.ds_solve_batch(ops, NEV[:len(ops)], NCV[:len(ops)], tol=1e-10, maxiter=3000,
                            seeds=[100 + i for i in range(len(ops))], **kw)


def test_batch_is_bitwise_the_standalone_solve(mixed):
    mats, ops, alone = mixed
    assert all(a.ret == 0 for a in alone)
    assert_same(run_batch(ops), alone)


def test_mfma_restarts_in_a_mixed_batch(mixed):
    """Solves whose restart V.Q runs on MFMA next to solves on the FMA kernel:
    each is bitwise its own standalone solve and matches the oracle."""
    mats, ops, _ = mixed
    mf = [i % 2 == 0 for i in range(len(ops))]
    alone = [H.ds_solver_gpu(op.nrows, NEV[i], NCV[i], op, tol=1e-10, maxiter=3000, seed=100 + i,
                             use_mfma=mf[i]) for i, op in enumerate(ops)]
    assert_same(run_batch(ops, use_mfma=mf), alone)
    for i, (rp, col, val) in enumerate(mats):
        if not mf[i]:
            continue
        o = O.ds_solver(O.Csr(rp, col, val), NEV[i], NCV[i], tol=1e-10, maxiter=3000)
        assert o.ret == 0 and alone[i].ret == 0
        np.testing.assert_allclose(alone[i].values, o.values, rtol=1e-10)


def test_batch_matches_oracle(mixed):
    mats, ops, alone = mixed
    res = run_batch(ops, vectors=False)
    for i, (rp, col, val) in enumerate(mats):
        o = O.ds_solver(O.Csr(rp, col, val), NEV[i], NCV[i], tol=1e-10, maxiter=3000)
        assert o.ret == 0 and res[i].status == 0
        np.testing.assert_allclose(res[i].values, o.values, rtol=1e-10)


@pytest.mark.parametrize("max_active", [1, 2, 3])
def test_concurrency_cap_and_queueing(mixed, max_active):
    mats, ops, alone = mixed
    res = run_batch(ops, max_active=max_active)
    assert_same(res, alone)
    # never more than max_active in flight: admissions wait for completions
    events = sorted([(r.t_admit, 1) for r in res] + [(r.t_done, -1) for r in res], key=lambda e: (e[0], e[1]))
    live = peak = 0
    for _, d in events:
        live += d
        peak = max(peak, live)
    assert peak <= max_active


def test_hbm_budget_admits_one_at_a_time(mixed):
    mats, ops, alone = mixed
    res = run_batch(ops, hbm_budget=1.0)  # smaller than any workspace: one solve in flight
    assert_same(res, alone)
    for a, b in zip(res, res[1:]):
        assert b.t_admit >= a.t_done


def test_values_only_batch(mixed):
    mats, ops, alone = mixed
    res = run_batch(ops, vectors=False)
    assert_same(res, alone, vectors=False)
    assert all(r.vectors is None for r in res)


def test_arrivals_delay_admission(mixed):
    mats, ops, alone = mixed
    arr = [0.0, 0.0, 0.05, 0.05, 0.1, 0.3, 0.3]
    res = run_batch(ops, arrivals=arr)
    assert_same(res, alone)
    for r, a in zip(res, arr):
        assert r.t_admit >= a and r.t_done >= r.t_admit
        assert r.latency > 0


@pytest.mark.parametrize("kind", ["stencil", "arrow"])
def test_same_operator_many_times_and_seeds(kind):
    # one operator shared by every solve of the launch; the arrow matrix's long
    # rows are split into chunks whose partial sums are per solve
    rp, col, val = W.stencil27(9, seed=12) if kind == "stencil" else arrow(20000, 15000, 7)
    op = op_for(rp, col, val)
    seeds = list(range(12))
    res = H.ds_solve_batch([op] * 12, 6, 18, tol=1e-10, maxiter=2000, seeds=seeds)
    for s, r in zip(seeds, res):
        a = H.ds_solver_gpu(op.nrows, 6, 18, op, tol=1e-10, maxiter=2000, seed=s)
        np.testing.assert_array_equal(r.values, a.values)
        np.testing.assert_array_equal(r.vectors, a.vectors)


def test_failures_and_bad_items_stay_local():
    rp, col, val = W.stencil27(9, seed=13)
    op = op_for(rp, col, val)
    rp2, col2, val2 = W.laplacian_2d(20)
    op2 = op_for(rp2, col2, val2)
    # item 1 cannot converge in 1 restart; item 2 has ncv > n (refused)
    res = H.ds_solve_batch([op, op, op2, op2], [6, 6, 4, 4], [18, 18, 500, 12], tol=1e-10,
                           maxiter=[2000, 1, 2000, 2000])
    a = H.ds_solver_gpu(op.nrows, 6, 18, op, tol=1e-10, maxiter=2000)
    b = H.ds_solver_gpu(op2.nrows, 4, 12, op2, tol=1e-10, maxiter=2000)
    assert res[0].status == 0 and res[1].status == 1 and res[2].status == -1 and res[3].status == 0
    np.testing.assert_array_equal(res[0].values, a.values)
    np.testing.assert_array_equal(res[3].values, b.values)
    assert res[1].info.restarts == 1


def test_profiled_batch_counts_every_class(mixed):
    import ctypes

    mats, ops, alone = mixed
    res = run_batch(ops, flags=H.HL_FLAG_PROFILE)
    assert_same(res, alone)
    pr = H.Profile()
    H.lib().hl_profile_last(ctypes.byref(pr))
    steps = max(r.info.spmv for r in res)
    # SpMV once, the two CGS passes and both DGKS correction launches per
    # lockstep step; a finalize kernel after each of the four passes
    assert pr.launches[0] >= steps and pr.launches[1] >= 2 * steps and pr.launches[4] >= 2 * steps
    assert pr.launches[3] >= 4 * steps
    assert pr.bytes[0] > 0 and pr.bytes[1] > 0 and pr.bytes[4] > 0 and pr.launches[2] > 0


def test_batch_with_and_without_graphs(mixed):
    mats, ops, alone = mixed
    a = run_batch(ops)
    b = run_batch(ops, flags=H.HL_FLAG_NO_GRAPH)
    assert_same(a, alone)
    assert_same(b, alone)
    assert all(r.info.graph_sweeps == 0 for r in b)


def test_many_similar_solves():
    # 24 solves of similar size in every launch; each result must equal its
    # standalone solve
    mats = [W.erdos_renyi(3000 + 37 * i, 12 + i % 5, seed=200 + i) for i in range(24)]
    ops = [op_for(*m) for m in mats]
    res = H.ds_solve_batch(ops, 6, 18, tol=1e-10, maxiter=2000, seeds=list(range(24)))
    for i, (op, r) in enumerate(zip(ops, res)):
        a = H.ds_solver_gpu(op.nrows, 6, 18, op, tol=1e-10, maxiter=2000, seed=i)
        assert r.status == a.ret == 0
        np.testing.assert_array_equal(r.values, a.values)
        np.testing.assert_array_equal(r.vectors, a.vectors)


@pytest.mark.parametrize("engines", [1, 2, 5])
def test_engine_count_changes_no_bits(mixed, engines):
    """A batch shares its queue among several engines (streams and host
    threads; tuning key "batch_engines", default 3): each solve's result is
    bitwise its standalone one whatever engine serves it, and the caps hold
    over all engines."""
    mats, ops, alone = mixed
    old = H.tune_get("batch_engines")
    H.tune_set("batch_engines", engines)
    try:
        assert_same(run_batch(ops), alone)
        res = run_batch(ops, max_active=2)
    finally:
        H.tune_set("batch_engines", old)
    assert_same(res, alone)
    events = sorted([(r.t_admit, 1) for r in res] + [(r.t_done, -1) for r in res], key=lambda e: (e[0], e[1]))
    live = peak = 0
    for _, d in events:
        live += d
        peak = max(peak, live)
    assert peak <= 2
