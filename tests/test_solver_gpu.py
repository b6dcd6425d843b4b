"""End-to-end GPU parity of ds_solver_gpu (HLanc.hpp:1868-1872 over 1649-1850).

Eigenvalues are checked against closed forms, dense eigh and the CPU oracle run
on the same matrix and start vector (1e-10 relative); every returned Ritz pair
must have ||A v - l v|| <= 1e-10 ||A|| (tol 1e-10 runs; machine-eps runs are
tighter), Ritz vectors must be orthonormal, and subspace angles to the exact
eigenspaces are bounded per degenerate cluster."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import hlanc_amd as H
import oracle as O
import refformats as R
from hlanc_amd import workloads as W

pytestmark = pytest.mark.gpu


def op_for(rp, col, val, sym=False, ncols=None):
    return H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val, is_symmetric=sym, ncols=ncols))


def check_pairs(A, vals, vecs, ref, nrmA, tol):
    nev = len(ref)
    assert np.max(np.abs(vals - ref) / np.maximum(np.abs(ref), 1e-300)) <= tol, (vals, ref)
    V = vecs[:, :nev]
    for i in range(nev):
        assert abs(np.linalg.norm(V[:, i]) - 1.0) < 1e-12
        assert np.linalg.norm(A @ V[:, i] - vals[i] * V[:, i]) <= tol * nrmA
    assert np.abs(V.T @ V - np.eye(nev)).max() < 1e-10


def subspace_angle(V, Z):
    """Largest principal angle between span(V) and span(Z) (orthonormal columns)."""
    s = np.linalg.svd(Z.T @ V, compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1, 1)))


def test_simple_pp_closed_form(golden):
    m = H.dcsr_matrix_t("pp", os.path.join(R.GOLDEN, "simple.pp"))
    op = H.dcsrmv_operator_gpu(m)
    r = H.ds_solver_gpu(16, 4, 12, op)
    assert r.ret == 0
    A = R.to_dense(m.row_ptr, m.col, m.val, 16, False)
    check_pairs(A, r.values, r.vectors, golden["simple_eigs"][-4:], golden["simple_eigs"][-1], 1e-13)


def test_simple_as_plain_function_operator(golden):
    # demo/HLanc_demo.cu:9-18 and 164: a bare function used as the operator
    def op_fun_simple(x, y):
        i = np.arange(16)
        y[:] = (np.minimum.outer(i, i) + 1.0) @ x

    r = H.ds_solver_gpu(16, 4, 12, op_fun_simple)
    assert r.ret == 0
    np.testing.assert_allclose(r.values, golden["simple_eigs"][-4:], rtol=1e-13)


@pytest.mark.parametrize("route", ["mm", "bin"])
def test_journals_dense(golden, tmp_path, route):
    m = H.dcsr_matrix_t("mm", os.path.join(R.GOLDEN, "Journals.mtx"))
    if route == "bin":
        f = str(tmp_path / "j.bin")
        m.save_bin(f)
        m = H.dcsr_matrix_t("bin", f)
    op = H.dcsrmv_operator_gpu(m)
    r = H.ds_solver_gpu(124, 4, 12, op)
    assert r.ret == 0
    A = R.to_dense(m.row_ptr, m.col, m.val, 124, m.symmetric())
    ref = golden["journals_eigs"][-4:]
    check_pairs(A, r.values, r.vectors, ref, ref[-1], 1e-12)


@pytest.mark.parametrize("v0", ["dlarnv", "splitmix"])
def test_journals_matches_oracle_same_start_vector(golden, v0):
    rp, col, val, nc, sym = R.journals()
    start = golden["dlarnv_124"] if v0 == "dlarnv" else None
    r = H.ds_solver_gpu(124, 4, 12, op_for(rp, col, val, sym=True), v0=start)
    o = O.ds_solver(O.Csr(rp, col, val, nc, sym), 4, 12, v0=start)
    assert r.ret == o.ret == 0
    np.testing.assert_allclose(r.values, o.values, rtol=1e-12)
    # same Ritz vectors up to sign
    dots = np.abs(np.sum(r.vectors[:, :4] * o.vectors[:, :4], axis=0))
    np.testing.assert_allclose(dots, 1.0, atol=1e-8)


def test_laplacian_32_closed_form_and_clusters():
    rp, col, val = W.laplacian_2d(32)
    A = sp.csr_matrix((val, col, rp))
    r = H.ds_solver_gpu(1024, 10, 30, op_for(rp, col, val), tol=1e-10, maxiter=2000)
    assert r.ret == 0
    lam = W.laplacian_2d_eigs(32)
    check_pairs(A, r.values, r.vectors, lam[-10:], 8.0, 1e-10)
    w, Z = np.linalg.eigh(A.toarray())
    # degenerate pairs: compare subspaces cluster by cluster
    vals = r.values
    i = 0
    while i < 10:
        j = i + 1
        while j < 10 and abs(vals[j] - vals[i]) < 1e-8:
            j += 1
        cl = np.flatnonzero(np.abs(w - vals[i]) < 1e-8)
        if len(cl) == j - i:
            assert subspace_angle(r.vectors[:, i:j], Z[:, cl]) < 1e-6
        i = j


def test_laplacian_256_config0():
    # n = 65,536, nnz = 326,656, nev = 10, ncv = 30, largest algebraic, tol 1e-10
    rp, col, val = W.laplacian_2d(256)
    assert len(val) == 326656
    A = sp.csr_matrix((val, col, rp))
    op = op_for(rp, col, val)
    r = H.ds_solver_gpu(65536, 10, 30, op, tol=1e-10, maxiter=1000)
    assert r.ret == 0, r.info.restarts
    check_pairs(A, r.values, r.vectors, W.laplacian_2d_eigs(256)[-10:], 8.0, 1e-10)
    # the drop-in defaults (tol = eps, 200 restarts) run out of restarts here:
    # return 1 with the outputs untouched, after exactly 200 restarts
    d = H.ds_solver_gpu(65536, 10, 30, op)
    assert d.ret == 1 and d.info.restarts == 200
    assert not d.values.any()


def test_random_spd_dense_and_oracle():
    rng = np.random.default_rng(11)
    n = 2000
    B = sp.random(n, n, density=0.004, random_state=12, data_rvs=lambda k: rng.uniform(-1, 1, k))
    A = (B + B.T + sp.diags(rng.uniform(0, 2, n))).tocsr()
    A.sort_indices()
    r = H.ds_solver_gpu(n, 8, 24, op_for(A.indptr, A.indices, A.data), tol=1e-10, maxiter=3000)
    assert r.ret == 0
    w = np.linalg.eigvalsh(A.toarray())
    check_pairs(A, r.values, r.vectors, w[-8:], np.abs(w).max(), 1e-10)
    o = O.ds_solver(O.Csr(A.indptr, A.indices, A.data), 8, 24, tol=1e-10, maxiter=3000)
    assert o.ret == 0
    np.testing.assert_allclose(r.values, o.values, rtol=1e-10)


def test_stencil27_matches_oracle():
    rp, col, val = W.stencil27(14, seed=3)
    A = sp.csr_matrix((val, col, rp))
    n = A.shape[0]
    r = H.ds_solver_gpu(n, 16, 48, op_for(rp, col, val), tol=1e-10, maxiter=2000)
    o = O.ds_solver(O.Csr(rp, col, val), 16, 48, tol=1e-10, maxiter=2000)
    assert r.ret == 0 and o.ret == 0
    np.testing.assert_allclose(r.values, o.values, rtol=1e-10)
    nrm = np.abs(r.values).max()
    for i in range(16):
        v = r.vectors[:, i]
        assert np.linalg.norm(A @ v - r.values[i] * v) <= 1e-10 * nrm


def test_failure_returns_one_and_leaves_outputs():
    rp, col, val = W.laplacian_2d(32)
    r = H.ds_solver_gpu(1024, 10, 30, op_for(rp, col, val), maxiter=2)
    assert r.ret == 1
    assert np.all(r.values == 0.0) and np.all(r.vectors == 0.0)
    assert r.info.restarts == 2 and r.info.nconv < 10


def test_invariant_subspace_n_equals_ncv():
    rng = np.random.default_rng(3)
    M = rng.normal(size=(12, 12))
    A = sp.csr_matrix(M + M.T)
    r = H.ds_solver_gpu(12, 4, 12, op_for(A.indptr, A.indices, A.data))
    assert r.ret == 0
    np.testing.assert_allclose(r.values, np.linalg.eigvalsh(A.toarray())[-4:], rtol=1e-12, atol=1e-12)


def test_repeatable_bitwise():
    rp, col, val = W.stencil27(10, seed=9)
    op = op_for(rp, col, val)
    r1 = H.ds_solver_gpu(1000, 6, 18, op, tol=1e-10, maxiter=500)
    r2 = H.ds_solver_gpu(1000, 6, 18, op, tol=1e-10, maxiter=500)
    assert r1.ret == r2.ret == 0
    np.testing.assert_array_equal(r1.values, r2.values)
    np.testing.assert_array_equal(r1.vectors, r2.vectors)


@pytest.mark.parametrize("mfma", [True, False])
def test_vq_backends_agree(mfma):
    rp, col, val = W.stencil27(12, seed=5)
    r = H.ds_solver_gpu(1728, 8, 24, op_for(rp, col, val), tol=1e-10, maxiter=1000, use_mfma=mfma)
    o = O.ds_solver(O.Csr(rp, col, val), 8, 24, tol=1e-10, maxiter=1000)
    assert r.ret == 0 and o.ret == 0
    np.testing.assert_allclose(r.values, o.values, rtol=1e-10)


def heavy_row_matrix(gen):
    if gen == "stencil":
        rp, col, val = W.stencil27(13, seed=8)
        return rp, col, val, 8, 24
    if gen == "powerlaw":  # rows of 2k-7k nnz: single-chunk long rows
        rp, col, val = W.chung_lu_laplacian(1 << 15, max_degree=20000, seed=4)
        return rp, col, val, 6, 20
    # an arrow matrix: two rows of > 8192 nnz, split into several chunks
    rng = np.random.default_rng(6)
    n = 30000
    cols = np.arange(1, n)
    far = np.setdiff1d(rng.choice(n, 9001, replace=False), [0, n // 2])[:9000]
    rows = np.concatenate([np.zeros(n - 1, int), cols, [n // 2] * 9000, far])
    cc = np.concatenate([cols, np.zeros(n - 1, int), far, [n // 2] * 9000])
    vv = np.concatenate([np.full(2 * (n - 1), 0.01), np.full(18000, 0.02)])
    A = sp.coo_matrix((vv, (rows, cc)), shape=(n, n)).tocsr() + sp.diags(rng.uniform(0, 1, n))
    A.sum_duplicates()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data, 6, 20


@pytest.mark.parametrize("gen", ["stencil", "powerlaw", "arrow"])
def test_heavy_rows_match_oracle(gen):
    # stream items, single-chunk long rows and multi-chunk long rows in a solve
    rp, col, val, nev, ncv = heavy_row_matrix(gen)
    n = len(rp) - 1
    r = H.ds_solver_gpu(n, nev, ncv, op_for(rp, col, val), tol=1e-10, maxiter=2000)
    o = O.ds_solver(O.Csr(rp, col, val), nev, ncv, tol=1e-10, maxiter=2000)
    assert r.ret == o.ret == 0
    np.testing.assert_allclose(r.values, o.values, rtol=1e-10)
    A = sp.csr_matrix((val, col, rp))
    nrm = np.abs(o.values).max()
    for i in range(nev):
        v = r.vectors[:, i]
        assert np.linalg.norm(A @ v - r.values[i] * v) <= 1e-10 * nrm


@pytest.mark.parametrize("blocks", [1, 7, 256, 2048, 4096])
def test_cgs_grid_cap_does_not_change_the_solve_beyond_rounding(blocks):
    # the CGS grid (and with it the partial-sum order) follows cgs_blocks_max
    rp, col, val = W.stencil27(12, seed=11)
    old = H.tune_get("cgs_blocks_max")
    try:
        H.tune_set("cgs_blocks_max", blocks)
        r = H.ds_solver_gpu(1728, 8, 24, op_for(rp, col, val), tol=1e-10, maxiter=1000)
    finally:
        H.tune_set("cgs_blocks_max", old)
    o = O.ds_solver(O.Csr(rp, col, val), 8, 24, tol=1e-10, maxiter=1000)
    assert r.ret == 0 and o.ret == 0
    np.testing.assert_allclose(r.values, o.values, rtol=1e-10)


def test_graph_replay_matches_launch_by_launch():
    # hundreds of restarts whose sweep shapes repeat: replayed HIP graphs must
    # give the same bits as launching every kernel on its own
    rp, col, val = W.laplacian_2d(48)
    n = len(rp) - 1
    op = op_for(rp, col, val)
    a = H.ds_solver_gpu(n, 8, 24, op, tol=1e-10, maxiter=3000)
    b = H.ds_solver_gpu(n, 8, 24, op, tol=1e-10, maxiter=3000, flags=H.HL_FLAG_NO_GRAPH)
    assert a.ret == b.ret == 0
    assert a.info.graph_sweeps > 0 and b.info.graph_sweeps == 0
    assert a.info.restarts == b.info.restarts
    np.testing.assert_array_equal(a.values, b.values)
    np.testing.assert_array_equal(a.vectors, b.vectors)
    np.testing.assert_allclose(a.values, W.laplacian_2d_eigs(48)[-8:], rtol=1e-10)


def test_pinned_vectors_output_is_the_same_solve():
    rp, col, val = W.stencil27(11, seed=21)
    n = len(rp) - 1
    op = op_for(rp, col, val)
    a = H.ds_solver_gpu(n, 6, 18, op, tol=1e-10, maxiter=1000)
    buf = H.pinned_empty(n + 5, 20)  # larger than needed: ldv = n + 5
    b = H.ds_solver_gpu(n, 6, 18, op, tol=1e-10, maxiter=1000, vectors_out=buf)
    assert a.ret == b.ret == 0
    np.testing.assert_array_equal(a.values, b.values)
    np.testing.assert_array_equal(a.vectors[:, :18], b.vectors[:, :18])
    assert np.shares_memory(b.vectors, buf)
    with pytest.raises(H.HLancError):
        H.ds_solver_gpu(n, 6, 18, op, vectors_out=np.zeros((n, 18)))  # row-major: refused


def test_device_vectors_output_is_the_same_solve():
    import torch

    rp, col, val = W.stencil27(11, seed=21)
    n = len(rp) - 1
    op = op_for(rp, col, val)
    a = H.ds_solver_gpu(n, 6, 18, op, tol=1e-10, maxiter=1000)
    dev = torch.full((20, n + 7), 7.0, dtype=torch.float64, device="cuda").T  # ldv = n + 7, padding kept
    b = H.ds_solver_gpu(n, 6, 18, op, tol=1e-10, maxiter=1000, vectors_out=dev)
    assert a.ret == b.ret == 0
    np.testing.assert_array_equal(a.values, b.values)
    np.testing.assert_array_equal(a.vectors[:, :18], b.vectors[:, :18].cpu().numpy())
    assert b.vectors.data_ptr() == dev.data_ptr()
    full = dev.cpu().numpy()
    assert (full[n:, :] == 7.0).all() and (full[:, 18:] == 7.0).all()  # nothing written past n x ncv
    # the batch takes device buffers per item too
    bb = H.ds_solve_batch([op, op], 6, 18, tol=1e-10, maxiter=1000,
                          vectors_out=[torch.empty((18, n), dtype=torch.float64, device="cuda").T, None])
    np.testing.assert_array_equal(bb[0].vectors.cpu().numpy(), a.vectors[:, :18])
    np.testing.assert_array_equal(bb[1].vectors, a.vectors[:, :18])
    with pytest.raises(H.HLancError):  # row-major device tensor: refused
        H.ds_solver_gpu(n, 6, 18, op, vectors_out=torch.empty((n, 18), dtype=torch.float64, device="cuda"))
