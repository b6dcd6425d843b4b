"""Pins the CPU oracle (oracle/hlanc_oracle.c) before anything is compared with it.

Each restated part is checked against MKL outputs recorded in tests/golden/ (the
library the reference links), against numpy/scipy, or against closed forms --
checks chosen so that a dropped term, a flipped sign or a transposed operand fails.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import oracle as O
import refformats as R


def test_dlartg_matches_mkl(golden):
    fg, ref = golden["dlartg_fg"], golden["dlartg_csr"]
    for (f, g), (c0, s0, r0) in zip(fg, ref):
        c, s, r = O.dlartg(f, g)
        assert np.isclose(c, c0, rtol=4e-16, atol=0) and np.isclose(s, s0, rtol=4e-16, atol=0)
        assert np.isclose(r, r0, rtol=4e-16, atol=0)
        # the rotation annihilates g
        assert abs(-s * f + c * g) <= 4e-16 * max(abs(f), abs(g))


@pytest.mark.parametrize("tag", ["t5", "t30", "t60", "t96", "deg"])
def test_tridiag_eig_matches_mkl_dsteqr(golden, tag):
    d, e = golden[f"steqr_{tag}_d"], golden[f"steqr_{tag}_e"]
    w_ref, z_ref = golden[f"steqr_{tag}_w"], golden[f"steqr_{tag}_z"]
    info, w, z = O.tridiag_eig(d, e)
    assert info == 0
    n = len(d)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    nrm = np.abs(w_ref).max()
    assert np.all(np.diff(w) >= 0)
    np.testing.assert_allclose(w, w_ref, rtol=0, atol=1e-13 * nrm)
    np.testing.assert_allclose(w, sla.eigh_tridiagonal(d, e, eigvals_only=True), rtol=0, atol=1e-13 * nrm)
    assert np.abs(T @ z - z * w).max() <= 1e-13 * nrm
    assert np.abs(z.T @ z - np.eye(n)).max() <= 1e-13
    if tag != "deg":  # simple spectrum: vectors equal up to sign
        dots = np.abs(np.sum(z * z_ref, axis=0))
        np.testing.assert_allclose(dots, 1.0, atol=1e-10)


def test_tridiag_eig_flags_nan():
    info, _, _ = O.tridiag_eig([1.0, np.nan, 2.0], [0.5, 0.5])
    assert info == 1


def test_householder_qr_matches_mkl_and_numpy(golden):
    a, af_ref, tau_ref = golden["qr_a"], golden["qr_af"], golden["qr_tau"]
    af, tau = O.dgeqr2(a)
    np.testing.assert_allclose(tau, tau_ref, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(af, af_ref, rtol=1e-12, atol=1e-13)
    # Q = I H(0)...H(k-1): numpy.linalg.qr uses the same dgeqrf convention
    m, k = a.shape
    Q = O.dorm2r_rn(af, tau, np.eye(m), k)
    Qn, Rn = np.linalg.qr(a, mode="complete")
    np.testing.assert_allclose(Q, Qn, atol=1e-13)
    np.testing.assert_allclose(np.triu(af)[:k], Rn[:k], atol=1e-12)
    # right application to a wide C, as dorm2r("R","N") does
    c = golden["qr_c"]
    np.testing.assert_allclose(O.dorm2r_rn(af, tau, c, k), golden["qr_cq"], atol=1e-12)
    np.testing.assert_allclose(O.dorm2r_rn(af, tau, c, k), c @ Qn, atol=1e-12)


@pytest.mark.parametrize("tag", ["simple", "journals", "rand"])
def test_spmv_matches_mkl_dcsrmv(golden, tag):
    if tag == "simple":
        rp, col, val, nc, sym = R.simple()
    elif tag == "journals":
        rp, col, val, nc, sym = R.journals()
    else:
        rp, col, val = golden["spmv_rand_rp"], golden["spmv_rand_col"], golden["spmv_rand_val"]
        nc, sym = len(golden["spmv_rand_x"]), False
    x, y_ref = golden[f"spmv_{tag}_x"], golden[f"spmv_{tag}_y"]
    A = O.Csr(rp, col, val, nc, sym)
    y = O.spmv(A, x)
    assert np.linalg.norm(y - y_ref) <= 1e-14 * np.linalg.norm(y_ref)
    y_dense = R.to_dense(rp, col, val, nc, sym) @ x
    assert np.linalg.norm(y - y_dense) <= 1e-14 * np.linalg.norm(y_dense)


def test_start_vector_is_splitmix64():
    lib = O.lib()
    import ctypes

    lib.hor_mix64.restype = ctypes.c_uint64
    lib.hor_mix64.argtypes = [ctypes.c_uint64]
    # first output of splitmix64 from state 0 (published known answer)
    assert lib.hor_mix64(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF
    x = O.start_vector(100000, seed=0)
    assert x.min() >= -1 and x.max() < 1
    assert abs(x.mean()) < 0.01 and abs(x.var() - 1 / 3) < 0.01
    assert not np.array_equal(x, O.start_vector(100000, seed=1))
    np.testing.assert_array_equal(O.start_vector(7, seed=3), O.start_vector(100, seed=3)[:7])


def laplacian_2d(m):
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(m, m))
    I = sp.eye(m)
    return (sp.kron(T, I) + sp.kron(I, T)).tocsr()


def laplacian_eigs(m):
    p = np.arange(1, m + 1)
    c = 2 - 2 * np.cos(p * np.pi / (m + 1))
    return np.sort((c[:, None] + c[None, :]).ravel())


def check_pairs(A, vals, vecs, ref_vals, nrmA, tol=1e-10):
    nev = len(ref_vals)
    assert np.max(np.abs(vals - ref_vals) / np.abs(ref_vals)) <= tol
    for i in range(nev):
        v = vecs[:, i]
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        assert np.linalg.norm(A @ v - vals[i] * v) <= tol * nrmA
    assert np.abs(vecs[:, :nev].T @ vecs[:, :nev] - np.eye(nev)).max() < 1e-10


def test_solver_simple_closed_form(golden):
    rp, col, val, nc, sym = R.simple()
    A = O.Csr(rp, col, val, nc, sym)
    r = O.ds_solver(A, 4, 12)
    assert r.ret == 0
    ref = golden["simple_eigs"][-4:]
    check_pairs(R.to_dense(rp, col, val, nc, sym), r.values, r.vectors, ref, ref[-1], tol=1e-13)


@pytest.mark.parametrize("v0", ["dlarnv", "splitmix"])
def test_solver_journals_dense(golden, v0):
    rp, col, val, nc, sym = R.journals()
    A = O.Csr(rp, col, val, nc, sym)
    start = golden["dlarnv_124"] if v0 == "dlarnv" else None
    r = O.ds_solver(A, 4, 12, v0=start)
    assert r.ret == 0
    ref = golden["journals_eigs"][-4:]
    check_pairs(R.to_dense(rp, col, val, nc, sym), r.values, r.vectors, ref, ref[-1], tol=1e-12)


def test_solver_simple_with_reference_start_vector(golden):
    rp, col, val, nc, sym = R.simple()
    r = O.ds_solver(O.Csr(rp, col, val, nc, sym), 4, 12, v0=golden["dlarnv_16"])
    assert r.ret == 0
    np.testing.assert_allclose(r.values, golden["simple_eigs"][-4:], rtol=1e-13)


def test_solver_laplacian_closed_form():
    L = laplacian_2d(32)
    r = O.ds_solver(O.Csr(L.indptr, L.indices, L.data), 10, 30, tol=1e-10, maxiter=1000)
    assert r.ret == 0
    check_pairs(L, r.values, r.vectors, laplacian_eigs(32)[-10:], 8.0)


def test_solver_random_spd_dense():
    rng = np.random.default_rng(7)
    n = 400
    B = sp.random(n, n, density=0.02, random_state=11, data_rvs=lambda k: rng.uniform(-1, 1, k))
    A = (B + B.T + sp.diags(rng.uniform(0, 3, n))).tocsr()
    r = O.ds_solver(O.Csr(A.indptr, A.indices, A.data), 8, 24, tol=1e-10, maxiter=2000)
    assert r.ret == 0
    w = np.linalg.eigvalsh(A.toarray())
    check_pairs(A, r.values, r.vectors, w[-8:], np.abs(w).max())


def test_solver_symmetric_lower_equals_general():
    L = laplacian_2d(12)
    low = sp.tril(L).tocsr()
    r1 = O.ds_solver(O.Csr(L.indptr, L.indices, L.data), 5, 15, tol=1e-12, maxiter=500)
    r2 = O.ds_solver(O.Csr(low.indptr, low.indices, low.data, sym=True), 5, 15, tol=1e-12, maxiter=500)
    assert r1.ret == r2.ret == 0
    np.testing.assert_allclose(r1.values, r2.values, rtol=1e-12)


def test_failure_returns_one_and_leaves_values():
    L = laplacian_2d(32)
    r = O.ds_solver(O.Csr(L.indptr, L.indices, L.data), 10, 30, maxiter=2)
    assert r.ret == 1 and r.nconv < 10
    assert np.all(r.values == 0.0)


def test_nan_matrix_fails_in_tridiagonal_solver():
    L = laplacian_2d(8)
    data = L.data.copy()
    data[5] = np.nan
    r = O.ds_solver(O.Csr(L.indptr, L.indices, data), 3, 9)
    assert r.ret == 1


def test_full_krylov_space_invariant_subspace():
    # n == ncv: the last Lanczos steps hit an invariant subspace and the DGKS
    # branch zeroes the residual (HLanc.hpp:1459-1463); eigenpairs stay exact.
    rng = np.random.default_rng(3)
    n = 12
    M = rng.normal(size=(n, n))
    A = sp.csr_matrix(M + M.T)
    r = O.ds_solver(O.Csr(A.indptr, A.indices, A.data), 4, 12)
    assert r.ret == 0
    w = np.linalg.eigvalsh(A.toarray())
    np.testing.assert_allclose(r.values, w[-4:], rtol=1e-12, atol=1e-12)


def test_dgks_counts_more_than_one_pass_per_step():
    L = laplacian_2d(32)
    r = O.ds_solver(O.Csr(L.indptr, L.indices, L.data), 10, 30, tol=1e-10, maxiter=1000)
    assert r.cgs > r.spmv  # the DGKS correction fires on most steps
