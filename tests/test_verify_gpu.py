"""The library's own residual check (hl_info.max_rel_resid, per-pair
residuals) and the Rayleigh-quotient refinement (HL_REFINE_RQ).

After the Ritz-vector step the library forms A v_i with the FP64 SpMV (or the
host operator) and reports ||A v_i - l_i v_i|| / (||A||_est ||v_i||), where
||A||_est is the largest |eigenvalue| of the final tridiagonal.  These tests
recompute that quantity on the host with scipy from the returned vectors and
values (agreement within 1e-13 absolute on the ratio, far below the 1e-10 bar
it is compared with), check that ||A||_est bounds ||A||_2 from below, and that
the refined values are the host Rayleigh quotients v^T A v / v^T v, with
residuals no larger than the unrefined ones (the Rayleigh quotient minimises
||A v - l v|| over l)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import hlanc_amd as H
from hlanc_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10


def make(gen):
    rp, col, val = gen()
    A = sp.csr_matrix((val, col, rp))
    return A, H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))


def host_resid(A, values, vectors, anorm):
    V = np.asarray(vectors[:, : len(values)])
    R = A @ V - V * values[None, :]
    return np.linalg.norm(R, axis=0) / (anorm * np.linalg.norm(V, axis=0))


CASES = {
    "stencil27_12": (lambda: W.stencil27(12, seed=7), 8, 24),
    "chunglu_16k": (lambda: W.chung_lu_laplacian(1 << 14, max_degree=2000, seed=11), 8, 24),
    "er_5000": (lambda: W.erdos_renyi(5000, 10, seed=2), 6, 20),
}


@pytest.mark.parametrize("prec", [H.HL_PREC_F64, H.HL_PREC_F64_PROJ32])
@pytest.mark.parametrize("name", sorted(CASES))
def test_library_residuals_match_host(name, prec):
    gen, nev, ncv = CASES[name]
    A, op = make(gen)
    r = H.ds_solver_gpu(A.shape[0], nev, ncv, op, tol=TOL, maxiter=3000, basis_prec=prec)
    assert r.ret == 0 and r.info.verified == nev
    assert 0.0 <= r.info.max_rel_resid <= TOL
    anorm = r.info.anorm_est
    lam_max = abs(sla.eigsh(A, k=1, which="LM", return_eigenvectors=False, tol=1e-12)[0])
    assert anorm <= lam_max * (1 + 1e-12) and anorm >= 0.999 * lam_max
    ref = host_resid(A, r.values, r.vectors, anorm)
    np.testing.assert_allclose(r.residuals, ref, rtol=0, atol=1e-13)
    assert r.info.max_rel_resid == r.residuals.max()


@pytest.mark.parametrize("name", sorted(CASES))
def test_rayleigh_quotient_refinement(name):
    gen, nev, ncv = CASES[name]
    A, op = make(gen)
    n = A.shape[0]
    a = H.ds_solver_gpu(n, nev, ncv, op, tol=TOL, maxiter=3000, basis_prec=H.HL_PREC_F64_PROJ32)
    b = H.ds_solver_gpu(n, nev, ncv, op, tol=TOL, maxiter=3000, basis_prec=H.HL_PREC_F64_PROJ32,
                        refine=H.HL_REFINE_RQ)
    assert a.ret == b.ret == 0
    # the same solve up to the values: the vectors are bitwise equal
    np.testing.assert_array_equal(a.vectors, b.vectors)
    V = np.asarray(b.vectors[:, :nev])
    rq = np.einsum("ij,ij->j", V, A @ V) / np.einsum("ij,ij->j", V, V)
    np.testing.assert_allclose(b.values, rq, rtol=1e-14)
    shift = np.abs(b.values - a.values) / np.abs(a.values)
    assert shift.max() <= b.info.rq_shift + 4 * np.finfo(float).eps  # |theta + delta - theta| rounds
    assert b.info.rq_shift < 1e-11
    assert np.all(b.residuals <= a.residuals * (1 + 1e-9) + 1e-16)
    np.testing.assert_allclose(b.residuals, host_resid(A, b.values, b.vectors, b.info.anorm_est), rtol=0,
                               atol=1e-13)


def test_batch_items_report_the_standalone_residuals():
    mats = [make(CASES[k][0]) for k in sorted(CASES)]
    nevs = [CASES[k][1] for k in sorted(CASES)]
    ncvs = [CASES[k][2] for k in sorted(CASES)]
    ops = [m[1] for m in mats]
    alone = [H.ds_solver_gpu(op.nrows, nevs[i], ncvs[i], op, tol=TOL, maxiter=3000, seed=3 + i,
                             basis_prec=H.HL_PREC_F64_PROJ32, refine=H.HL_REFINE_RQ) for i, op in enumerate(ops)]
    batch = H.ds_solve_batch(ops, nevs, ncvs, tol=TOL, maxiter=3000, seeds=[3 + i for i in range(len(ops))],
                             basis_prec=H.HL_PREC_F64_PROJ32, refine=H.HL_REFINE_RQ)
    values_only = H.ds_solve_batch(ops, nevs, ncvs, tol=TOL, maxiter=3000, seeds=[3 + i for i in range(len(ops))],
                                   basis_prec=H.HL_PREC_F64_PROJ32, refine=H.HL_REFINE_RQ, vectors=False)
    for a, b, c in zip(alone, batch, values_only):
        assert a.ret == 0 and b.status == 0 and c.status == 0
        np.testing.assert_array_equal(b.values, a.values)
        np.testing.assert_array_equal(b.residuals, a.residuals)
        np.testing.assert_array_equal(c.values, a.values)  # checked without a copy-out
        np.testing.assert_array_equal(c.residuals, a.residuals)
        assert b.info.max_rel_resid == a.info.max_rel_resid <= TOL


def test_no_verify_flag_and_failures_report_no_residual():
    A, op = make(CASES["stencil27_12"][0])
    n = A.shape[0]
    r = H.ds_solver_gpu(n, 8, 24, op, tol=TOL, maxiter=3000, flags=H.HL_FLAG_NO_VERIFY)
    assert r.ret == 0 and r.info.max_rel_resid == -1.0 and r.info.verified == 0
    assert np.isnan(r.residuals).all()
    f = H.ds_solver_gpu(n, 8, 24, op, tol=TOL, maxiter=1)
    assert f.ret == 1 and f.info.max_rel_resid == -1.0 and f.info.verified == 0
    with pytest.raises(H.HLancError):  # refinement needs the check
        H.ds_solver_gpu(n, 8, 24, op, tol=TOL, flags=H.HL_FLAG_NO_VERIFY, refine=H.HL_REFINE_RQ)


def test_host_operator_residual_check():
    rp, col, val = W.stencil5(20, seed=1)  # seeded coefficients: a simple spectrum
    A = sp.csr_matrix((val, col, rp))

    def apply(x, y):
        y[:] = A @ x

    r = H.ds_solver_gpu(400, 4, 12, apply, tol=TOL, maxiter=2000, refine=H.HL_REFINE_RQ)
    assert r.ret == 0 and r.info.verified == 4 and r.info.max_rel_resid <= TOL
    np.testing.assert_allclose(r.residuals, host_resid(A, r.values, r.vectors, r.info.anorm_est), rtol=0,
                               atol=1e-13)
    np.testing.assert_allclose(r.values, np.linalg.eigvalsh(A.toarray())[-4:], rtol=1e-10)
