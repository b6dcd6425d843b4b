"""Mixed precision.  Two modes, both returning FP64 results (eigenvalues
within 1e-10 of the plain FP64 solve and of the CPU oracle, every pair with
||A v - l v|| <= 1e-10 ||A||, orthonormal vectors), and both bitwise the same
alone and inside a batch next to solves of the other modes:

* HL_PREC_F64_PROJ32: the FP64 basis plus an FP32 copy that the projection-only
  Gram-Schmidt passes read; the DGKS correction runs every step.
* HL_PREC_F32: a first phase keeps the Krylov basis in FP32 (and the values in
  FP32 when the operator holds them) to a 1e-6 Ritz estimate; the solve then
  restarts in FP64 from the sum of that phase's wanted Ritz vectors."""
import numpy as np
import pytest
import scipy.sparse as sp

import hlanc_amd as H
import oracle as O
from hlanc_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10


def op_for(rp, col, val, fp32_values=False):
    return H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val), fp32_values=fp32_values)


CASES = {
    "stencil27_12": (lambda: W.stencil27(12, seed=7), 8, 24),
    "laplace2d_48": (lambda: W.laplacian_2d(48), 6, 20),
    "chunglu_16k": (lambda: W.chung_lu_laplacian(1 << 14, max_degree=2000, seed=11), 8, 24),
}


def check_fp64_quality(A, r, nev, ref_values):
    nrm = np.abs(r.values).max()
    np.testing.assert_allclose(r.values, ref_values, rtol=TOL)
    V = r.vectors[:, :nev]
    for i in range(nev):
        assert np.linalg.norm(A @ V[:, i] - r.values[i] * V[:, i]) <= TOL * nrm
    assert np.abs(V.T @ V - np.eye(nev)).max() < 1e-10


@pytest.mark.parametrize("fp32_values", [True, False])
@pytest.mark.parametrize("case", sorted(CASES))
def test_mixed_solve_returns_fp64_results(case, fp32_values):
    gen, nev, ncv = CASES[case]
    rp, col, val = gen()
    A = sp.csr_matrix((val, col, rp))
    n = A.shape[0]
    op = op_for(rp, col, val, fp32_values)
    ref = H.ds_solver_gpu(n, nev, ncv, op, tol=TOL, maxiter=3000)
    mix = H.ds_solver_gpu(n, nev, ncv, op, tol=TOL, maxiter=3000, basis_prec=H.HL_PREC_F32)
    assert ref.ret == 0 and mix.ret == 0
    check_fp64_quality(A, mix, nev, ref.values)
    # the FP32 phase did part of the work, and its values were close already
    assert mix.info.spmv_f32 > 0 and mix.info.spmv > mix.info.spmv_f32
    assert 0.0 <= mix.info.refine_shift < 1e-4
    assert ref.info.spmv_f32 == 0 and ref.info.refine_shift == 0.0


@pytest.mark.parametrize("case", sorted(CASES))
def test_fp32_copy_projections_return_fp64_results(case):
    gen, nev, ncv = CASES[case]
    rp, col, val = gen()
    A = sp.csr_matrix((val, col, rp))
    n = A.shape[0]
    op = op_for(rp, col, val)
    ref = H.ds_solver_gpu(n, nev, ncv, op, tol=TOL, maxiter=3000, basis_prec=H.HL_PREC_F64)
    pj = H.ds_solver_gpu(n, nev, ncv, op, tol=TOL, maxiter=3000, basis_prec=H.HL_PREC_F64_PROJ32)
    assert ref.ret == 0 and pj.ret == 0
    check_fp64_quality(A, pj, nev, ref.values)
    assert pj.info.spmv_f32 == 0
    # the correction is skipped where its exact coefficients are negligible
    # (dgks_skip_exp, 1e-13 by default); with skipping off it runs every step
    assert pj.info.spmv <= pj.info.cgs_passes < 2 * pj.info.spmv
    old = H.tune_get("dgks_skip_exp")
    H.tune_set("dgks_skip_exp", 0)
    try:
        every = H.ds_solver_gpu(n, nev, ncv, op, tol=TOL, maxiter=3000, basis_prec=H.HL_PREC_F64_PROJ32)
    finally:
        H.tune_set("dgks_skip_exp", old)
    assert every.ret == 0 and every.info.cgs_passes >= 2 * every.info.spmv
    check_fp64_quality(A, every, nev, ref.values)


def test_fp32_copy_projections_match_oracle():
    rp, col, val = W.stencil27(14, seed=3)
    A = sp.csr_matrix((val, col, rp))
    n = A.shape[0]
    pj = H.ds_solver_gpu(n, 16, 48, op_for(rp, col, val), tol=TOL, maxiter=2000, basis_prec=H.HL_PREC_F64_PROJ32)
    o = O.ds_solver(O.Csr(rp, col, val), 16, 48, tol=TOL, maxiter=2000)
    assert pj.ret == 0 and o.ret == 0
    check_fp64_quality(A, pj, 16, o.values)


def test_fp32_copy_projections_with_a_host_operator_run_fp64():
    rp, col, val = W.laplacian_2d(16)
    A = sp.csr_matrix((val, col, rp))

    def apply(x, y):
        y[:] = A @ x

    a = H.ds_solver_gpu(256, 4, 12, apply, tol=TOL, maxiter=500, basis_prec=H.HL_PREC_F64_PROJ32)
    b = H.ds_solver_gpu(256, 4, 12, apply, tol=TOL, maxiter=500)
    assert a.ret == 0
    np.testing.assert_array_equal(a.values, b.values)
    np.testing.assert_array_equal(a.vectors, b.vectors)


def test_mixed_solve_matches_oracle():
    rp, col, val = W.stencil27(10, seed=5)
    A = sp.csr_matrix((val, col, rp))
    n = A.shape[0]
    mix = H.ds_solver_gpu(n, 10, 30, op_for(rp, col, val, True), tol=TOL, maxiter=3000, basis_prec=H.HL_PREC_F32)
    o = O.ds_solver(O.Csr(rp, col, val), 10, 30, tol=TOL, maxiter=3000)
    assert mix.ret == 0 and o.ret == 0
    check_fp64_quality(A, mix, 10, o.values)


def test_all_precision_modes_in_one_batch_are_bitwise_their_standalone_solves():
    mats = [W.stencil27(9, seed=1), W.chung_lu_laplacian(1 << 13, max_degree=1500, seed=2),
            W.laplacian_2d(30), W.stencil27(7, seed=4)]
    mats += [W.chung_lu_laplacian(1 << 12, max_degree=800, seed=8), W.stencil27(8, seed=9)]
    ops = [op_for(*m, fp32_values=(i % 2 == 0)) for i, m in enumerate(mats)]
    precs = [H.HL_PREC_F32, H.HL_PREC_F64, H.HL_PREC_F32, H.HL_PREC_F64_PROJ32, H.HL_PREC_F64_PROJ32,
             H.HL_PREC_F64]
    nev, ncv = [6, 6, 5, 4, 5, 6], [18, 18, 15, 12, 16, 18]
    k = len(ops)
    alone = [H.ds_solver_gpu(op.nrows, nev[i], ncv[i], op, tol=TOL, maxiter=3000, seed=50 + i, basis_prec=precs[i])
             for i, op in enumerate(ops)]
    batch = H.ds_solve_batch(ops, nev, ncv, tol=TOL, maxiter=3000, seeds=[50 + i for i in range(k)],
                             basis_prec=precs)
    for b, a in zip(batch, alone):
        assert a.ret == 0 and b.status == 0
        np.testing.assert_array_equal(b.values, a.values)
        np.testing.assert_array_equal(b.vectors, a.vectors)
        assert b.info.spmv == a.info.spmv and b.info.spmv_f32 == a.info.spmv_f32
        assert b.info.restarts == a.info.restarts


def test_mixed_solve_needs_a_device_operator():
    rp, col, val = W.laplacian_2d(8)
    A = sp.csr_matrix((val, col, rp))

    def apply(x, y):
        y[:] = A @ x

    with pytest.raises(H.HLancError):
        H.ds_solver_gpu(64, 4, 12, apply, basis_prec=H.HL_PREC_F32)
