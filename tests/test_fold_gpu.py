"""Folded Lanczos steps (tuning key "fold_partials_max", opt-in): FP64 solves
with few CGS partials close each pass in the next launch instead of a finalize
kernel (5 launches per step instead of 9).  Folding changes the launches, not the
arithmetic: every result must equal, bit for bit, the unfolded solve, alone,
in mixed batches (folded and unfolded solves, FP32-copy solves, a wide basis
beside them), under graph replay, and through the DGKS paths of
HLanc.hpp:1442-1465 (the invariant-subspace case runs the second correction
and its failure branch)."""
import ctypes
from contextlib import contextmanager

import numpy as np
import pytest
import scipy.sparse as sp

import hlanc_amd as H
import oracle as O
from hlanc_amd import workloads as W

pytestmark = pytest.mark.gpu


@contextmanager
def tuned(key, value):
    old = H.tune_get(key)
    H.tune_set(key, value)
    try:
        yield
    finally:
        H.tune_set(key, old)


def op_for(rp, col, val, form=None):
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))
    if form == "sell":
        op.set_sell(True)
    elif form == "slices":
        op.set_slices(1024)
        op.set_items(256, 512)
    return op


def same(a, b, vectors=True):
    assert a.ret == b.ret
    np.testing.assert_array_equal(a.values, b.values)
    if vectors:
        np.testing.assert_array_equal(a.vectors, b.vectors)
    if a.residuals is not None or b.residuals is not None:
        np.testing.assert_array_equal(a.residuals, b.residuals)
    for f in ("restarts", "spmv", "cgs_passes", "nconv", "rnorm", "max_rel_resid"):
        assert getattr(a.info, f) == getattr(b.info, f), f


FOLD = 8192  # folds every solve of these tests (at most 256 CGS blocks)


def solve_both(op, nev, ncv, **kw):
    kw.setdefault("tol", 1e-10)
    kw.setdefault("maxiter", 3000)
    with tuned("fold_partials_max", FOLD):
        folded = H.ds_solver_gpu(op.nrows, nev, ncv, op, basis_prec=H.HL_PREC_F64, **kw)
    with tuned("fold_partials_max", 0):
        plain = H.ds_solver_gpu(op.nrows, nev, ncv, op, basis_prec=H.HL_PREC_F64, **kw)
    return folded, plain


CASES = [
    ("lap2d_64_sell", lambda: W.laplacian_2d(64), "sell", 6, 20),
    ("lap2d_64_items", lambda: W.laplacian_2d(64), None, 6, 20),
    ("st27_16", lambda: W.stencil27(16, seed=2), None, 8, 24),
    ("cl_8k_slices", lambda: W.chung_lu_laplacian(1 << 13, max_degree=2000, seed=3), "slices", 8, 24),
    ("er_3000_ncv128", lambda: W.erdos_renyi(3000, 10, seed=4), None, 40, 128),
    ("er_20000", lambda: W.erdos_renyi(20000, 14, seed=5), None, 10, 30),
]


@pytest.mark.parametrize("name,gen,form,nev,ncv", CASES, ids=[c[0] for c in CASES])
def test_folded_solve_is_bitwise_the_unfolded_one(name, gen, form, nev, ncv):
    rp, col, val = gen()
    op = op_for(rp, col, val, form)
    a, b = solve_both(op, nev, ncv, seed=7, refine=H.HL_REFINE_RQ)
    assert a.ret == 0
    same(a, b)
    assert a.info.max_rel_resid <= 1e-10
    o = O.ds_solver(O.Csr(rp, col, val), nev, ncv, tol=1e-10, maxiter=3000, seed=7)
    assert o.ret == 0
    np.testing.assert_allclose(np.sort(a.values), np.sort(o.values), rtol=1e-9, atol=1e-12)


def test_config0_shape_folds_and_matches():
    """The 2-D Laplacian 256^2, nev 10 / ncv 30 (256 CGS blocks x 31 partials):
    folded, bitwise the unfolded solve, and 4 finalize launches per step fewer
    in the profile."""
    rp, col, val = W.laplacian_2d(256)
    op = op_for(rp, col, val)
    a, b = solve_both(op, 10, 30, seed=0)
    assert a.ret == 0
    same(a, b)
    assert a.info.graph_sweeps > 100 and b.info.graph_sweeps > 100  # replayed sweeps, folded and not
    pa, pb = H.Profile(), H.Profile()
    with tuned("fold_partials_max", FOLD):
        H.ds_solver_gpu(op.nrows, 10, 30, op, tol=1e-10, maxiter=3000, seed=0, basis_prec=H.HL_PREC_F64,
                        flags=H.HL_FLAG_PROFILE)
        H.lib().hl_profile_last(ctypes.byref(pa))
    with tuned("fold_partials_max", 0):
        H.ds_solver_gpu(op.nrows, 10, 30, op, tol=1e-10, maxiter=3000, seed=0, basis_prec=H.HL_PREC_F64,
                        flags=H.HL_FLAG_PROFILE)
        H.lib().hl_profile_last(ctypes.byref(pb))
    steps = a.info.spmv
    assert pa.launches[0] == pb.launches[0] and pa.launches[1] == pb.launches[1]
    assert pb.launches[3] - pa.launches[3] >= 3 * steps


def test_invariant_subspace_runs_the_second_correction_folded():
    # n == ncv: the last steps hit an invariant subspace, where the second DGKS
    # correction runs and fails (the residual is zeroed)
    rng = np.random.default_rng(3)
    M = rng.normal(size=(12, 12))
    A = sp.csr_matrix(M + M.T)
    op = op_for(A.indptr, A.indices, A.data)
    a, b = solve_both(op, 4, 12, tol=0.0, maxiter=200)
    assert a.ret == 0
    same(a, b)
    np.testing.assert_allclose(a.values, np.linalg.eigvalsh(A.toarray())[-4:], rtol=1e-12, atol=1e-12)


def test_folded_and_unfolded_solves_share_a_batch():
    mats = [W.laplacian_2d(40), W.stencil27(12, seed=8), W.erdos_renyi(150000, 8, seed=9),
            W.chung_lu_laplacian(1 << 13, max_degree=2000, seed=10), W.erdos_renyi(5000, 12, seed=11)]
    ops = [op_for(*m) for m in mats]
    nev, ncv = [6, 8, 6, 8, 5], [20, 24, 20, 24, 16]
    prec = [H.HL_PREC_F64, H.HL_PREC_F64, H.HL_PREC_F64, H.HL_PREC_F64_PROJ32, H.HL_PREC_F64]
    kw = dict(tol=1e-10, maxiter=3000, seeds=[20 + i for i in range(5)], basis_prec=prec, refine=H.HL_REFINE_RQ)
    with tuned("fold_partials_max", FOLD):
        a = H.ds_solve_batch(ops, nev, ncv, **kw)
    b = H.ds_solve_batch(ops, nev, ncv, **kw)
    for x, y in zip(a, b):
        assert x.status == y.status == 0
        np.testing.assert_array_equal(x.values, y.values)
        np.testing.assert_array_equal(x.vectors, y.vectors)
        assert x.info.spmv == y.info.spmv and x.info.cgs_passes == y.info.cgs_passes
    for i, op in enumerate(ops):  # and each is its standalone solve
        r = H.ds_solver_gpu(op.nrows, nev[i], ncv[i], op, tol=1e-10, maxiter=3000, seed=20 + i,
                            basis_prec=prec[i], refine=H.HL_REFINE_RQ)
        np.testing.assert_array_equal(a[i].values, r.values)
        np.testing.assert_array_equal(a[i].vectors, r.vectors)


def test_a_wide_basis_in_the_batch_turns_folding_off():
    # ncv 160 sends the FP64 launches to the two-kernel passes, which do not fold:
    # the small solve beside it then runs unfolded, with the same bits
    mats = [W.erdos_renyi(4000, 10, seed=12), W.stencil27(10, seed=13)]
    ops = [op_for(*m) for m in mats]
    kw = dict(tol=1e-10, maxiter=3000, seeds=[1, 2], basis_prec=H.HL_PREC_F64)
    with tuned("fold_partials_max", FOLD):
        a = H.ds_solve_batch(ops, [60, 8], [160, 24], **kw)
    alone = H.ds_solver_gpu(ops[1].nrows, 8, 24, ops[1], tol=1e-10, maxiter=3000, seed=2, basis_prec=H.HL_PREC_F64)
    assert a[0].status == 0 and a[1].status == 0
    np.testing.assert_array_equal(a[1].values, alone.values)
    np.testing.assert_array_equal(a[1].vectors, alone.vectors)
