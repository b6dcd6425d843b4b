"""GPU parity of the staged-slice SpMV form (hl_csr_set_staged) against the
oracle's naive SpMV (HLanc.hpp:403-431): ||y - y_ref|| <= 1e-14 ||y_ref||,
bitwise repeatable, and solves on staged operators against the oracle, alone
and batched."""
import numpy as np
import pytest
import scipy.sparse as sp

import hlanc_amd as H
import oracle as O
import refformats as R
from hlanc_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-14


def pieces_of(rp, col):
    """Pieces of the staged form, counted independently of the library: per row
    and slice of 16384 columns, ceil(entries / 64)."""
    n = len(rp) - 1
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    key = rows * (1 << 20) + (col.astype(np.int64) >> 14)
    _, cnt = np.unique(key, return_counts=True)
    return int(((cnt + 63) // 64).sum())


def ragged_matrix(seed=5, n=777, ncols=50000):
    """Empty rows, rows of 2049, 8192, 8193 and 40000 entries (pieces cut every
    64 entries of a slice), unsorted columns with duplicates, n % 64 != 0."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 40, n)
    lens[::9] = 0
    lens[10], lens[20], lens[30], lens[40] = 2049, 8192, 8193, 40000
    lens[-1] = 0
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = rng.integers(0, ncols, rp[-1]).astype(np.int32)
    val = rng.uniform(-1, 1, rp[-1])
    return rp, col, val, ncols


def case_matrix(case):
    if case == "ragged":
        return ragged_matrix()
    if case == "wide":  # more columns than rows: 31 slices, one row in every slice
        rng = np.random.default_rng(3)
        n, ncols = 300, 500_000
        lens = rng.integers(0, 200, n)
        lens[7] = 0
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        col = rng.integers(0, ncols, rp[-1]).astype(np.int32)
        col[rp[5]:rp[6]] = np.arange(rp[6] - rp[5]) * (ncols // (rp[6] - rp[5]))
        return rp, col, rng.uniform(-1, 1, rp[-1]), ncols
    if case == "tall":  # fewer columns than rows: one slice
        rng = np.random.default_rng(4)
        n, ncols = 20_000, 3_000
        lens = rng.integers(0, 12, n)
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        col = rng.integers(0, ncols, rp[-1]).astype(np.int32)
        return rp, col, rng.uniform(-1, 1, rp[-1]), ncols
    if case == "chunglu":
        rp, col, val = W.chung_lu_laplacian(1 << 16, max_degree=9000, seed=3)
    elif case == "stencil":
        rp, col, val = W.stencil27(24, seed=2)
    elif case == "er":
        rp, col, val = W.erdos_renyi(40_000, 20.0, seed=5)
    else:
        rp, col, val, nc, _ = R.journals()
        return rp, col, val, nc
    return rp, col, val, len(rp) - 1


@pytest.mark.parametrize("case", ["ragged", "wide", "tall", "chunglu", "stencil", "er", "journals"])
def test_staged_spmv_matches_oracle(case):
    rp, col, val, ncols = case_matrix(case)
    n = len(rp) - 1
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val, ncols=ncols))
    op.set_staged(True)
    on, pieces, entries, bins = op.staged()
    assert on and pieces == pieces_of(rp, col)
    # every bin holds at most 3584 pieces, so there are at least pieces / 3584 of them
    assert bins >= max(1, -(-pieces // 3584))
    assert entries >= rp[-1] and entries % 64 == 0
    x = np.random.default_rng(7).uniform(-1, 1, ncols)
    y_ref = O.spmv(O.Csr(rp, col, val, ncols, False), x)
    scale = max(np.linalg.norm(y_ref), 1e-300)
    y = op(x)
    assert np.linalg.norm(y - y_ref) <= TOL * scale, np.linalg.norm(y - y_ref) / scale
    np.testing.assert_array_equal(op(x), y)  # repeatable
    assert (y[np.diff(rp) == 0] == 0.0).all()
    op.set_staged(False)
    assert op.staged() == (False, 0, 0, 0)
    assert np.linalg.norm(op(x) - y_ref) <= TOL * scale


@pytest.mark.parametrize("shape", ["no_entries", "one_row", "five_rows", "one_entry_per_slice"])
def test_staged_spmv_small_and_empty(shape):
    """Edge shapes of the staged form: a matrix with no entries (one bin, no
    chunks), a single row, fewer rows than a chunk's 64 lanes, and a row with
    one entry in each of 200 slices (200 one-entry pieces added in slice
    order)."""
    rng = np.random.default_rng(11)
    if shape == "no_entries":
        n, ncols = 100, 70_000
        rp, col = np.zeros(n + 1, np.int32), np.zeros(0, np.int32)
    elif shape == "one_row":
        n, ncols = 1, 20
        rp, col = np.array([0, 3], np.int32), np.array([19, 0, 7], np.int32)
    elif shape == "five_rows":
        n, ncols = 5, 3
        rp = np.array([0, 2, 2, 5, 6, 9], np.int32)
        col = rng.integers(0, ncols, 9).astype(np.int32)
    else:
        n, ncols = 3, 200 << 14
        rp = np.array([0, 200, 201, 201], np.int32)
        col = np.concatenate([(np.arange(200) << 14) + rng.integers(0, 1 << 14, 200), [5]]).astype(np.int32)
    val = rng.uniform(-1, 1, len(col))
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val, ncols=ncols))
    op.set_staged(True)
    on, pieces, entries, bins = op.staged()
    assert on and pieces == (pieces_of(rp, col) if len(col) else 0) and bins >= 1
    x = rng.uniform(-1, 1, ncols)
    y = op(x)
    if shape == "one_entry_per_slice":  # row 0's pieces: one product each, added in slice order
        prods = val[:200] * x[col[:200]]
        acc = 0.0
        for p_ in prods:
            acc += p_
        assert y[0] == acc and y[1] == val[200] * x[5] and y[2] == 0.0
    y_ref = O.spmv(O.Csr(rp, col, val, ncols, False), x)
    assert np.linalg.norm(y - y_ref) <= TOL * max(np.linalg.norm(y_ref), 1e-300)
    if not len(col):
        assert (y == 0.0).all()


def test_staged_spmv_matches_mkl_golden_vectors(golden):
    """The staged form against the mkl_dcsrmv outputs recorded in
    tests/golden/ (simple.pp, Journals, a random 1000 x 1500 CSR)."""
    cases = []
    for tag, (rp, col, val, nc, sym) in (("simple", R.simple()), ("journals", R.journals())):
        cases.append((rp, col, val, nc, sym, golden[f"spmv_{tag}_x"], golden[f"spmv_{tag}_y"]))
    x = golden["spmv_rand_x"]
    cases.append((golden["spmv_rand_rp"], golden["spmv_rand_col"], golden["spmv_rand_val"], len(x), False, x,
                  golden["spmv_rand_y"]))
    for rp, col, val, nc, sym, x, y_mkl in cases:
        op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val, ncols=nc, is_symmetric=sym))
        op.set_staged(True)
        assert op.staged()[0]
        y = op(x)
        assert np.linalg.norm(y - y_mkl) <= TOL * np.linalg.norm(y_mkl)


def test_staged_rows_sum_in_slice_order():
    """A row's partials are added in slice order, each piece (64 entries of a
    slice, in entry order) summed in entry order: the result is bitwise that
    sequential sum."""
    ncols = 5 << 14
    # one row of 70 entries in each of 5 slices, magnitudes 1 .. 2^-40
    cols, vals = [], []
    for s in range(5):
        for k in range(70):  # two pieces in each slice
            cols.append((s << 14) + 3 * k)
            vals.append(2.0 ** (-10 * s) * (1 + k / 128))
    rp = np.array([0, len(cols)], dtype=np.int32)
    col = np.array(cols, dtype=np.int32)
    val = np.array(vals)
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val, ncols=ncols))
    op.set_staged(True)
    assert op.staged()[1] == 10
    x = np.ones(ncols)
    want = 0.0
    for s in range(5):
        for p0 in (0, 64):
            part = 0.0
            for k in range(p0, min(p0 + 64, 70)):
                part += vals[s * 70 + k]
            want += part
    assert op(x)[0] == want


def test_staged_forms_exclude_each_other_and_limits():
    rp, col, val = W.stencil27(12, seed=1)
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))
    op.set_slices(256)
    op.set_staged(True)  # drops the slices
    assert op.staged()[0] and op.slices() == (0, 0)
    op.set_sell(True)  # drops the staged form
    assert op.sell()[0] and not op.staged()[0]
    op.set_staged(True)
    assert not op.sell()[0] and op.staged()[0]
    op.set_slices(512)
    assert not op.staged()[0] and op.slices()[0] == 512
    # more than 4000 slices of 16384 columns
    wide = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=np.array([0, 2], dtype=np.int32),
                                                 col=np.array([0, 70_000_000], dtype=np.int32),
                                                 val=np.array([1.0, 2.0]), ncols=70_000_001))
    with pytest.raises(H.HLancError):
        wide.set_staged(True)
    assert not wide.staged()[0]
    # a row of more pieces than a bin holds (3584): 230,000 entries in one slice
    # make 3594 pieces; 229,000 make 3579
    for m, ok in ((230_000, False), (229_000, True)):
        long_row = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=np.array([0, m], dtype=np.int32),
                                                         col=(np.arange(m) % 16384).astype(np.int32),
                                                         val=np.ones(m), ncols=16384))
        if not ok:
            with pytest.raises(H.HLancError):
                long_row.set_staged(True)
            assert not long_row.staged()[0]
        else:
            long_row.set_staged(True)
            assert long_row.staged()[1] == 3579
            assert long_row(np.ones(16384))[0] == float(m)


def test_staged_operator_solves_match_oracle_alone_and_batched():
    """Solves on a staged operator (its bins kernel writes w and the basis
    column, and the FP32 copy in the FP32-copy mode) against the oracle; each
    bitwise the same in a batch beside a solve on an unstaged operator and an
    FP32-basis solve on the same operator (which takes the work items)."""
    rp, col, val = W.chung_lu_laplacian(1 << 16, max_degree=4000, seed=11)
    n = len(rp) - 1
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))
    op.set_staged(True)
    o = O.ds_solver(O.Csr(rp, col, val), 8, 24, tol=1e-10, maxiter=1000)
    assert o.ret == 0
    A = sp.csr_matrix((val, col, rp))
    alone = []
    for prec in (H.HL_PREC_F64, H.HL_PREC_F64_PROJ32, H.HL_PREC_F32):
        r = H.ds_solver_gpu(n, 8, 24, op, tol=1e-10, maxiter=1000, basis_prec=prec)
        assert r.ret == 0
        np.testing.assert_allclose(r.values, o.values, rtol=1e-10)
        assert r.info.max_rel_resid <= 1e-10
        for i in range(8):
            v = r.vectors[:, i]
            assert np.linalg.norm(A @ v - r.values[i] * v) <= 1e-10 * np.abs(r.values).max()
        alone.append(r)
    plain = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))
    precs = [H.HL_PREC_F64, H.HL_PREC_F64_PROJ32, H.HL_PREC_F32, H.HL_PREC_F64]
    b = H.ds_solve_batch([op, op, op, plain], 8, 24, tol=1e-10, maxiter=1000, basis_prec=precs)
    for k in range(3):
        np.testing.assert_array_equal(b[k].values, alone[k].values)
        np.testing.assert_array_equal(b[k].vectors, alone[k].vectors)
    r_plain = H.ds_solver_gpu(n, 8, 24, plain, tol=1e-10, maxiter=1000, basis_prec=H.HL_PREC_F64)
    np.testing.assert_array_equal(b[3].values, r_plain.values)


def test_staged_solve_with_an_empty_slice():
    """A square operator whose middle slice of 16384 columns (and so, being
    symmetric, those rows) holds no entry: no piece and no chunk there, while
    the basis column is still written over those rows.  Values and vectors
    against the oracle and dense residuals, alone and batched bitwise."""
    rng = np.random.default_rng(12)
    n, blk = 3 << 14, 1 << 14
    keep = np.r_[0:blk, 2 * blk:3 * blk]  # rows/columns of slices 0 and 2
    m = len(keep)
    a = rng.integers(0, m, 8 * m)
    b = rng.integers(0, m, 8 * m)
    v = rng.uniform(-1, 1, 8 * m)
    B = sp.coo_matrix((np.r_[v, v], (np.r_[a, b], np.r_[b, a])), shape=(m, m)).tocsr()
    d = np.linspace(0.0, 10.0, m)
    d[-8:] = 30.0 + 2.0 * np.arange(8)  # separated top of the spectrum: a quick solve
    B = B + sp.diags(d)
    P = sp.csr_matrix((np.ones(m), (keep, np.arange(m))), shape=(n, m))
    A = (P @ B @ P.T).tocsr()
    A.sort_indices()
    rp, col, val = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))
    op.set_staged(True)
    assert op.staged()[0]
    o = O.ds_solver(O.Csr(rp, col, val), 4, 16, tol=1e-10, maxiter=2000)
    assert o.ret == 0
    alone = []
    for prec in (H.HL_PREC_F64, H.HL_PREC_F64_PROJ32):
        r = H.ds_solver_gpu(n, 4, 16, op, tol=1e-10, maxiter=2000, basis_prec=prec)
        assert r.ret == 0
        np.testing.assert_allclose(r.values, o.values, rtol=1e-10)
        for i in range(4):
            x = r.vectors[:, i]
            assert np.linalg.norm(A @ x - r.values[i] * x) <= 1e-9 * np.abs(r.values).max()
        alone.append(r)
    bt = H.ds_solve_batch([op, op], 4, 16, tol=1e-10, maxiter=2000,
                          basis_prec=[H.HL_PREC_F64, H.HL_PREC_F64_PROJ32])
    for k in range(2):
        np.testing.assert_array_equal(bt[k].values, alone[k].values)
        np.testing.assert_array_equal(bt[k].vectors, alone[k].vectors)


def test_staged_ncv96_solve_matches_oracle():
    """A wider basis (ncv 96) over a staged operator, FP32-copy mode, against
    the oracle and bitwise the same batched with a second staged operator."""
    rp, col, val = W.chung_lu_laplacian(1 << 15, max_degree=3000, seed=4)
    rp2, col2, val2 = W.erdos_renyi(30_000, 16.0, seed=9)
    n, n2 = len(rp) - 1, len(rp2) - 1
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))
    op2 = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp2, col=col2, val=val2))
    op.set_staged(True)
    op2.set_staged(True)
    o = O.ds_solver(O.Csr(rp, col, val), 32, 96, tol=1e-10, maxiter=1000)
    assert o.ret == 0
    r = H.ds_solver_gpu(n, 32, 96, op, tol=1e-10, maxiter=1000, basis_prec=H.HL_PREC_F64_PROJ32)
    assert r.ret == 0 and r.info.max_rel_resid <= 1e-10
    np.testing.assert_allclose(r.values, o.values, rtol=1e-10)
    r2 = H.ds_solver_gpu(n2, 8, 24, op2, tol=1e-10, maxiter=1000, basis_prec=H.HL_PREC_F64)
    b = H.ds_solve_batch([op, op2], [32, 8], [96, 24], tol=1e-10, maxiter=1000,
                         basis_prec=[H.HL_PREC_F64_PROJ32, H.HL_PREC_F64])
    np.testing.assert_array_equal(b[0].values, r.values)
    np.testing.assert_array_equal(b[1].values, r2.values)


def test_autotune_can_pick_the_staged_form():
    """The autotuner times the staged form on operators of >= 65536 nonzeros
    and leaves exactly one form in place, whichever wins; the SpMV matches the
    oracle afterwards."""
    rp, col, val = W.chung_lu_laplacian(1 << 18, seed=1)
    n = len(rp) - 1
    op = H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))
    nc, rc, t = op.autotune(reps=3)
    forms = [op.slices()[0] > 0, op.sell()[0], op.staged()[0]]
    assert sum(forms) <= 1 and t > 0
    x = np.random.default_rng(8).uniform(-1, 1, n)
    y_ref = O.spmv(O.Csr(rp, col, val, n, False), x)
    assert np.linalg.norm(op(x) - y_ref) <= TOL * np.linalg.norm(y_ref)
