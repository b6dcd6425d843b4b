"""Wide Krylov bases on the GPU: ncv = 96 (the 27-point workload's nev 32 /
ncv 96, and the nev-32 requests of the serving stream), ncv = 128, and bases
wider than the fused kernels hold (ncv = 160, 256, 400), in the FP64 basis mode
and with the FP32 copy for the projection passes.

Up to 128 columns these solves run the widest fused instances: the CGS passes
over 65..128 columns (24 and 32 columns per lane group; FP32 passes of 12 and 16
column pairs), and restarts whose kept block spans 3..5 output tiles of 16
columns on the MFMA V.Q kernel next to restarts whose block of Q no longer fits
its LDS (the FMA kernel).  Beyond 128 columns a CGS pass runs as a subtracting
and a projecting kernel over 128-column chunks, restarts run a K-chunked V.Q
with the residual update on its own, and the Ritz-vector step of ncv = 400 runs
in two K-chunks.  Each is checked against the CPU oracle (values to 1e-10,
residuals, orthogonality), alone and bitwise inside a batch; narrow solves keep
their bits when a wide one shares their launches.

The full-size workloads (Chung-Lu n = 4.1 M, nev 20 / ncv 60, and the 27-point
stencil 128^3, nev 32 / ncv 96) are checked by size-independent properties:
every returned pair at ||A v - l v|| <= 1e-10 ||A||, orthonormal vectors, and
the FP32-copy mode's values within 1e-10 of the FP64 mode's."""
import numpy as np
import pytest
import scipy.sparse as sp

import hlanc_amd as H
import oracle as O
from hlanc_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10
PRECS = {"f64": H.HL_PREC_F64, "proj32": H.HL_PREC_F64_PROJ32}

CASES = {
    # name: (generator, nev, ncv)
    "st27_24_ncv96": (lambda: W.stencil27(24, seed=3), 32, 96),
    "st27_16_ncv128": (lambda: W.stencil27(16, seed=4), 48, 128),
    "er6000_ncv128": (lambda: W.erdos_renyi(6000, 12, seed=5), 40, 128),
    # wider than the fused CGS pass (two kernels per pass, 128-column chunks) and
    # than the segmented restart kernels (K-chunked V.Q; K > 384 on the Ritz step)
    "st27_16_ncv160": (lambda: W.stencil27(16, seed=4), 48, 160),
    "er4000_ncv256": (lambda: W.erdos_renyi(4000, 12, seed=6), 80, 256),
    "st27_14_ncv400": (lambda: W.stencil27(14, seed=2), 100, 400),
}
WIDE = ["st27_16_ncv160", "er4000_ncv256", "st27_14_ncv400"]

_mats = {}


def case(name):
    if name not in _mats:
        gen, nev, ncv = CASES[name]
        rp, col, val = gen()
        o = O.ds_solver(O.Csr(rp, col, val), nev, ncv, tol=TOL, maxiter=3000, nthreads=8)
        assert o.ret == 0
        _mats[name] = (rp, col, val, nev, ncv, o)
    return _mats[name]


def op_for(rp, col, val):
    return H.dcsrmv_operator_gpu(H.dcsr_matrix_t(row_ptr=rp, col=col, val=val))


def check_pairs(A, values, vectors, nev, ref):
    np.testing.assert_allclose(values, ref, rtol=TOL)
    nrm = np.abs(ref).max()
    V = np.asarray(vectors[:, :nev])
    R = A @ V - V * values[None, :]
    assert np.linalg.norm(R, axis=0).max() <= TOL * nrm
    assert np.abs(V.T @ V - np.eye(nev)).max() < 1e-10


@pytest.mark.parametrize("prec", sorted(PRECS))
@pytest.mark.parametrize("name", sorted(CASES))
def test_wide_basis_matches_oracle(name, prec):
    rp, col, val, nev, ncv, o = case(name)
    A = sp.csr_matrix((val, col, rp))
    r = H.ds_solver_gpu(A.shape[0], nev, ncv, op_for(rp, col, val), tol=TOL, maxiter=3000,
                        basis_prec=PRECS[prec])
    assert r.ret == 0
    check_pairs(A, r.values, r.vectors, nev, o.values)
    assert r.info.restarts >= 2  # restarts ran (the V.Q kernels under test)


@pytest.mark.parametrize("name", ["st27_24_ncv96", "er6000_ncv128", "st27_16_ncv160", "st27_14_ncv400"])
def test_wide_basis_fma_restart_agrees(name):
    rp, col, val, nev, ncv, o = case(name)
    A = sp.csr_matrix((val, col, rp))
    r = H.ds_solver_gpu(A.shape[0], nev, ncv, op_for(rp, col, val), tol=TOL, maxiter=3000, use_mfma=False,
                        basis_prec=H.HL_PREC_F64_PROJ32)
    assert r.ret == 0
    check_pairs(A, r.values, r.vectors, nev, o.values)


def test_wide_bases_in_one_batch_are_bitwise_their_standalone_solves():
    names = ["st27_24_ncv96", "st27_24_ncv96", "er6000_ncv128", "st27_16_ncv128"]
    precs = [H.HL_PREC_F64, H.HL_PREC_F64_PROJ32, H.HL_PREC_F64_PROJ32, H.HL_PREC_F64]
    mats = [case(nm) for nm in names]
    ops = [op_for(m[0], m[1], m[2]) for m in mats]
    # a narrow solve in the same launches
    rp, col, val = W.stencil27(10, seed=9)
    ops.append(op_for(rp, col, val))
    precs.append(H.HL_PREC_F64_PROJ32)
    nev = [m[3] for m in mats] + [6]
    ncv = [m[4] for m in mats] + [18]
    alone = [H.ds_solver_gpu(op.nrows, nev[i], ncv[i], op, tol=TOL, maxiter=3000, seed=7 + i, basis_prec=precs[i])
             for i, op in enumerate(ops)]
    batch = H.ds_solve_batch(ops, nev, ncv, tol=TOL, maxiter=3000, seeds=[7 + i for i in range(len(ops))],
                             basis_prec=precs)
    for i, (b, a) in enumerate(zip(batch, alone)):
        assert a.ret == 0 and b.status == 0, i
        np.testing.assert_array_equal(b.values, a.values)
        np.testing.assert_array_equal(b.vectors, a.vectors)
        assert b.info.restarts == a.info.restarts and b.info.spmv == a.info.spmv
    for i, m in enumerate(mats):
        np.testing.assert_allclose(alone[i].values, m[5].values, rtol=TOL)


def test_narrow_solves_beside_a_wide_one_keep_their_bits():
    """A batch step of more than 128 columns runs every solve of its launch
    through the two-kernel pass; a narrow solve's sums are the fused pass's, so
    it stays bitwise its standalone solve."""
    rp, col, val, nev, ncv, o = case("st27_16_ncv160")
    wide = op_for(rp, col, val)
    narrow = [op_for(*W.stencil27(10, seed=9)), op_for(*W.erdos_renyi(3000, 10, seed=4))]
    for prec in (H.HL_PREC_F64, H.HL_PREC_F64_PROJ32):
        alone = [H.ds_solver_gpu(op.nrows, 8, 24, op, tol=TOL, maxiter=3000, seed=11 + i, basis_prec=prec)
                 for i, op in enumerate(narrow)]
        batch = H.ds_solve_batch(narrow + [wide], [8, 8, nev], [24, 24, ncv], tol=TOL, maxiter=3000,
                                 seeds=[11, 12, 0], basis_prec=prec)
        for a, b in zip(alone, batch[:2]):
            assert a.ret == 0 and b.status == 0
            np.testing.assert_array_equal(b.values, a.values)
            np.testing.assert_array_equal(b.vectors, a.vectors)
        assert batch[2].status == 0
        np.testing.assert_allclose(batch[2].values, o.values, rtol=TOL)


def test_basis_beyond_the_build_width_is_an_argument_error():
    rp, col, val = W.stencil27(12, seed=1)
    op = op_for(rp, col, val)
    with pytest.raises(H.HLancError):
        H.ds_solver_gpu(op.nrows, 100, 600, op)
    with pytest.raises(H.HLancError):  # the two-phase FP32 mode stops at 128
        H.ds_solver_gpu(op.nrows, 50, 160, op, basis_prec=H.HL_PREC_F32)


# ------------------------------------------------------------ full size

def device_check(op, values, vecs_dev, nev):
    """max_i ||A v_i - l_i v_i|| / max|l| and max |V^T V - I| with the GPU SpMV."""
    import torch

    V = vecs_dev[:, :nev].T.contiguous()  # nev x n, rows contiguous
    AV = torch.empty_like(V)
    for i in range(nev):
        op.spmv(V[i].data_ptr(), AV[i].data_ptr())
    torch.cuda.synchronize()
    lam = torch.tensor(values, device=V.device, dtype=torch.float64)[:, None]
    res = torch.linalg.norm(AV - lam * V, dim=1).max().item() / float(np.abs(values).max())
    G = V @ V.T
    ortho = (G - torch.eye(nev, device=V.device, dtype=torch.float64)).abs().max().item()
    return res, ortho


@pytest.mark.parametrize("workload", ["chunglu_cfg1", "stencil27_cfg2"])
def test_full_size_workload_properties(workload):
    import torch

    if workload == "chunglu_cfg1":
        rp, col, val = W.chung_lu_laplacian(1 << 22, seed=0)
        nev, ncv = 20, 60
    else:
        rp, col, val = W.stencil27(128, seed=0)
        nev, ncv = 32, 96
    n = len(rp) - 1
    op = op_for(rp, col, val)
    del rp, col, val
    op.autotune(reps=5)  # the form the bench runs
    out = {}
    for prec in ("proj32", "f64"):
        vec = torch.empty((ncv, n), dtype=torch.float64, device="cuda").T
        r = H.ds_solver_gpu(n, nev, ncv, op, tol=TOL, maxiter=1000, vectors_out=vec, basis_prec=PRECS[prec])
        assert r.ret == 0, prec
        res, ortho = device_check(op, r.values, vec, nev)
        assert res <= TOL, (prec, res)
        assert ortho < 1e-10, (prec, ortho)
        out[prec] = r.values
        del vec
    np.testing.assert_allclose(out["proj32"], out["f64"], rtol=TOL)
