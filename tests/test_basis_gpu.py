"""GPU parity of the basis kernels (mat_v_gpu_t's dense work, HLanc.hpp:1000-1107)
against float64 numpy: the fused CGS pass and the V.Q product (MFMA and FMA)."""
import numpy as np
import pytest

import hlanc_amd as H

pytestmark = pytest.mark.gpu


def dev(a):
    import torch

    return torch.tensor(np.asfortranarray(a).ravel(order="F"), device="cuda")


@pytest.mark.parametrize("n", [1000, 65536 + 13])
@pytest.mark.parametrize("m", [0, 1, 5, 30, 64, 96, 128])
@pytest.mark.parametrize("with_h", [False, True])
def test_cgs_pass_matches_numpy(n, m, with_h):
    import torch

    rng = np.random.default_rng(n + m)
    ldv = n + 7
    V = np.zeros((ldv, max(m, 1)))
    V[:n, :m] = rng.uniform(-1, 1, (n, m))
    w = rng.uniform(-1, 1, n)
    h = rng.uniform(-1, 1, max(m, 1))
    Vd, wd, hd = dev(V), torch.tensor(w, device="cuda"), torch.tensor(h, device="cuda")
    wout = torch.empty_like(wd)
    hout = torch.zeros(max(m, 1), dtype=torch.float64, device="cuda")
    nrm = torch.zeros(1, dtype=torch.float64, device="cuda")
    H.cgs_pass(0, n, m, Vd.data_ptr(), ldv, wd.data_ptr(), wout.data_ptr(), hd.data_ptr() if with_h else None,
               hout.data_ptr(), nrm.data_ptr())
    torch.cuda.synchronize()
    Vn = V[:n, :m]
    w_ref = w - Vn @ h[:m] if with_h else w
    h_ref = Vn.T @ w_ref
    scale = np.abs(Vn).sum(axis=1).max() * np.abs(h[:m]).max() if (with_h and m) else 1.0
    np.testing.assert_allclose(wout.cpu().numpy(), w_ref, rtol=0, atol=1e-14 * max(scale, 1.0) * 4)
    if m:
        err = np.abs(hout.cpu().numpy()[:m] - h_ref)
        assert np.all(err <= 1e-13 * (np.abs(Vn).T @ np.abs(w_ref)))
    assert abs(nrm.item() - w_ref @ w_ref) <= 1e-13 * (w_ref @ w_ref)


@pytest.mark.parametrize("flags", [0, 1, 2, 3, 8, 9, 11])
def test_cgs_pass_variants_match_numpy(flags):
    # every tunable variant of the standalone pass (non-temporal loads,
    # prefetch, split finalize) computes the same pass
    import torch

    n, m = 70001, 45
    rng = np.random.default_rng(flags)
    V = rng.uniform(-1, 1, (n, m))
    w = rng.uniform(-1, 1, n)
    h = rng.uniform(-1, 1, m)
    Vd, wd, hd = dev(V), torch.tensor(w, device="cuda"), torch.tensor(h, device="cuda")
    wout = torch.empty_like(wd)
    hout = torch.zeros(m, dtype=torch.float64, device="cuda")
    nrm = torch.zeros(1, dtype=torch.float64, device="cuda")
    old = H.tune_get("cgs_flags")
    try:
        H.tune_set("cgs_flags", flags)
        H.cgs_pass(0, n, m, Vd.data_ptr(), n, wd.data_ptr(), wout.data_ptr(), hd.data_ptr(), hout.data_ptr(),
                   nrm.data_ptr())
        torch.cuda.synchronize()
    finally:
        H.tune_set("cgs_flags", old)
    w_ref = w - V @ h
    h_ref = V.T @ w_ref
    np.testing.assert_allclose(wout.cpu().numpy(), w_ref, rtol=0, atol=1e-13 * np.abs(V).sum(axis=1).max())
    assert np.all(np.abs(hout.cpu().numpy() - h_ref) <= 1e-13 * (np.abs(V).T @ np.abs(w_ref)))
    assert abs(nrm.item() - w_ref @ w_ref) <= 1e-13 * (w_ref @ w_ref)


def test_cgs_flag_bit2_is_refused():
    with pytest.raises(H.HLancError):
        H.tune_set("cgs_flags", 4)


@pytest.mark.parametrize("mfma", [True, False])
@pytest.mark.parametrize("n,K,N", [(1000, 12, 1), (1000, 12, 11), (4099, 30, 11), (4099, 30, 30),
                                   (20000, 96, 33), (20000, 96, 96), (3001, 60, 64), (3001, 60, 65),
                                   (513, 128, 128), (17, 7, 3)])
def test_vq_matches_numpy(mfma, n, K, N):
    import torch

    rng = np.random.default_rng(K * 1000 + N)
    ldv, ldo, ldq = n + 5, n + 3, K + 2
    V = np.zeros((ldv, K))
    V[:n] = rng.uniform(-1, 1, (n, K))
    Q = np.zeros((ldq, N))
    Q[:K] = rng.uniform(-1, 1, (K, N))
    Vd, Qd = dev(V), dev(Q)
    Od = torch.full((ldo * N,), np.nan, dtype=torch.float64, device="cuda")
    H.vq(0, n, K, N, Vd.data_ptr(), ldv, Qd.data_ptr(), ldq, Od.data_ptr(), ldo, use_mfma=mfma)
    torch.cuda.synchronize()
    out = Od.cpu().numpy().reshape(N, ldo).T
    ref = V[:n] @ Q[:K]
    bound = 4 * K * 1.2e-16 * (np.abs(V[:n]) @ np.abs(Q[:K]))
    assert np.all(np.abs(out[:n] - ref) <= bound + 1e-300)
    assert np.all(np.isnan(out[n:]))  # padding rows untouched


@pytest.mark.parametrize("mfma", [True, False])
def test_vq_identity_places_every_column(mfma):
    # Q = I (K = N = 48): Out must reproduce V column by column; a wrong MFMA
    # output layout shows up as permuted or missing columns.
    import torch

    n, K = 100, 48
    V = np.arange(n * K, dtype=np.float64).reshape(K, n).T.copy() + 1.0
    Q = np.eye(K)
    Vd, Qd = dev(V), dev(Q)
    Od = torch.zeros(n * K, dtype=torch.float64, device="cuda")
    H.vq(0, n, K, K, Vd.data_ptr(), n, Qd.data_ptr(), K, Od.data_ptr(), n, use_mfma=mfma)
    torch.cuda.synchronize()
    out = Od.cpu().numpy().reshape(K, n).T
    got = [int(round(out[0, c] - 1.0)) // n if out[0, c] else -1 for c in range(K)]
    assert got == list(range(K)), got
    np.testing.assert_array_equal(out, V)
