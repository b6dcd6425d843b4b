"""hl_cgs_pass_noalloc, the allocation-free variant of hl_cgs_pass, called
through the C ABI: on two calls in a row (the second with a smaller basis) it
must give bitwise the outputs of hl_cgs_pass on the same operands (same
kernels, same grid), and both must match the pass computed in numpy
(w_out = w_in - V h_in, h_out = V^T w_out, nrm2 = ||w_out||^2) to FP64
rounding.  Its limit (m <= 128) is an argument error with a message."""
import numpy as np
import pytest

import hlanc_amd as H

pytestmark = pytest.mark.gpu


def operands(n, m, ldv, seed):
    import torch

    rng = np.random.default_rng(seed)
    Vh = rng.standard_normal((m, ldv))  # row c = column c of the column-major n x m basis (ld ldv)
    wh = rng.standard_normal(n)
    hh = rng.standard_normal(m)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    return (Vh, wh, hh), (dev(Vh), dev(wh), dev(hh))


def call(name, n, m, ldv, dv, with_outputs=True):
    import torch

    V, w, h = dv
    w_out = torch.full((max(n, 1),), float("nan"), dtype=torch.float64, device="cuda")
    h_out = torch.full((max(m, 1),), float("nan"), dtype=torch.float64, device="cuda")
    nrm = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    fn = getattr(H.lib(), name)
    rc = fn(0, n, m, V.data_ptr(), ldv, w.data_ptr(), w_out.data_ptr(), h.data_ptr(),
            h_out.data_ptr() if with_outputs else None, nrm.data_ptr() if with_outputs else None, s)
    torch.cuda.synchronize()
    return rc, w_out.cpu().numpy()[:n], h_out.cpu().numpy()[:m], nrm.cpu().numpy()[0]


def numpy_pass(n, m, hv):
    Vh, wh, hh = hv
    V = Vh[:, :n].T
    w_out = wh - V @ hh
    return w_out, V.T @ w_out, float(w_out @ w_out)


@pytest.mark.parametrize("shapes", [[(300_000, 100, 300_064), (5_000, 7, 5_000)],
                                    [(70_000, 128, 70_001), (64, 1, 64)]])
def test_no_alloc_pass_equals_the_allocating_pass(shapes):
    for k, (n, m, ldv) in enumerate(shapes):
        hv, dv = operands(n, m, ldv, seed=10 * k + m)
        rc0, w0, h0, s0 = call("hl_cgs_pass", n, m, ldv, dv)
        rc1, w1, h1, s1 = call("hl_cgs_pass_noalloc", n, m, ldv, dv)
        assert rc0 == 0 and rc1 == 0, H.lib().hl_last_error()
        np.testing.assert_array_equal(w1, w0)
        np.testing.assert_array_equal(h1, h0)
        assert s1 == s0
        wr, hr, sr = numpy_pass(n, m, hv)
        scale = np.abs(hv[0][:, :n]).max() * np.abs(wr).max() * n
        np.testing.assert_allclose(w1, wr, rtol=0, atol=1e-12 * np.abs(wr).max() * m)
        np.testing.assert_allclose(h1, hr, rtol=0, atol=1e-13 * scale)
        assert abs(s1 - sr) <= 1e-12 * sr


def test_no_alloc_pass_optional_outputs_and_limit():
    n, m, ldv = 20_000, 40, 20_000
    hv, dv = operands(n, m, ldv, seed=3)
    rc0, w0, _, _ = call("hl_cgs_pass", n, m, ldv, dv)
    rc1, w1, h1, s1 = call("hl_cgs_pass_noalloc", n, m, ldv, dv, with_outputs=False)
    assert rc0 == 0 and rc1 == 0
    np.testing.assert_array_equal(w1, w0)  # w_out written, h_out / nrm2 left alone
    assert np.isnan(h1).all() and np.isnan(s1)
    hv, dv = operands(1000, 129, 1000, seed=4)
    rc, _, _, _ = call("hl_cgs_pass_noalloc", 1000, 129, 1000, dv)
    assert rc == -1  # HL_ERR_ARG, before any launch
    assert b"128" in H.lib().hl_last_error()
