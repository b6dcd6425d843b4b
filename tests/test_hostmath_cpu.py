"""The product's own host FP64 kernels of the restart, checked on the CPU through
their C-ABI exports (hl_host_*), apart from any solve:

* tridiag_eig (dsteqr("I"), HLanc.hpp:1691,1805) against MKL dsteqr outputs in
  tests/golden/golden.npz, scipy's eigh_tridiagonal and the invariants
  T Z = Z Lambda, Z^T Z = I; its last-row form bitwise against the full form;
* dlartg (HLanc.hpp:1514,1536) against MKL dlartg outputs, sign convention
  included;
* householder_q (dgeqr2 + dorm2r, HLanc.hpp:1837-1840) against MKL outputs and
  numpy.linalg.qr;
* restart_chase (the shift sweeps of dsapps, HLanc.hpp:1492-1589), for the
  product and the oracle alike, by the implicit-restart theorem on a Lanczos
  factorization A V = V T + r e^T of a small dense A: Q is orthogonal and banded,
  the updated T is Q^T T Q (still tridiagonal), V Q e_1 is parallel to
  prod_i (A - mu_i I) v_1, and the truncated factorization holds again with
  the residual the restart forms (HLanc.hpp:1605-1614).
A dropped term, a flipped sign or a transposed rotation fails one of these."""
import numpy as np
import pytest
import scipy.linalg as sla

import hlanc_amd as H
import oracle as O

TAGS = ["t5", "t30", "t60", "t96", "deg"]


@pytest.mark.parametrize("tag", TAGS)
def test_tridiag_eig_matches_mkl_dsteqr(golden, tag):
    d, e = golden[f"steqr_{tag}_d"], golden[f"steqr_{tag}_e"]
    w_ref, z_ref = golden[f"steqr_{tag}_w"], golden[f"steqr_{tag}_z"]
    info, w, z = H.host_tridiag_eig(d, e)
    assert info == 0
    n = len(d)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    nrm = np.abs(w_ref).max()
    assert np.all(np.diff(w) >= 0)
    np.testing.assert_allclose(w, w_ref, rtol=0, atol=1e-13 * nrm)
    np.testing.assert_allclose(w, sla.eigh_tridiagonal(d, e, eigvals_only=True), rtol=0, atol=1e-13 * nrm)
    assert np.abs(T @ z - z * w).max() <= 1e-13 * nrm
    assert np.abs(z.T @ z - np.eye(n)).max() <= 1e-13
    if tag != "deg":  # simple spectrum: the vectors equal MKL's up to sign
        np.testing.assert_allclose(np.abs(np.sum(z * z_ref, axis=0)), 1.0, atol=1e-10)
    # the last-row form: the same eigenvalues and Z's last row, bit for bit
    info2, w2, zl = H.host_tridiag_eig(d, e, last_row_only=True)
    assert info2 == 0
    np.testing.assert_array_equal(w2, w)
    np.testing.assert_array_equal(zl, z[-1, :])


def hard_tridiagonals():
    rng = np.random.default_rng(5)
    out = {
        "wilkinson21": (np.abs(np.arange(-10, 11)).astype(float), np.ones(20)),  # close pairs
        "diagonal": (rng.uniform(-1, 1, 9), np.zeros(8)),  # already reduced
        "split": (rng.uniform(-1, 1, 12), np.where(np.arange(11) == 5, 0.0, rng.uniform(0.1, 1, 11))),
        "one": (np.array([3.0]), np.zeros(0)),
        "two": (np.array([1.0, -2.0]), np.array([0.5])),
        "huge": (rng.uniform(-1, 1, 20) * 1e150, rng.uniform(-1, 1, 19) * 1e150),
        "tiny": (rng.uniform(-1, 1, 20) * 1e-150, rng.uniform(-1, 1, 19) * 1e-150),
        "graded": (10.0 ** -np.arange(16, dtype=float), 10.0 ** -np.arange(0.5, 15.5, dtype=float)),
        "lanczos_like": (rng.uniform(0, 2, 60), rng.uniform(1e-9, 1, 59)),
    }
    return out


@pytest.mark.parametrize("name", sorted(hard_tridiagonals()))
def test_tridiag_eig_hard_cases(name):
    d, e = hard_tridiagonals()[name]
    info, w, z = H.host_tridiag_eig(d, e)
    assert info == 0
    n = len(d)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    nrm = max(np.abs(np.linalg.eigvalsh(T)).max(), np.finfo(float).tiny)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(T), rtol=0, atol=4e-15 * n * nrm)
    assert np.abs(T @ z - z * w).max() <= 4e-15 * n * nrm
    assert np.abs(z.T @ z - np.eye(n)).max() <= 4e-15 * n


def test_tridiag_eig_flags_nonfinite():
    assert H.host_tridiag_eig([1.0, np.nan, 2.0], [0.5, 0.5])[0] == 1
    assert H.host_tridiag_eig([1.0, 2.0, 3.0], [np.inf, 0.5])[0] == 1


def test_dlartg_matches_mkl(golden):
    fg, ref = golden["dlartg_fg"], golden["dlartg_csr"]
    for (f, g), (c0, s0, r0) in zip(fg, ref):
        c, s, r = H.host_dlartg(f, g)
        assert np.isclose(c, c0, rtol=4e-16, atol=0) and np.isclose(s, s0, rtol=4e-16, atol=0)
        assert np.isclose(r, r0, rtol=4e-16, atol=0)
        assert abs(-s * f + c * g) <= 4e-16 * max(abs(f), abs(g))


def test_householder_q_matches_mkl_and_numpy(golden):
    a, af_ref, tau_ref = golden["qr_a"], golden["qr_af"], golden["qr_tau"]
    af, tau, Q = H.host_householder_q(a)
    np.testing.assert_allclose(tau, tau_ref, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(af, af_ref, rtol=1e-12, atol=1e-13)
    m, k = a.shape
    Qn, Rn = np.linalg.qr(a, mode="complete")  # the same dgeqrf reflector convention
    np.testing.assert_allclose(Q, Qn, atol=1e-13)
    np.testing.assert_allclose(np.triu(af)[:k], Rn[:k], atol=1e-12)
    np.testing.assert_allclose(golden["qr_c"] @ Q, golden["qr_cq"], atol=1e-12)  # dorm2r("R","N")


# ------------------------------------------------------- the restart chase

def lanczos(A, m, seed):
    """A V = V T + r e_m^T with full reorthogonalization (numpy, FP64)."""
    n = A.shape[0]
    rng = np.random.default_rng(seed)
    V = np.zeros((n, m))
    T = np.zeros((m, m))
    v = rng.standard_normal(n)
    V[:, 0] = v / np.linalg.norm(v)
    beta = 0.0
    for j in range(m):
        w = A @ V[:, j]
        h = V[:, : j + 1].T @ w
        w -= V[:, : j + 1] @ h
        h2 = V[:, : j + 1].T @ w
        w -= V[:, : j + 1] @ h2
        T[j, j] = h[j] + h2[j]
        if j > 0:
            T[j, j - 1] = T[j - 1, j] = beta
        beta = np.linalg.norm(w)
        if j + 1 < m:
            V[:, j + 1] = w / beta
    return V, T, w


CHASES = {"product": H.host_restart_chase, "oracle": O.dsapps_chase}


def check_chase(chase, A, V, T, r, nev, shifts, tol_dir=1e-8):
    m = T.shape[0]
    np_ = m - nev
    sub = np.concatenate([[0.0], np.diag(T, -1)])
    dia = np.diag(T).copy()
    sub2, dia2, Q = chase(nev, shifts, sub, dia)
    nT = np.abs(T).max()
    # Q orthogonal, lower bandwidth np
    assert np.abs(Q.T @ Q - np.eye(m)).max() <= 1e-14 * m
    low = np.tril(np.abs(Q), -(np_ + 1))
    assert low.max() == 0.0
    # the updated T is Q^T T Q, which stays tridiagonal
    QTQ = Q.T @ T @ Q
    T2 = np.diag(dia2) + np.diag(sub2[1:], -1) + np.diag(sub2[1:], 1)
    assert np.abs(QTQ - T2).max() <= 1e-13 * nT
    assert np.all(sub2[1:] >= 0.0)
    # implicit restart theorem: V Q e_1 is parallel to prod (A - mu_i) v_1
    p = V[:, 0].copy()
    for mu in shifts:
        p = A @ p - mu * p
        p /= np.linalg.norm(p)
    q1 = V @ Q[:, 0]
    assert abs(abs(p @ q1) / np.linalg.norm(q1) - 1.0) <= tol_dir
    # the truncated factorization A V+ = V+ T+ + r+ e_k^T with the restart's residual
    k = nev
    Vp = V @ Q[:, :k]
    rp = (V @ Q[:, k]) * sub2[k] + r * Q[m - 1, k - 1]
    E = A @ Vp - Vp @ T2[:k, :k]
    E[:, k - 1] -= rp
    assert np.abs(E).max() <= 1e-12 * np.abs(np.linalg.eigvalsh(A)).max()
    return sub2, dia2, Q


@pytest.mark.parametrize("who", sorted(CHASES))
@pytest.mark.parametrize("nev,ncv", [(4, 12), (8, 20), (10, 30)])
def test_restart_chase_implicit_restart_theorem(who, nev, ncv):
    rng = np.random.default_rng(nev * 100 + ncv)
    n = 60
    M = rng.standard_normal((n, n))
    A = (M + M.T) / 2
    V, T, r = lanczos(A, ncv, seed=ncv)
    assert np.abs(A @ V - V @ T - np.outer(r, np.eye(ncv)[-1])).max() < 1e-12
    theta = np.linalg.eigvalsh(T)
    shifts = theta[: ncv - nev]  # exact shifts: the unwanted (smallest) Ritz values
    check_chase(CHASES[who], A, V, T, r, nev, shifts)


@pytest.mark.parametrize("who", sorted(CHASES))
def test_restart_chase_deflation_branch(who):
    # a negligible sub-diagonal entry splits T: the sweep restarts below it
    rng = np.random.default_rng(7)
    m, nev = 14, 6
    d = rng.uniform(-1, 1, m)
    e = rng.uniform(0.1, 1, m - 1)
    e[4] = 1e-20
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    sub = np.concatenate([[0.0], e])
    shifts = np.sort(np.linalg.eigvalsh(T))[: m - nev]
    sub2, dia2, Q = CHASES[who](nev, shifts, sub, d)
    assert sub2[5] == 0.0  # the deflated entry is zeroed
    assert np.abs(Q.T @ Q - np.eye(m)).max() <= 1e-14 * m
    T2 = np.diag(dia2) + np.diag(sub2[1:], -1) + np.diag(sub2[1:], 1)
    assert np.abs(Q.T @ T @ Q - T2).max() <= 1e-13


def test_product_chase_equals_oracle_chase():
    rng = np.random.default_rng(3)
    n, nev, ncv = 80, 8, 24
    M = rng.standard_normal((n, n))
    A = (M + M.T) / 2
    V, T, r = lanczos(A, ncv, seed=1)
    sub = np.concatenate([[0.0], np.diag(T, -1)])
    shifts = np.linalg.eigvalsh(T)[: ncv - nev]
    a = H.host_restart_chase(nev, shifts, sub, np.diag(T))
    b = O.dsapps_chase(nev, shifts, sub, np.diag(T))
    # what the restart keeps: T(0:nev, 0:nev), Q(:, 0:nev) and the residual term
    # sub[nev] Q(:, nev) (HLanc.hpp:1605-1614); the discarded block is what exact
    # shifts leave ill-conditioned, and Q(:, nev) alone is too when sub[nev] ~ 0
    nT = np.abs(T).max()
    np.testing.assert_allclose(a[0][:nev], b[0][:nev], rtol=0, atol=1e-12 * nT)
    np.testing.assert_allclose(a[1][:nev], b[1][:nev], rtol=0, atol=1e-12 * nT)
    np.testing.assert_allclose(a[2][:, :nev], b[2][:, :nev], rtol=0, atol=1e-12)
    np.testing.assert_allclose(a[0][nev] * a[2][:, nev], b[0][nev] * b[2][:, nev], rtol=0, atol=1e-12 * nT)
