"""Seeded synthetic matrices for the solve workloads (all symmetric, FP64 CSR, int32 indices).

* ``laplacian_2d(m)``: 5-point Dirichlet Laplacian on an m x m grid (diagonal 4,
  off-diagonal -1); the 256 x 256 case has n = 65,536 and nnz = 326,656.
* ``chung_lu_laplacian(n, ...)``: normalized Laplacian I - D^-1/2 W D^-1/2 of a
  symmetric Chung-Lu power-law graph (exponent 2.1, mean degree 16, max expected
  degree 2e4), edge weights U(0,1], restricted to its giant component (the
  smaller components are mostly trees, and every bipartite component adds an
  exact eigenvalue 2, which would make the wanted spectrum degenerate).
* ``stencil27(m, seed)``: 27-point stencil on an m^3 grid, off-diagonal
  -(1 + 0.1 u), u ~ U(0,1) per undirected pair, diagonal = row abs-sum + 1e-3;
  nnz = (3m - 2)^3.
* ``erdos_renyi(n, deg, seed)``: symmetric random graph with mean degree
  ``deg``, weights U(-1,1), diagonal shifted to make it positive definite.
* ``stencil5(m, seed)``: 5-point stencil on an m x m grid with the coefficient
  rule of ``stencil27``.
* ``batch_mix(count, seed)``: ``count`` instances with n log-uniform in
  [1e3, 2e5]: a quarter 2-D 5-point and a quarter 3-D 27-point stencils, half
  Erdos-Renyi graphs with degree uniform in 8..32.
* ``serve_mix(count, seed)``: a pool of request matrices with n log-uniform in
  [4e3, 2e6]: a quarter 27-point stencils, a quarter Chung-Lu Laplacians, half
  Erdos-Renyi graphs, by seed.

Each returns ``(row_ptr, col, val)`` as numpy arrays of a general (both
triangles stored) CSR matrix with sorted column indices.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _csr(A) -> tuple:
    A = A.tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return (A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))


def laplacian_2d(m: int):
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(m, m))
    I = sp.eye(m)
    return _csr(sp.kron(T, I) + sp.kron(I, T))


def laplacian_2d_eigs(m: int) -> np.ndarray:
    """Closed form: 4 - 2 cos(p pi/(m+1)) - 2 cos(q pi/(m+1)), ascending."""
    c = 2.0 - 2.0 * np.cos(np.arange(1, m + 1) * np.pi / (m + 1))
    return np.sort((c[:, None] + c[None, :]).ravel())


def stencil27(m: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    n = m ** 3
    idx = np.arange(n, dtype=np.int64).reshape(m, m, m)
    rows, cols, vals = [], [], []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                if (dx, dy, dz) <= (0, 0, 0):
                    continue  # one direction per undirected pair
                sl_a = tuple(slice(max(0, -d), m - max(0, d)) for d in (dx, dy, dz))
                sl_b = tuple(slice(max(0, d), m - max(0, -d)) for d in (dx, dy, dz))
                a = idx[sl_a].ravel()
                b = idx[sl_b].ravel()
                v = -(1.0 + 0.1 * rng.random(a.size))
                rows += [a, b]
                cols += [b, a]
                vals += [v, v]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    diag = np.bincount(r, weights=np.abs(v), minlength=n) + 1e-3
    r = np.concatenate([r, np.arange(n)])
    c = np.concatenate([c, np.arange(n)])
    v = np.concatenate([v, diag])
    return _csr(sp.coo_matrix((v, (r, c)), shape=(n, n)))


def chung_lu_laplacian(n: int, mean_degree: float = 16.0, exponent: float = 2.1, max_degree: float = 2.0e4,
                       seed: int = 0, giant_only: bool = True):
    from scipy.sparse.csgraph import connected_components

    rng = np.random.default_rng(seed)
    gamma = 1.0 / (exponent - 1.0)
    wmax = min(max_degree, 0.5 * n)
    # expected degrees w_i = wmax (1 + i/i0)^-gamma with i0 set for the mean
    lo, hi = 1e-6, float(n)
    i = np.arange(n, dtype=np.float64)
    for _ in range(100):
        i0 = np.sqrt(lo * hi)
        mean = (wmax * (1.0 + i / i0) ** -gamma).mean()
        if mean > mean_degree:
            hi = i0
        else:
            lo = i0
    w = wmax * (1.0 + i / i0) ** -gamma
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    m_edges = int(round(w.sum() / 2))
    a = np.searchsorted(cdf, rng.random(m_edges)).astype(np.int64)
    b = np.searchsorted(cdf, rng.random(m_edges)).astype(np.int64)
    keep = a != b
    a, b = np.minimum(a[keep], b[keep]), np.maximum(a[keep], b[keep])
    key = np.unique(a * n + b)
    a, b = key // n, key % n
    wt = 1.0 - rng.random(a.size)  # U(0, 1]
    W = sp.coo_matrix((np.concatenate([wt, wt]), (np.concatenate([a, b]), np.concatenate([b, a]))),
                      shape=(n, n)).tocsr()
    if giant_only:
        _, lab = connected_components(W, directed=False)
        big = np.argmax(np.bincount(lab))
        sel = np.flatnonzero(lab == big)
        W = W[sel][:, sel]
    d = np.asarray(W.sum(axis=1)).ravel()
    dinv = 1.0 / np.sqrt(d)
    Dh = sp.diags(dinv)
    L = sp.eye(W.shape[0]) - Dh @ W @ Dh
    return _csr(L)


def erdos_renyi(n: int, deg: float, seed: int = 0):
    rng = np.random.default_rng(seed)
    m_edges = int(round(n * deg / 2))
    a = rng.integers(0, n, m_edges)
    b = rng.integers(0, n, m_edges)
    keep = a != b
    a, b = a[keep], b[keep]
    v = rng.uniform(-1.0, 1.0, a.size)
    A = sp.coo_matrix((np.concatenate([v, v]), (np.concatenate([a, b]), np.concatenate([b, a]))), shape=(n, n))
    A = A.tocsr()
    A.sum_duplicates()
    shift = np.abs(A).sum(axis=1).A.ravel().max() + 1.0
    return _csr(A + sp.eye(n) * shift)


def stencil5(m: int, seed: int = 0):
    rng = np.random.default_rng(seed)
    n = m * m
    idx = np.arange(n, dtype=np.int64).reshape(m, m)
    rows, cols, vals = [], [], []
    for a, b in ((idx[:-1, :].ravel(), idx[1:, :].ravel()), (idx[:, :-1].ravel(), idx[:, 1:].ravel())):
        v = -(1.0 + 0.1 * rng.random(a.size))
        rows += [a, b]
        cols += [b, a]
        vals += [v, v]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals)
    diag = np.bincount(r, weights=np.abs(v), minlength=n) + 1e-3
    r = np.concatenate([r, np.arange(n)])
    c = np.concatenate([c, np.arange(n)])
    v = np.concatenate([v, diag])
    return _csr(sp.coo_matrix((v, (r, c)), shape=(n, n)))


def mix_instance(kind: str, n: float, seed: int, deg: int = 16):
    """One instance of the batch / serving mixes: kind 'st5', 'st27', 'er' or
    'cl' (Chung-Lu normalized Laplacian)."""
    if kind == "st5":
        return stencil5(max(2, int(round(np.sqrt(n)))), seed)
    if kind == "st27":
        return stencil27(max(2, int(round(np.cbrt(n)))), seed)
    if kind == "cl":
        return chung_lu_laplacian(int(round(n)), seed=seed)
    return erdos_renyi(int(round(n)), deg, seed)


def _build(kinds, ns, seeds, degs, threads):
    """The instances, built on `threads` host threads (each from its own seed, so
    the result does not depend on the thread count)."""
    def one(i):
        return kinds[i], mix_instance(kinds[i], ns[i], int(seeds[i]), int(degs[i]))

    if threads <= 1:
        return [one(i) for i in range(len(kinds))]
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, range(len(kinds))))


def batch_mix(count: int = 512, seed: int = 0, nmin: float = 1e3, nmax: float = 2e5, threads: int = 1):
    """[(kind, (row_ptr, col, val))] of the batched uneven workload."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))
    degs = rng.integers(8, 33, count)
    kinds = [("st5", "er", "st27", "er")[i % 4] for i in range(count)]
    return _build(kinds, ns, [seed * 100003 + i for i in range(count)], degs, threads)


def serve_mix(count: int = 32, seed: int = 0, nmin: float = 4e3, nmax: float = 2e6, threads: int = 1):
    """[(kind, (row_ptr, col, val))]: the request matrices of the serving stream."""
    rng = np.random.default_rng(seed + 7)
    ns = np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))
    degs = rng.integers(8, 33, count)
    kinds = [("st27", "er", "cl", "er")[i % 4] for i in range(count)]
    return _build(kinds, ns, [seed * 100019 + i for i in range(count)], degs, threads)
