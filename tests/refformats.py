"""Test-side readers for the two matrix files in tests/golden/.

They follow dcsr_matrix_t's reading rules (HLanc.hpp:119-247): zero values are
dropped, a symmetric file keeps only its lower triangle, and columns stay in
file order within a row.  They exist so the oracle and the product's own
readers are both checked against a third, independent parse.
"""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_pp(path, is_symmetric=False):
    """HLanc.hpp:119-169: one row per line of `col:val` tokens."""
    rows = []
    with open(path) as f:
        for line in f:
            toks = line.split()
            r = len(rows)
            row = []
            for t in toks:
                c, v = t.split(":")
                c, v = int(c), float(v)
                if v == 0.0:
                    continue
                if is_symmetric and c > r:
                    continue
                row.append((c, v))
            rows.append(row)
    return _pack(rows, ncols=None, sym=is_symmetric)


def read_mm(path):
    """HLanc.hpp:171-247: coordinate Matrix Market, lower triangle if symmetric."""
    with open(path) as f:
        banner = f.readline()
        sym = "symmetric" in banner
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        nr, nc, ne = (int(x) for x in line.split())
        rows = [[] for _ in range(nr)]
        data = np.loadtxt(f, ndmin=2)
    for r, c, v in data[:ne]:
        r, c = int(r) - 1, int(c) - 1
        if v == 0.0 or (sym and c > r):
            continue
        rows[r].append((c, v))
    return _pack(rows, ncols=nc, sym=sym)


def _pack(rows, ncols, sym):
    rp = np.zeros(len(rows) + 1, dtype=np.int32)
    cols, vals = [], []
    for i, row in enumerate(rows):
        rp[i + 1] = rp[i] + len(row)
        cols += [c for c, _ in row]
        vals += [v for _, v in row]
    col = np.array(cols, dtype=np.int32)
    val = np.array(vals, dtype=np.float64)
    if ncols is None:
        ncols = max(int(col.max()) + 1 if len(col) else 0, len(rows) if sym else 0)
    return rp, col, val, int(ncols), bool(sym)


def to_dense(rp, col, val, ncols, sym):
    n = len(rp) - 1
    A = np.zeros((n, ncols))
    for r in range(n):
        for p in range(rp[r], rp[r + 1]):
            A[r, col[p]] += val[p]
            if sym and col[p] != r:
                A[col[p], r] += val[p]
    return A


def simple():
    return read_pp(os.path.join(GOLDEN, "simple.pp"))


def journals():
    return read_mm(os.path.join(GOLDEN, "Journals.mtx"))
