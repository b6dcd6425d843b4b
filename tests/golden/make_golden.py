"""Writes the golden fixtures in tests/golden/ (run once, in a container with MKL).

The reference links Intel MKL for every BLAS/LAPACK/sparse call on its host path
(HLanc.hpp:9-12; call sites HLanc.hpp:452, 911-942, 1372-1378, 1514, 1536, 1691,
1805, 1837).  This script calls the installed MKL 2021.4 runtime library
(libmkl_rt.so) through ctypes -- as a library, nothing taken apart -- and stores
its outputs for the oracle's restatements to be checked against.  Dense and
closed-form answers come from numpy.  Output: golden.npz (data only).

    python tests/golden/make_golden.py [--mkl /opt/conda/lib/libmkl_rt.so]
"""
from __future__ import annotations

import argparse
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import refformats as R  # noqa: E402

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def D(a):
    return a.ctypes.data_as(_dp)


def I(v):
    return ctypes.byref(ctypes.c_int(v))


def mkl_dlartg(mkl, f, g):
    c, s, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    mkl.dlartg_(ctypes.byref(ctypes.c_double(f)), ctypes.byref(ctypes.c_double(g)),
                ctypes.byref(c), ctypes.byref(s), ctypes.byref(r))
    return c.value, s.value, r.value


def mkl_dsteqr(mkl, d, e):
    n = len(d)
    d = np.array(d, dtype=np.float64)
    e = np.array(list(e) + [0.0], dtype=np.float64)
    z = np.zeros((n, n), order="F")
    work = np.zeros(max(1, 2 * n - 2))
    info = ctypes.c_int()
    mkl.dsteqr_(ctypes.c_char_p(b"I"), I(n), D(d), D(e), z.ctypes.data_as(_dp), I(n), D(work),
                ctypes.byref(info))
    assert info.value == 0
    return d, z


def mkl_dgeqr2_dorm2r(mkl, a, c):
    m, k = a.shape
    a = np.array(a, order="F")
    tau = np.zeros(k)
    work = np.zeros(max(k, c.shape[0]))
    info = ctypes.c_int()
    mkl.dgeqr2_(I(m), I(k), a.ctypes.data_as(_dp), I(m), D(tau), D(work), ctypes.byref(info))
    assert info.value == 0
    cc = np.array(c, order="F")
    nr, nc = cc.shape
    mkl.dorm2r_(ctypes.c_char_p(b"R"), ctypes.c_char_p(b"N"), I(nr), I(nc), I(k),
                a.ctypes.data_as(_dp), I(m), D(tau), cc.ctypes.data_as(_dp), I(nr), D(work),
                ctypes.byref(info))
    assert info.value == 0
    return a, tau, cc


def mkl_dcsrmv(mkl, rp, col, val, ncols, sym, x):
    n = len(rp) - 1
    y = np.zeros(n)
    desc = ctypes.create_string_buffer(b"SLNC" if sym else b"GLNC", 6)
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    mkl.mkl_dcsrmv(ctypes.c_char_p(b"N"), I(n), I(ncols), ctypes.byref(ctypes.c_double(1.0)), desc,
                   D(np.ascontiguousarray(val)), np.ascontiguousarray(col, np.int32).ctypes.data_as(_ip),
                   rp.ctypes.data_as(_ip), rp[1:].ctypes.data_as(_ip), D(np.ascontiguousarray(x)),
                   ctypes.byref(ctypes.c_double(0.0)), D(y))
    return y


def mkl_dlarnv(mkl, n):
    iseed = (ctypes.c_int * 4)(1, 3, 5, 7)
    x = np.zeros(n)
    mkl.dlarnv_(I(2), iseed, I(n), D(x))
    return x


def random_csr(rng, n, ncols, dens_rows):
    """Messy CSR: empty rows, unsorted and duplicate columns, one heavy row."""
    rp = [0]
    col, val = [], []
    for r in range(n):
        k = int(rng.integers(0, dens_rows)) if r % 7 else 0
        if r == n // 2:
            k = 3 * ncols // 2
        cs = rng.integers(0, ncols, size=k)
        col += list(cs)
        val += list(rng.uniform(-1, 1, size=k))
        rp.append(len(col))
    return np.array(rp, np.int32), np.array(col, np.int32), np.array(val), ncols


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mkl", default="/opt/conda/lib/libmkl_rt.so")
    args = ap.parse_args()
    mkl = ctypes.CDLL(args.mkl)
    rng = np.random.default_rng(20240601)
    out = {}

    # dlartg: sign conventions over all quadrants and the zero cases.
    fg = [(3.0, 4.0), (-3.0, 4.0), (3.0, -4.0), (-3.0, -4.0), (5.0, 1.0), (-5.0, 1.0), (-5.0, -1.0),
          (0.0, 2.0), (0.0, -2.0), (2.0, 0.0), (-2.0, 0.0), (1e-300, 1e-300), (1e300, -1e299)]
    fg += [tuple(x) for x in rng.normal(size=(40, 2)) * np.exp(rng.uniform(-20, 20, size=(40, 1)))]
    out["dlartg_fg"] = np.array(fg)
    out["dlartg_csr"] = np.array([mkl_dlartg(mkl, f, g) for f, g in fg])

    # dsteqr("I") on random and degenerate tridiagonals.
    for tag, n in (("t5", 5), ("t30", 30), ("t60", 60), ("t96", 96)):
        d = rng.normal(size=n)
        e = rng.normal(size=n - 1)
        w, z = mkl_dsteqr(mkl, d, e)
        out[f"steqr_{tag}_d"], out[f"steqr_{tag}_e"], out[f"steqr_{tag}_w"] = d, e, w
        out[f"steqr_{tag}_z"] = z
    d = np.array([2.0, 2.0, 2.0, 1.0, 3.0, 3.0, 0.5, 4.0])
    e = np.array([0.0, 1e-3, 0.0, 0.0, 2.0, 0.0, 1e-9])
    w, z = mkl_dsteqr(mkl, d, e)
    out["steqr_deg_d"], out["steqr_deg_e"], out["steqr_deg_w"], out["steqr_deg_z"] = d, e, w, z

    # dgeqr2 + dorm2r("R","N"): Householder storage and its application.
    a = rng.normal(size=(30, 10))
    c = rng.normal(size=(50, 30))
    af, tau, cc = mkl_dgeqr2_dorm2r(mkl, a, c)
    out["qr_a"], out["qr_af"], out["qr_tau"], out["qr_c"], out["qr_cq"] = a, af, tau, c, cc

    # mkl_dcsrmv exactly as dcsrmv_operator_mkl calls it (HLanc.hpp:446-465).
    for tag, (rp, col, val, nc, sym) in (("simple", R.simple()), ("journals", R.journals())):
        x = rng.uniform(-1, 1, size=nc)
        out[f"spmv_{tag}_x"] = x
        out[f"spmv_{tag}_y"] = mkl_dcsrmv(mkl, rp, col, val, nc, sym, x)
    rp, col, val, nc = random_csr(rng, 301, 257, 40)
    x = rng.uniform(-1, 1, size=nc)
    out.update(spmv_rand_rp=rp, spmv_rand_col=col, spmv_rand_val=val, spmv_rand_x=x,
               spmv_rand_y=mkl_dcsrmv(mkl, rp, col, val, nc, False, x))

    # dgetv0's start vector: dlarnv(idist=2, iseed={1,3,5,7}) (HLanc.hpp:1368-1374).
    out["dlarnv_16"] = mkl_dlarnv(mkl, 16)
    out["dlarnv_124"] = mkl_dlarnv(mkl, 124)

    # Known spectra.
    k = np.arange(1, 17)
    out["simple_eigs"] = np.sort(1.0 / (4.0 * np.sin((2 * k - 1) * np.pi / 66.0) ** 2))
    rp, col, val, nc, sym = R.journals()
    out["journals_eigs"] = np.linalg.eigvalsh(R.to_dense(rp, col, val, nc, sym))

    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
