"""ctypes front end of the CPU oracle (oracle/hlanc_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as part of the product path.  See hlanc_oracle.c for what each
function restates (HLanc.hpp file:line) and what pins it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libhlanc_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
OP_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, _dp, _dp)


class _Csr(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int), ("ncols", ctypes.c_int), ("row_ptr", _ip),
                ("col", _ip), ("val", _dp), ("sym", ctypes.c_int)]


class _Opts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("maxiter", ctypes.c_int), ("v0", _dp),
                ("seed", ctypes.c_uint64), ("nthreads", ctypes.c_int),
                ("max_restarts_then_stop", ctypes.c_int)]


class _Info(ctypes.Structure):
    _fields_ = [("ret", ctypes.c_int), ("restarts", ctypes.c_int), ("nconv", ctypes.c_int),
                ("spmv", ctypes.c_longlong), ("cgs", ctypes.c_longlong), ("rnorm", ctypes.c_double)]


def build() -> None:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.hor_csr_spmv.argtypes = [ctypes.POINTER(_Csr), _dp, _dp]
        L.hor_dlartg.argtypes = [ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]
        L.hor_dsapps_chase.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.hor_tridiag_eig.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
        L.hor_tridiag_eig.restype = ctypes.c_int
        L.hor_dgeqr2.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp]
        L.hor_dorm2r_rn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                    _dp, _dp, ctypes.c_int]
        L.hor_start_vector.argtypes = [ctypes.c_int, ctypes.c_uint64, _dp]
        L.hor_ds_solver.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(_Opts),
                                    ctypes.POINTER(_Info)]
        L.hor_ds_solver.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class Csr:
    """Keeps numpy CSR arrays alive for the C struct."""

    def __init__(self, row_ptr, col, val, ncols=None, sym=False):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.nrows = len(self.row_ptr) - 1
        self.ncols = int(ncols if ncols is not None else self.nrows)
        self.sym = bool(sym)
        self.c = _Csr(self.nrows, self.ncols, self.row_ptr.ctypes.data_as(_ip),
                      self.col.ctypes.data_as(_ip), _d(self.val), int(sym))


def spmv(A: Csr, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(A.nrows)
    lib().hor_csr_spmv(ctypes.byref(A.c), _d(x), _d(y))
    return y


def dlartg(f: float, g: float):
    c, s, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().hor_dlartg(f, g, ctypes.byref(c), ctypes.byref(s), ctypes.byref(r))
    return c.value, s.value, r.value


def dsapps_chase(nev: int, shifts, sub, dia):
    """(sub, dia, Q) after the oracle's dsapps shift sweeps; sub[i] = T(i, i-1)."""
    sub = np.array(sub, dtype=np.float64)
    dia = np.array(dia, dtype=np.float64)
    sh = np.ascontiguousarray(shifts, dtype=np.float64)
    ncv = len(dia)
    q = np.zeros((ncv, ncv), order="F")
    lib().hor_dsapps_chase(nev, ncv - nev, _d(sh) if len(sh) else None, _d(sub), _d(dia), q.ctypes.data_as(_dp))
    return sub, dia, q


def tridiag_eig(d, e):
    d = np.array(d, dtype=np.float64)
    n = len(d)
    e2 = np.zeros(max(n, 1))
    e2[: n - 1] = e
    z = np.zeros((n, n), order="F")
    info = lib().hor_tridiag_eig(n, _d(d), _d(e2), z.ctypes.data_as(_dp), n)
    return info, d, z


def dgeqr2(a):
    a = np.array(a, dtype=np.float64, order="F")
    m, k = a.shape
    tau = np.zeros(k)
    lib().hor_dgeqr2(m, k, a.ctypes.data_as(_dp), m, _d(tau))
    return a, tau


def dorm2r_rn(a, tau, c, k):
    a = np.asfortranarray(a, dtype=np.float64)
    c = np.array(c, dtype=np.float64, order="F")
    nr, nc = c.shape
    lib().hor_dorm2r_rn(nr, nc, k, a.ctypes.data_as(_dp), a.shape[0], _d(np.ascontiguousarray(tau)),
                        c.ctypes.data_as(_dp), nr)
    return c


def start_vector(n: int, seed: int = 0) -> np.ndarray:
    x = np.zeros(n)
    lib().hor_start_vector(n, seed, _d(x))
    return x


@dataclass
class SolveResult:
    ret: int
    values: np.ndarray
    vectors: np.ndarray
    restarts: int
    nconv: int
    spmv: int
    cgs: int
    rnorm: float


def ds_solver(A: Csr, nev: int, ncv: int, tol: float = 0.0, maxiter: int = 0, v0=None, seed: int = 0,
              nthreads: int = 1, max_restarts_then_stop: int = 0) -> SolveResult:
    """ds_solver_mkl restated (HLanc.hpp:1649-1861) over the oracle's CSR operator."""
    n = A.nrows
    values = np.zeros(ncv)
    vectors = np.zeros((n, ncv), order="F")
    v0a = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64)
    o = _Opts(tol, maxiter, _d(v0a) if v0a is not None else None, seed, nthreads,
              max_restarts_then_stop)
    info = _Info()
    fn = ctypes.cast(lib().hor_csr_op, ctypes.c_void_p)
    ret = lib().hor_ds_solver(n, nev, ncv, _d(values), vectors.ctypes.data_as(_dp), n, fn,
                              ctypes.cast(ctypes.pointer(A.c), ctypes.c_void_p), ctypes.byref(o),
                              ctypes.byref(info))
    return SolveResult(ret, values[:nev].copy(), vectors, info.restarts, info.nconv, info.spmv,
                       info.cgs, info.rnorm)
