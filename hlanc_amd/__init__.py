"""Python binding of libhlanc_hip.so, the MI355X implementation of HLanc's IRLM hot path.

The names mirror the reference's C++ interface (HLanc.hpp): ``dcsr_matrix_t``
(HLanc.hpp:74-100), ``dcsrmv_operator_gpu`` (HLanc.hpp:476-572),
``dcsrmv_operator_mgpu`` (HLanc.hpp:750-893, one GPU) and ``ds_solver_gpu``
(HLanc.hpp:1868-1872).  Everything below calls the C ABI declared in
include/hlanc_capi.h through ctypes; there is no CPU fallback: if the library is
not built, importing works but every call raises ``HLancError``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhlanc_hip.so")

HL_OK = 0
HL_NOT_CONVERGED = 1
HL_PREC_F64 = 0
HL_PREC_F32 = 1
HL_PREC_F64_PROJ32 = 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
HOST_OP = ctypes.CFUNCTYPE(None, ctypes.c_void_p, _dp, _dp)


class HLancError(RuntimeError):
    def __init__(self, code: int, msg: str = ""):
        super().__init__(f"hlanc error {code}: {msg}")
        self.code = code


class _Matrix(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int), ("ncols", ctypes.c_int), ("nnz", ctypes.c_int), ("sym", ctypes.c_int),
                ("row_ptr", _ip), ("col", _ip), ("val", _dp)]


class Options(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("maxiter", ctypes.c_int), ("v0", _dp), ("seed", ctypes.c_uint64),
                ("basis_prec", ctypes.c_int), ("use_mfma", ctypes.c_int), ("flags", ctypes.c_int),
                ("refine", ctypes.c_int), ("residuals", _dp), ("reserved", ctypes.c_int * 4)]


class Info(ctypes.Structure):
    _fields_ = [("ret", ctypes.c_int), ("restarts", ctypes.c_int), ("nconv", ctypes.c_int),
                ("graph_sweeps", ctypes.c_int),
                ("spmv", ctypes.c_longlong), ("cgs_passes", ctypes.c_longlong), ("rnorm", ctypes.c_double),
                ("seconds", ctypes.c_double), ("max_rel_resid", ctypes.c_double),
                ("refine_shift", ctypes.c_double), ("spmv_f32", ctypes.c_longlong),
                ("restarts_f32", ctypes.c_int), ("proj64_steps", ctypes.c_int),
                ("anorm_est", ctypes.c_double), ("rq_shift", ctypes.c_double), ("verified", ctypes.c_int),
                ("pad", ctypes.c_int)]


class Profile(ctypes.Structure):
    """hl_profile: per kernel class (0 SpMV, 1 CGS passes, 2 V.Q, 3 other kernels,
    4 DGKS correction passes, 5 host work between sweeps) launches, seconds,
    algorithmic bytes."""
    _fields_ = [("launches", ctypes.c_longlong * 6), ("seconds", ctypes.c_double * 6), ("bytes", ctypes.c_double * 6)]


PROFILE_CLASSES = ["spmv", "cgs", "vq", "other", "dgks", "host"]
HL_FLAG_NO_GRAPH = 1
HL_FLAG_PROFILE = 2
HL_FLAG_NO_VERIFY = 4
HL_REFINE_NONE = 0
HL_REFINE_RQ = 1


class SolveItem(ctypes.Structure):
    """hl_solve_item: one solve of hl_ds_solve_batch."""
    _fields_ = [("A", ctypes.c_void_p), ("nev", ctypes.c_int), ("ncv", ctypes.c_int), ("values", _dp),
                ("vectors", _dp), ("ldv", ctypes.c_int), ("pad0", ctypes.c_int), ("opts", Options),
                ("arrival", ctypes.c_double), ("info", Info), ("status", ctypes.c_int), ("pad1", ctypes.c_int),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]

class ServeResult(ctypes.Structure):
    """hl_serve_result: a finished request of a server."""
    _fields_ = [("status", ctypes.c_int), ("pad", ctypes.c_int), ("info", Info), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
                ("footprint", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("done", ctypes.c_longlong), ("queued", ctypes.c_int),
                ("in_flight", ctypes.c_int), ("bytes_in_use", ctypes.c_double), ("seconds", ctypes.c_double)]


HL_PENDING = 2
HL_SERVE_SELL_EVEN = 1
HL_FORM_ITEMS, HL_FORM_SLICES, HL_FORM_SELL, HL_FORM_STAGED = 0, 1, 2, 3

_lib = None

# (name, restype, argtypes) of every entry point of include/hlanc_capi.h
_SIGS = [
    ("hl_last_error", ctypes.c_char_p, []),
    ("hl_version", ctypes.c_char_p, []),
    ("hl_device_count", ctypes.c_int, [_ip]),
    ("hl_host_alloc", ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    ("hl_host_free", ctypes.c_int, [ctypes.c_void_p]),
    ("hl_matrix_read", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(_Matrix)]),
    ("hl_matrix_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(_Matrix)]),
    ("hl_matrix_free", None, [ctypes.POINTER(_Matrix)]),
    ("hl_csr_create", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp,
                                     ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("hl_csr_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("hl_csr_info", ctypes.c_int, [ctypes.c_void_p, _ip, _ip, _ip, ctypes.POINTER(ctypes.c_longlong)]),
    ("hl_csr_apply", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("hl_csr_spmv", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("hl_csr_spmv_f32val", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("hl_csr_spmv_bytes", ctypes.c_double, [ctypes.c_void_p, ctypes.c_int]),
    ("hl_csr_set_items", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("hl_csr_autotune", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _ip, _ip, ctypes.POINTER(ctypes.c_double)]),
    ("hl_csr_set_slices", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hl_csr_get_slices", ctypes.c_int, [ctypes.c_void_p, _ip, _ip]),
    ("hl_csr_set_sell", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hl_csr_get_sell", ctypes.c_int, [ctypes.c_void_p, _ip, ctypes.POINTER(ctypes.c_longlong)]),
    ("hl_csr_get_sell_col_bytes", ctypes.c_int, [ctypes.c_void_p, _ip]),
    ("hl_csr_pick_form", ctypes.c_int, [ctypes.c_void_p, _ip]),
    ("hl_csr_device_bytes", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    ("hl_csr_set_staged", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hl_csr_get_staged", ctypes.c_int, [ctypes.c_void_p, _ip, ctypes.POINTER(ctypes.c_longlong),
                                         ctypes.POINTER(ctypes.c_longlong), _ip]),
    ("hl_cgs_pass", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_void_p]),
    ("hl_cgs_pass_noalloc", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_longlong, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("hl_vq", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                             ctypes.c_longlong, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong,
                             ctypes.c_int, ctypes.c_void_p]),
    ("hl_tune_set", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    ("hl_tune_get", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]),
    ("hl_bench_cgs", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double)]),
    ("hl_bench_gather", ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int,
                                       ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    ("hl_bench_grid_barrier", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    ("hl_bench_graph_nodes", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    ("hl_bench_stream", ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]),
    ("hl_host_tridiag_eig", ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]),
    ("hl_host_tridiag_eig_last", ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp]),
    ("hl_host_dlartg", None, [ctypes.c_double, ctypes.c_double, _dp, _dp, _dp]),
    ("hl_host_householder_q", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
                                             ctypes.c_int]),
    ("hl_host_restart_chase", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]),
    ("hl_options_default", None, [ctypes.POINTER(Options)]),
    ("hl_profile_last", ctypes.c_int, [ctypes.POINTER(Profile)]),
    ("hl_ds_solve_csr", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                       ctypes.POINTER(Options), ctypes.POINTER(Info)]),
    ("hl_ds_solve_host_op", ctypes.c_int, [ctypes.c_int, ctypes.c_int, HOST_OP, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.POINTER(Options),
                                           ctypes.POINTER(Info)]),
    ("hl_ds_solve_batch", ctypes.c_int, [ctypes.c_int, ctypes.POINTER(SolveItem), ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double]),
    ("hl_serve_create", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    ("hl_serve_submit", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(Options), _dp, _dp, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_longlong)]),
    ("hl_serve_submit_csr", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.POINTER(Options), _dp, _dp,
                                           ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]),
    ("hl_serve_poll", ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ServeResult)]),
    ("hl_serve_wait", ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_double,
                                     ctypes.POINTER(ServeResult)]),
    ("hl_serve_release", ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong]),
    ("hl_serve_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServeStats)]),
    ("hl_serve_destroy", ctypes.c_int, [ctypes.c_void_p]),
]


def build(verbose: bool = False) -> None:
    """Compile libhlanc_hip.so for gfx950 with hipcc (no GPU needed)."""
    subprocess.run(["make", "-s", "-j8", "-C", os.path.join(_HERE, "csrc")], check=True,
                   stdout=None if verbose else subprocess.DEVNULL)


def lib():
    """The loaded C library; raises HLancError when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HLancError(-2, f"{LIB_PATH} is not built; run hlanc_amd.build() or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return [s[0] for s in _SIGS]


def _check(rc: int, allow=(HL_OK,)):
    if rc not in allow:
        raise HLancError(rc, lib().hl_last_error().decode(errors="replace"))
    return rc


def device_count() -> int:
    c = ctypes.c_int()
    lib().hl_device_count(ctypes.byref(c))
    return c.value


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


# ----------------------------------------------------------------- matrices

class dcsr_matrix_t:
    """Host CSR matrix (HLanc.hpp:74-100): reads "pp"/"mm"/"bin" files, or wraps arrays."""

    def __init__(self, format=None, filename=None, is_symmetric=False, *, row_ptr=None, col=None, val=None,
                 ncols=None):
        if format is not None:
            m = _Matrix()
            _check(lib().hl_matrix_read(format.encode(), os.fsencode(filename), int(is_symmetric), ctypes.byref(m)))
            try:
                self.row_ptr = np.ctypeslib.as_array(m.row_ptr, shape=(m.nrows + 1,)).copy()
                self.col = np.ctypeslib.as_array(m.col, shape=(max(m.nnz, 1),))[: m.nnz].copy()
                self.val = np.ctypeslib.as_array(m.val, shape=(max(m.nnz, 1),))[: m.nnz].copy()
                self.num_columns = m.ncols
                self.is_symmetric = bool(m.sym)
            finally:
                lib().hl_matrix_free(ctypes.byref(m))
        else:
            self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
            self.col = np.ascontiguousarray(col, dtype=np.int32)
            self.val = np.ascontiguousarray(val, dtype=np.float64)
            n = len(self.row_ptr) - 1
            self.num_columns = int(ncols if ncols is not None else n)
            self.is_symmetric = bool(is_symmetric)

    def get_num_rows(self):
        return len(self.row_ptr) - 1

    def get_nnz(self):
        return len(self.val)

    def get_num_columns(self):
        return self.num_columns

    def get_row_ptr(self):
        return self.row_ptr

    def get_column_index(self):
        return self.col

    def get_value(self):
        return self.val

    def symmetric(self):
        return self.is_symmetric

    def _as_struct(self):
        self._keep = (np.ascontiguousarray(self.row_ptr, np.int32), np.ascontiguousarray(self.col, np.int32),
                      np.ascontiguousarray(self.val, np.float64))
        rp, c, v = self._keep
        return _Matrix(self.get_num_rows(), self.num_columns, len(v), int(self.is_symmetric), _i(rp), _i(c), _d(v))

    def save(self, format, filename):
        m = self._as_struct()
        _check(lib().hl_matrix_write(format.encode(), os.fsencode(filename), ctypes.byref(m)))

    def save_pp(self, filename):
        self.save("pp", filename)

    def save_mm(self, filename):
        self.save("mm", filename)

    def save_bin(self, filename):
        self.save("bin", filename)


# ----------------------------------------------------------------- operator

class dcsrmv_operator_gpu:
    """y = A x on one MI355X (HLanc.hpp:476-572).  Callable with numpy arrays
    (host staging, synchronous) or, through ``spmv``, with device pointers."""

    def __init__(self, m: dcsr_matrix_t, dev: int = 0, fp32_values: bool = False):
        self.m = m
        self.dev = dev
        h = ctypes.c_void_p()
        rp = np.ascontiguousarray(m.row_ptr, np.int32)
        c = np.ascontiguousarray(m.col, np.int32)
        v = np.ascontiguousarray(m.val, np.float64)
        _check(lib().hl_csr_create(dev, len(rp) - 1, m.num_columns, len(v), _i(rp), _i(c), _d(v),
                                   int(m.is_symmetric), HL_PREC_F32 if fp32_values else HL_PREC_F64,
                                   ctypes.byref(h)))
        self.h = h
        self.nrows = len(rp) - 1
        self.ncols = m.num_columns
        nnz = ctypes.c_longlong()
        lib().hl_csr_info(h, None, None, None, ctypes.byref(nnz))
        self.nnz_general = nnz.value

    def close(self):
        if getattr(self, "h", None):
            lib().hl_csr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __call__(self, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if y is None:
            y = np.empty(self.nrows)
        _check(lib().hl_csr_apply(self.h, x.ctypes.data, y.ctypes.data))
        return y

    def spmv(self, x_ptr: int, y_ptr: int, stream: int = 0, fp32_values: bool = False):
        f = lib().hl_csr_spmv_f32val if fp32_values else lib().hl_csr_spmv
        _check(f(self.h, x_ptr, y_ptr, stream))

    def spmv_bytes(self, fp32_values: bool = False) -> float:
        return lib().hl_csr_spmv_bytes(self.h, HL_PREC_F32 if fp32_values else HL_PREC_F64)

    def set_items(self, nnz_cap: int, rows_cap: int) -> None:
        """Work-item shape of the SpMV (hl_csr_set_items)."""
        _check(lib().hl_csr_set_items(self.h, int(nnz_cap), int(rows_cap)))

    def autotune(self, reps: int = 20):
        """Pick the fastest work-item shape and column slicing (hl_csr_autotune);
        returns (nnz_cap, rows_cap, seconds per SpMV); slices() tells the slicing."""
        nc, rc_, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().hl_csr_autotune(self.h, int(reps), ctypes.byref(nc), ctypes.byref(rc_), ctypes.byref(t)))
        return nc.value, rc_.value, t.value

    def set_slices(self, width: int) -> None:
        """Column-sliced SpMV with slices of `width` columns; 0 drops it (hl_csr_set_slices)."""
        _check(lib().hl_csr_set_slices(self.h, int(width)))

    def slices(self):
        """(slice width, virtual rows) of the sliced form; (0, 0) when unsliced."""
        w, npr = ctypes.c_int(), ctypes.c_int()
        _check(lib().hl_csr_get_slices(self.h, ctypes.byref(w), ctypes.byref(npr)))
        return w.value, npr.value

    def set_sell(self, on: bool) -> None:
        """Sliced-ELL SpMV (64-row chunks, a lane per row) on or off (hl_csr_set_sell)."""
        _check(lib().hl_csr_set_sell(self.h, int(bool(on))))

    def sell(self):
        """(on, stored entries) of the sliced-ELL form."""
        on, ne = ctypes.c_int(), ctypes.c_longlong()
        _check(lib().hl_csr_get_sell(self.h, ctypes.byref(on), ctypes.byref(ne)))
        return bool(on.value), ne.value

    def pick_form(self) -> int:
        """Builds the SpMV form chosen from the matrix's structure, without timing
        (hl_csr_pick_form): HL_FORM_SELL, HL_FORM_STAGED or HL_FORM_ITEMS."""
        f = ctypes.c_int()
        _check(lib().hl_csr_pick_form(self.h, ctypes.byref(f)))
        return f.value

    def device_bytes(self) -> float:
        """Device bytes the operator holds, its SpMV form's arrays included (hl_csr_device_bytes)."""
        b = ctypes.c_double()
        _check(lib().hl_csr_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def sell_col_bytes(self) -> int:
        """Bytes per stored column index of the sliced-ELL form: 2 for banded operators (offsets
        from the row), 4 otherwise, 0 when the form is off."""
        b = ctypes.c_int()
        _check(lib().hl_csr_get_sell_col_bytes(self.h, ctypes.byref(b)))
        return b.value

    def set_staged(self, on: bool) -> None:
        """Staged-slice SpMV (x slices of 16384 columns in LDS, partials per row bin) on or off
        (hl_csr_set_staged)."""
        _check(lib().hl_csr_set_staged(self.h, int(bool(on))))

    def staged(self):
        """(on, pieces, stored entries, row bins) of the staged-slice form."""
        on, npc, ne, nb = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_int()
        _check(lib().hl_csr_get_staged(self.h, ctypes.byref(on), ctypes.byref(npc), ctypes.byref(ne),
                                       ctypes.byref(nb)))
        return bool(on.value), npc.value, ne.value, nb.value


class dcsrmv_operator_mgpu(dcsrmv_operator_gpu):
    """dcsrmv_operator_mgpu (HLanc.hpp:750-893) on a one-GPU build: the first
    device in ``mask`` holds the whole matrix; symmetric storage is refused as
    the reference refuses it (HLanc.hpp:798-800)."""

    def __init__(self, m: dcsr_matrix_t, mask: int = 0xFFFFFFFF):
        if m.symmetric():
            raise HLancError(-1, "dcsrmv_operator_mgpu can only be constructed with a general matrix.")
        ndev = device_count()
        devs = [d for d in range(ndev) if mask & (1 << d)]
        if not devs:
            raise HLancError(-2, "no device selected by mask")
        super().__init__(m, devs[0])


# ------------------------------------------------------------------- solver

class _PinnedBuffer:
    """Page-locked host memory from hl_host_alloc, freed with the last array
    that views it."""

    def __init__(self, count: int):
        p = ctypes.c_void_p()
        _check(lib().hl_host_alloc(8 * max(count, 1), ctypes.byref(p)))
        self.ptr = p.value
        self.__array_interface__ = {"shape": (max(count, 1),), "typestr": "<f8", "data": (self.ptr, False),
                                    "version": 3}

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.hl_host_free(self.ptr)
            self.ptr = None


def pinned_empty(rows: int, cols: int) -> np.ndarray:
    """A column-major float64 rows x cols array in page-locked host memory, for
    the ``vectors`` output: the final copy of the basis then runs at full link
    rate."""
    buf = np.asarray(_PinnedBuffer(rows * cols))
    return buf[: rows * cols].reshape(cols, rows).T


@dataclass
class SolveResult:
    ret: int
    values: np.ndarray
    vectors: np.ndarray
    info: Info
    residuals: Optional[np.ndarray] = None  # per returned pair (the library's residual check)


def options(tol: float = 0.0, maxiter: int = 200, v0=None, seed: int = 0, use_mfma: Optional[bool] = None,
            basis_prec: Optional[int] = None, flags: int = 0, refine: int = HL_REFINE_NONE,
            residuals: Optional[np.ndarray] = None) -> Options:
    """hl_options from the library defaults (hl_options_default).  ``residuals``
    (float64, >= nev entries) receives each returned pair's relative residual."""
    o = Options()
    lib().hl_options_default(ctypes.byref(o))
    o.tol = tol
    o.maxiter = maxiter
    o.seed = seed
    if use_mfma is not None:  # None keeps the library default (hl_options_default)
        o.use_mfma = int(use_mfma)
    if basis_prec is not None:  # None keeps the library default (hl_options_default)
        o.basis_prec = basis_prec
    o.flags = flags
    o.refine = int(refine)
    if v0 is not None:
        o._v0 = np.ascontiguousarray(v0, dtype=np.float64)
        o.v0 = _d(o._v0)
    if residuals is not None:
        if residuals.dtype != np.float64 or not residuals.flags.c_contiguous:
            raise HLancError(-1, "residuals must be a contiguous float64 array")
        o._res = residuals
        o.residuals = _d(residuals)
    return o


def _is_device_array(out) -> bool:
    return hasattr(out, "data_ptr") and bool(getattr(out, "is_cuda", False))


def _vectors_buffer(out, ldv: int, ncv: int):
    if out is None:
        return np.zeros((ncv, ldv)).T  # column-major ldv x ncv
    if _is_device_array(out):  # a device tensor: the vectors stay in HBM
        if (str(out.dtype) != "torch.float64" or out.dim() != 2 or out.shape[0] < ldv or out.shape[1] < ncv
                or out.stride(0) != 1 or out.stride(1) < out.shape[0]):
            raise HLancError(-1, f"device vectors_out must be a column-major float64 tensor of at least "
                                 f"{ldv} x {ncv} (e.g. torch.empty(ncv, ldv).T)")
        return out
    if out.dtype != np.float64 or out.shape[0] < ldv or out.shape[1] < ncv or not out.flags.f_contiguous:
        raise HLancError(-1, f"vectors_out must be a column-major float64 array of at least {ldv} x {ncv}")
    return out


def _vectors_args(vectors):
    """(pointer, leading dimension) of a host array or device tensor from _vectors_buffer."""
    if _is_device_array(vectors):
        return ctypes.cast(vectors.data_ptr(), _dp), vectors.stride(1)
    return vectors.ctypes.data_as(_dp), vectors.shape[0]


def ds_solver_gpu(n: int, nev: int, ncv: int, op, dev: int = 0, *, tol: float = 0.0, maxiter: int = 200, v0=None,
                  seed: int = 0, use_mfma: Optional[bool] = None, ldv: Optional[int] = None,
                  flags: int = 0, vectors_out: Optional[np.ndarray] = None,
                  basis_prec: Optional[int] = None, refine: int = HL_REFINE_NONE) -> SolveResult:
    """ds_solver_gpu (HLanc.hpp:1868-1872).  ``op`` is a dcsrmv_operator_gpu
    (device-resident path) or any callable ``op(x, y)`` on numpy arrays (host
    path, as the reference calls arbitrary operators with host pointers).
    ``ret`` is 0 on success and 1 on the reference's numerical failures; values
    and vectors are written only on success.  ``vectors_out`` (column-major
    float64, at least ldv x ncv: a host array, e.g. from pinned_empty, or a
    device tensor such as ``torch.empty(ncv, ldv, dtype=torch.float64,
    device="cuda").T``, which keeps the result in HBM) receives the vectors
    instead of a fresh array.  ``basis_prec`` (None: the library default,
    HL_PREC_F64, the reference's Gram-Schmidt) selects the basis storage; every
    mode returns FP64 results to ``tol`` (hlanc_capi.h, hl_options.basis_prec).
    ``refine=HL_REFINE_RQ`` returns Rayleigh quotients of the returned vectors.
    ``residuals`` of the result holds the library's per-pair residual check
    (NaN where it did not run)."""
    ldv = n if ldv is None else ldv
    values = np.zeros(ncv)
    vectors = _vectors_buffer(vectors_out, ldv, ncv)
    vptr, ldv = _vectors_args(vectors)
    res = np.full(max(nev, 1), np.nan)
    o = options(tol, maxiter, v0, seed, use_mfma, basis_prec=basis_prec, flags=flags, refine=refine, residuals=res)
    info = Info()
    if isinstance(op, dcsrmv_operator_gpu):
        if op.nrows != n:
            raise HLancError(-1, f"operator has {op.nrows} rows, n={n}")
        rc = lib().hl_ds_solve_csr(op.h, nev, ncv, _d(values), vptr, ldv, ctypes.byref(o), ctypes.byref(info))
    else:
        def tramp(ctx, x, y, _op=op, _n=n):
            xa = np.ctypeslib.as_array(x, shape=(_n,))
            ya = np.ctypeslib.as_array(y, shape=(_n,))
            _op(xa, ya)

        cb = HOST_OP(tramp)
        rc = lib().hl_ds_solve_host_op(dev, n, cb, None, nev, ncv, _d(values), vptr, ldv, ctypes.byref(o),
                                       ctypes.byref(info))
    _check(rc, allow=(HL_OK, HL_NOT_CONVERGED))
    vecs = vectors[:n] if _is_device_array(vectors) else np.asarray(vectors[:n])
    return SolveResult(rc, values[:nev].copy(), vecs, info, res[:nev].copy())


@dataclass
class BatchResult:
    status: int
    values: np.ndarray
    vectors: Optional[np.ndarray]
    info: Info
    t_admit: float
    t_done: float
    arrival: float
    residuals: Optional[np.ndarray] = None  # per returned pair (the library's residual check)

    @property
    def latency(self) -> float:
        """Seconds from arrival to completion."""
        return self.t_done - self.arrival


def ds_solve_batch(ops, nev: int, ncv: int, dev: int = 0, *, tol: float = 0.0, maxiter: int = 200, seeds=None,
                   v0s=None, use_mfma: Optional[bool] = None, vectors: bool = True, arrivals=None,
                   max_active: int = 0, hbm_budget: float = 0.0, flags: int = 0, vectors_out=None,
                   basis_prec=None, refine=HL_REFINE_NONE):
    """Many solves on one device with continuous batching (hl_ds_solve_batch).

    ``ops`` are dcsrmv_operator_gpu on ``dev``; nev, ncv, tol, maxiter,
    use_mfma and basis_prec are one value for all solves or one per solve.  ``arrivals`` (seconds after the
    call, non-decreasing) delays admission, as a serving loop sees requests
    arrive.  Each result is bitwise what ds_solver_gpu returns for the same
    operator and options.  ``vectors_out`` optionally lists one vectors buffer
    per solve (as ds_solver_gpu's)."""
    k = len(ops)
    per = lambda x, i: x[i] if isinstance(x, (list, tuple, np.ndarray)) else x  # noqa: E731
    items = (SolveItem * k)()
    keep = []
    for i, op in enumerate(ops):
        if not isinstance(op, dcsrmv_operator_gpu):
            raise HLancError(-1, "ds_solve_batch takes dcsrmv_operator_gpu operators")
        n = op.nrows
        nev_i, ncv_i = int(per(nev, i)), int(per(ncv, i))
        vals = np.zeros(max(ncv_i, 1))
        vecs = None
        if vectors:
            vecs = _vectors_buffer(None if vectors_out is None else vectors_out[i], n, max(ncv_i, 1))
        seed = 0 if seeds is None else int(seeds[i])
        v0 = None if v0s is None else v0s[i]
        res = np.full(max(nev_i, 1), np.nan)
        o = options(float(per(tol, i)), int(per(maxiter, i)), v0, seed, per(use_mfma, i),
                    basis_prec=per(basis_prec, i), flags=flags, refine=per(refine, i), residuals=res)
        it = items[i]
        it.A = op.h
        it.nev = nev_i
        it.ncv = ncv_i
        it.values = _d(vals)
        it.vectors, it.ldv = _vectors_args(vecs) if vecs is not None else (None, n)
        it.opts = o
        it.arrival = 0.0 if arrivals is None else float(arrivals[i])
        keep.append((vals, vecs, o, res))
    rc = lib().hl_ds_solve_batch(dev, items, k, int(max_active), float(hbm_budget))
    _check(rc)
    out = []
    for i in range(k):
        it = items[i]
        vals, vecs, _, res = keep[i]
        info = Info.from_buffer_copy(it.info)
        if vecs is not None:
            vecs = vecs[: ops[i].nrows] if _is_device_array(vecs) else np.asarray(vecs[: ops[i].nrows])
        out.append(BatchResult(it.status, vals[:it.nev].copy(), vecs, info, it.t_admit, it.t_done, it.arrival,
                               res[:it.nev].copy()))
    return out


# ------------------------------------------------------------------ serving

@dataclass
class ServeOutcome:
    status: int
    values: np.ndarray
    vectors: Optional[np.ndarray]
    residuals: np.ndarray
    info: Info
    t_submit: float
    t_admit: float
    t_start: float
    t_done: float
    footprint: float

    @property
    def latency(self) -> float:
        """Seconds from submission to completion."""
        return self.t_done - self.t_submit


class Server:
    """The serving loop (hl_serve_*): requests submitted one by one while others
    run, continuously batched on one device, admitted by HBM footprint.  Each
    request's result is bitwise what ds_solver_gpu returns for the same operator
    and options."""

    def __init__(self, dev: int = 0, max_active: int = 0, hbm_budget: float = 0.0, sell_even: bool = False):
        h = ctypes.c_void_p()
        _check(lib().hl_serve_create(dev, int(max_active), float(hbm_budget),
                                     HL_SERVE_SELL_EVEN if sell_even else 0, ctypes.byref(h)))
        self.h = h
        self._req = {}  # ticket -> (values, vectors, residuals, options, kept host arrays)

    def submit(self, A, nev: int, ncv: int, *, tol: float = 0.0, maxiter: int = 200, seed: int = 0, v0=None,
               use_mfma: Optional[bool] = None, basis_prec: Optional[int] = None, refine: int = HL_REFINE_NONE,
               flags: int = 0, vectors: bool = True) -> int:
        """A request on ``A``: a dcsrmv_operator_gpu on the server's device, or a host
        matrix (dcsr_matrix_t or a (row_ptr, col, val) triple) that the server uploads
        after admission.  Returns the ticket."""
        res = np.full(max(nev, 1), np.nan)
        o = options(tol, maxiter, v0, seed, use_mfma, basis_prec=basis_prec, flags=flags, refine=refine,
                    residuals=res)
        t = ctypes.c_longlong()
        if isinstance(A, dcsrmv_operator_gpu):
            n, keep = A.nrows, A
            vals = np.zeros(max(ncv, 1))
            vecs = np.zeros((max(ncv, 1), n)).T if vectors else None
            rc = lib().hl_serve_submit(self.h, A.h, nev, ncv, ctypes.byref(o), _d(vals),
                                       vecs.ctypes.data_as(_dp) if vecs is not None else None, n, ctypes.byref(t))
        else:
            if isinstance(A, dcsr_matrix_t):
                rp, col, val, sym = A.row_ptr, A.col, A.val, A.is_symmetric
            else:
                (rp, col, val), sym = A, False
            rp = np.ascontiguousarray(rp, np.int32)
            col = np.ascontiguousarray(col, np.int32)
            val = np.ascontiguousarray(val, np.float64)
            n, keep = len(rp) - 1, (rp, col, val)
            vals = np.zeros(max(ncv, 1))
            vecs = np.zeros((max(ncv, 1), n)).T if vectors else None
            rc = lib().hl_serve_submit_csr(self.h, n, len(val), _i(rp), _i(col), _d(val), int(bool(sym)), nev, ncv,
                                           ctypes.byref(o), _d(vals),
                                           vecs.ctypes.data_as(_dp) if vecs is not None else None, n,
                                           ctypes.byref(t))
        _check(rc)
        self._req[t.value] = (vals, vecs, res, o, keep, nev, n)
        return t.value

    def _outcome(self, ticket, r: ServeResult) -> ServeOutcome:
        vals, vecs, res, _, _, nev, n = self._req[ticket]
        return ServeOutcome(r.status, vals[:nev].copy(), None if vecs is None else np.asarray(vecs[:n]),
                            res[:nev].copy(), Info.from_buffer_copy(r.info), r.t_submit, r.t_admit, r.t_start,
                            r.t_done, r.footprint)

    def poll(self, ticket: int) -> Optional[ServeOutcome]:
        r = ServeResult()
        rc = _check(lib().hl_serve_poll(self.h, ticket, ctypes.byref(r)), allow=(HL_OK, HL_PENDING))
        return None if rc == HL_PENDING else self._outcome(ticket, r)

    def wait(self, ticket: int, timeout: Optional[float] = None) -> Optional[ServeOutcome]:
        r = ServeResult()
        rc = _check(lib().hl_serve_wait(self.h, ticket, -1.0 if timeout is None else float(timeout),
                                        ctypes.byref(r)), allow=(HL_OK, HL_PENDING))
        return None if rc == HL_PENDING else self._outcome(ticket, r)

    def release(self, ticket: int) -> None:
        """Forget a finished request (its host arrays and outputs are dropped)."""
        _check(lib().hl_serve_release(self.h, ticket))
        self._req.pop(ticket, None)

    def stats(self) -> ServeStats:
        st = ServeStats()
        lib().hl_serve_stats(self.h, ctypes.byref(st))
        return st

    def close(self) -> None:
        """Finish every submitted request, then stop the server."""
        if getattr(self, "h", None):
            rc = lib().hl_serve_destroy(self.h)
            self.h = None
            _check(rc)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- basis operations

def cgs_pass(dev, n, m, V_ptr, ldv, w_in_ptr, w_out_ptr, h_in_ptr, h_out_ptr, nrm2_ptr, stream=0):
    """One fused CGS pass on device pointers (hl_cgs_pass)."""
    _check(lib().hl_cgs_pass(dev, n, m, V_ptr, ldv, w_in_ptr, w_out_ptr, h_in_ptr, h_out_ptr, nrm2_ptr, stream))


def vq(dev, n, K, N, V_ptr, ldv, Q_ptr, ldq, Out_ptr, ldo, use_mfma=True, stream=0):
    """Out = V Q on device pointers (hl_vq)."""
    _check(lib().hl_vq(dev, n, K, N, V_ptr, ldv, Q_ptr, ldq, Out_ptr, ldo, int(use_mfma), stream))


# --------------------------------------------------- host math of the restart

def host_tridiag_eig(d, e, last_row_only: bool = False):
    """(info, ascending eigenvalues, eigenvectors or their last row) of the
    symmetric tridiagonal (d, e) by the library's host QR (hl_host_tridiag_eig)."""
    d = np.array(d, dtype=np.float64)
    n = len(d)
    e2 = np.zeros(max(n, 1))
    e2[: max(n - 1, 0)] = e
    if last_row_only:
        z = np.zeros(max(n, 1))
        info = lib().hl_host_tridiag_eig_last(n, _d(d), _d(e2), _d(z))
        return info, d, z[:n]
    z = np.zeros((n, n), order="F")
    info = lib().hl_host_tridiag_eig(n, _d(d), _d(e2), z.ctypes.data_as(_dp), max(n, 1))
    return info, d, z


def host_dlartg(f: float, g: float):
    c, s, r = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().hl_host_dlartg(f, g, ctypes.byref(c), ctypes.byref(s), ctypes.byref(r))
    return c.value, s.value, r.value


def host_householder_q(a):
    """(factored a, tau, Q = H(0)...H(k-1) as m x m) of the library's Householder QR."""
    a = np.array(a, dtype=np.float64, order="F")
    m, k = a.shape
    tau = np.zeros(max(k, 1))
    q = np.zeros((m, m), order="F")
    _check(lib().hl_host_householder_q(m, k, a.ctypes.data_as(_dp), m, _d(tau), q.ctypes.data_as(_dp), m))
    return a, tau[:k], q


def host_restart_chase(nev: int, shifts, sub, dia):
    """(sub, dia, Q) after the library's dsapps shift sweeps (hl_host_restart_chase);
    sub[i] = T(i, i-1) with sub[0] unused."""
    sub = np.array(sub, dtype=np.float64)
    dia = np.array(dia, dtype=np.float64)
    sh = np.ascontiguousarray(shifts, dtype=np.float64)
    ncv = len(dia)
    q = np.zeros((ncv, ncv), order="F")
    _check(lib().hl_host_restart_chase(nev, ncv - nev, _d(sh) if len(sh) else None, _d(sub), _d(dia),
                                       q.ctypes.data_as(_dp)))
    return sub, dia, q


# ------------------------------------------------------------------ tuning

def tune_set(key: str, value: int) -> None:
    _check(lib().hl_tune_set(key.encode(), int(value)))


def tune_get(key: str) -> int:
    v = ctypes.c_int()
    _check(lib().hl_tune_get(key.encode(), ctypes.byref(v)))
    return v.value


def bench_grid_barrier(blocks: int, barriers: int = 1000, dev: int = 0) -> float:
    """Seconds per grid-wide barrier of a cooperative launch (hl_bench_grid_barrier)."""
    t = ctypes.c_double()
    _check(lib().hl_bench_grid_barrier(dev, int(blocks), int(barriers), ctypes.byref(t)))
    return t.value


def bench_graph_nodes(nodes: int = 20, blocks: int = 256, reps: int = 100, dev: int = 0) -> float:
    """Seconds per node of a replayed graph of empty kernels (hl_bench_graph_nodes)."""
    t = ctypes.c_double()
    _check(lib().hl_bench_graph_nodes(dev, int(nodes), int(blocks), int(reps), ctypes.byref(t)))
    return t.value


def bench_cgs(n: int, m: int, tiled: bool = True, reps: int = 20, dev: int = 0) -> float:
    """Seconds per launch of one fused CGS pass over an n x m FP64 basis."""
    t = ctypes.c_double()
    _check(lib().hl_bench_cgs(dev, n, m, int(tiled), reps, ctypes.byref(t)))
    return t.value


def bench_stream(nbytes: int = 4 << 30, reps: int = 10, dev: int = 0) -> float:
    """Seconds per read-only streaming pass over nbytes of HBM."""
    t = ctypes.c_double()
    _check(lib().hl_bench_stream(dev, int(nbytes), reps, ctypes.byref(t)))
    return t.value


def bench_gather(nx: int, m: int, random: bool = True, reps: int = 10, dev: int = 0) -> float:
    """Seconds per pass of m gathers from a vector of nx doubles (hl_bench_gather)."""
    t = ctypes.c_double()
    _check(lib().hl_bench_gather(dev, int(nx), int(m), int(random), reps, ctypes.byref(t)))
    return t.value
