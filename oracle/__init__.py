"""CPU parity oracle for the CSR SpMV / batched SpMV / BLAS-1 / CG path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import this package, and only as
the checker (or the timed CPU baseline).  The product package ``num_amd``
never imports it.

The arithmetic lives in ``oracle.c`` (a C restatement of the reference's
Kokkos CPU path; every function there cites the reference file:line it
follows).  This module is a thin numpy/ctypes wrapper around ``liboracle.so``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_set_mutant.argtypes = [C.c_int]
        L.orc_get_mutant.restype = C.c_int
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_spmv.argtypes = [C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_spmv_f32.argtypes = [C.c_int32, _i32p, _i32p, _f32p, _f32p, _f32p]
        L.orc_batch_spmv.argtypes = [C.c_int32, C.c_int32, C.c_int32, _i32p, _i32p, _f64p,
                                     _f64p, _f64p]
        L.orc_axpy.argtypes = [C.c_int64, C.c_double, _f64p, _f64p, _f64p]
        L.orc_axpy2.argtypes = [C.c_int64, C.c_double, _f64p, C.c_double, _f64p, _f64p]
        L.orc_dot.argtypes = [C.c_int64, _f64p, _f64p]
        L.orc_dot.restype = C.c_double
        L.orc_norm2.argtypes = [C.c_int64, _f64p]
        L.orc_norm2.restype = C.c_double
        L.orc_copy.argtypes = [C.c_int64, _f64p, _f64p]
        L.orc_fill.argtypes = [C.c_int64, _f64p, C.c_double]
        L.orc_scale.argtypes = [C.c_int64, _f64p, C.c_double]
        L.orc_cg_solve_local.argtypes = [C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p,
                                         C.c_double, C.c_int32, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.orc_cg_solve_local.restype = C.c_int
        L.orc_cg_mixed.argtypes = [C.c_int32, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double,
                                   C.c_int32, C.c_double, C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                   C.POINTER(C.c_int32)]
        L.orc_cg_mixed.restype = C.c_int
        _lib = L
    return _lib


class mutant:
    """Context manager planting mistake ``m`` in the oracle (tests only)."""

    def __init__(self, m: int):
        self.m = m

    def __enter__(self):
        lib().orc_set_mutant(self.m)
        return self

    def __exit__(self, *a):
        lib().orc_set_mutant(0)


def set_threads(t: int) -> None:
    lib().orc_set_threads(int(t))


def threads() -> int:
    return int(lib().orc_get_threads())


def _csr(rowptr, colind, vals):
    return (np.ascontiguousarray(rowptr, np.int32), np.ascontiguousarray(colind, np.int32),
            np.ascontiguousarray(vals, np.float64))


def spmv(rowptr, colind, vals, x, nrows=None):
    rp, ci, v = _csr(rowptr, colind, vals)
    n = len(rp) - 1 if nrows is None else nrows
    y = np.empty(n, np.float64)
    lib().orc_spmv(n, rp, ci, v, np.ascontiguousarray(x, np.float64), y)
    return y


def spmv_f32(rowptr, colind, vals32, x32):
    rp = np.ascontiguousarray(rowptr, np.int32)
    n = len(rp) - 1
    y = np.empty(n, np.float32)
    lib().orc_spmv_f32(n, rp, np.ascontiguousarray(colind, np.int32),
                       np.ascontiguousarray(vals32, np.float32),
                       np.ascontiguousarray(x32, np.float32), y)
    return y


def batch_spmv(rowptr, colind, vals, X):
    rp, ci, v = _csr(rowptr, colind, vals)
    X = np.ascontiguousarray(X, np.float64)
    n = len(rp) - 1
    ncols, k = X.shape
    Y = np.empty((n, k), np.float64)
    lib().orc_batch_spmv(n, ncols, k, rp, ci, v, X, Y)
    return Y


def axpy(alpha, x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    z = np.empty_like(x)
    lib().orc_axpy(len(x), alpha, x, y, z)
    return z


def axpy2(alpha, x, beta, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    z = np.empty_like(x)
    lib().orc_axpy2(len(x), alpha, x, beta, y, z)
    return z


def dot(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return float(lib().orc_dot(len(x), x, y))


def norm2(x):
    x = np.ascontiguousarray(x, np.float64)
    return float(lib().orc_norm2(len(x), x))


def copy(src, dst):
    """dst[:] = src (orc_copy, vec_ops.hpp:74-79); returns dst."""
    src = np.ascontiguousarray(src, np.float64)
    if not (isinstance(dst, np.ndarray) and dst.dtype == np.float64 and dst.flags.c_contiguous
            and dst.shape == src.shape):
        raise ValueError("copy: dst must be a contiguous float64 array shaped like src")
    lib().orc_copy(len(src), src, dst)
    return dst


def fill(x, val):
    """x[:] = val in place (orc_fill, vec_ops.hpp:84-88); returns x."""
    if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous):
        raise ValueError("fill: x must be a contiguous float64 array")
    lib().orc_fill(len(x), x, float(val))
    return x


def scale(x, alpha):
    x = np.array(x, np.float64, copy=True)
    lib().orc_scale(len(x), x, alpha)
    return x


def cg_solve_local(rowptr, colind, vals, b, x0=None, tol=1e-10, max_iter=1000):
    """Returns (x, {'iters','residual','converged'}) like bindings.cpp:283-287."""
    rp, ci, v = _csr(rowptr, colind, vals)
    n = len(rp) - 1
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n, np.float64) if x0 is None else np.array(x0, np.float64, copy=True)
    it, res, conv = C.c_int32(), C.c_double(), C.c_int32()
    rc = lib().orc_cg_solve_local(n, rp, ci, v, b, x, tol, max_iter, C.byref(it), C.byref(res),
                                  C.byref(conv))
    if rc != 0:
        raise MemoryError("oracle cg: allocation failed")
    return x, {"iters": it.value, "residual": res.value, "converged": bool(conv.value)}


def cg_mixed(rowptr, colind, vals, b, x0=None, tol=1e-10, max_iter=100000, delta=0.1):
    """Reliable-update mixed-precision CG; returns (x, {'outer' (FP64
    refinements), 'inner_total' (FP32 steps), 'true_residual', 'converged',
    'breakdown'})."""
    rp, ci, v = _csr(rowptr, colind, vals)
    n = len(rp) - 1
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n, np.float64) if x0 is None else np.array(x0, np.float64, copy=True)
    oi, tot, res, conv = C.c_int32(), C.c_int32(), C.c_double(), C.c_int32()
    rc = lib().orc_cg_mixed(n, rp, ci, v, b, x, tol, max_iter, delta, C.byref(oi),
                            C.byref(tot), C.byref(res), C.byref(conv))
    if rc < 0:
        raise MemoryError("oracle mixed cg: allocation failed")
    return x, {"outer": oi.value, "inner_total": tot.value, "true_residual": res.value,
               "converged": bool(conv.value), "breakdown": rc == 1}
