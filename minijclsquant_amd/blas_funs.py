"""The vector helpers of the reference namespace, computed by the HIP library.

The reference's package namespace carries the OpenMP vector loops of
jclsquant/cython_modules/blas_funs.pyx (scal :15-25, axpy :31-43, axpy2
:47-59, aAxby :64-75, vdot :80-91, dot :94-105, vec_mul :113-121, vec_mul2
:125-137, vec_mul3 :145-157) through the star imports of observables.py:14,
recurrence_functions.py:3-4 and modifiers.py:9.  These functions keep those
names, argument orders and in-place semantics on host complex128 arrays (the
first ``size`` elements; ``n_threads`` is accepted and unused) and run on the
device through jkpm_zvec / jkpm_zdot / jkpm_aAxby.  Where the reference reads
out of bounds (``size`` beyond an array) or on a wrong dtype raises from its
typed memoryviews, these raise ValueError.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat

__all__ = ["scal", "axpy", "axpy2", "aAxby", "vdot", "dot", "vec_mul", "vec_mul2", "vec_mul3"]


def _host(name, a, size=None, out=False):
    if not isinstance(a, np.ndarray) or a.dtype != np.complex128 or a.ndim != 1 \
            or not a.flags.c_contiguous:
        raise ValueError(f"{name} must be a 1-D C-contiguous complex128 numpy array")
    if size is not None and not 0 <= int(size) <= a.size:
        raise ValueError(f"size {size} outside [0, len({name})={a.size}]")
    if out and not a.flags.writeable:
        raise ValueError(f"{name} is read-only")
    return a


def _dev(torch, a, size):
    return torch.from_numpy(a[:size]).to("cuda")


def _pair(z):
    z = complex(z)
    return (ctypes.c_double * 2)(z.real, z.imag)


def _zvec(op, size, a, b, x, y, w=None, z=None, out=None):
    """Runs one elementwise op on device copies and writes `out` back."""
    torch = nat.require_device()
    size = int(size)
    if size == 0:
        return
    dx = _dev(torch, x, size)
    dy = dx if y is x else _dev(torch, y, size)
    dw = _dev(torch, w, size) if w is not None else None
    dz = _dev(torch, z, size) if z is not None else None
    nat.check(nat.lib().jkpm_zvec(op, size, _pair(a), _pair(b), dx.data_ptr(), dy.data_ptr(),
                                  dw.data_ptr() if dw is not None else None,
                                  dz.data_ptr() if dz is not None else None,
                                  nat.stream_ptr(None)))
    res = dz if out is z else dy
    out[:size] = res.cpu().numpy()


def scal(a, x, size, n_threads=None):
    """x <- a x (in place)."""
    _host("x", x, size, out=True)
    _zvec(nat.JKPM_VOP_AXPBY, size, 0.0, a, x, x, out=x)


def axpy(a, x, y, size, n_threads=None):
    """y <- a x + y (in place)."""
    _host("x", x, size)
    _host("y", y, size, out=True)
    _zvec(nat.JKPM_VOP_AXPBY, size, a, 1.0, x, y, out=y)


def axpy2(a, b, x, y, size, n_threads=None):
    """y <- a x + b y (in place)."""
    _host("x", x, size)
    _host("y", y, size, out=True)
    _zvec(nat.JKPM_VOP_AXPBY, size, a, b, x, y, out=y)


def vec_mul(a, x, y, z, size, n_threads=None):
    """z <- a y x (in place on z)."""
    _host("x", x, size)
    _host("y", y, size)
    _host("z", z, size, out=True)
    _zvec(nat.JKPM_VOP_MUL, size, a, 0.0, x, y, z=z, out=z)


def vec_mul2(a, x, y, size, n_threads=None):
    """y <- a x y (in place on y)."""
    _host("x", x, size)
    _host("y", y, size, out=True)
    _zvec(nat.JKPM_VOP_MUL2, size, a, 0.0, x, y, out=y)


def vec_mul3(a, x, y, w, z, size, n_threads=None):
    """z <- a y x w (in place on z)."""
    for name, arr in (("x", x), ("y", y), ("w", w)):
        _host(name, arr, size)
    _host("z", z, size, out=True)
    _zvec(nat.JKPM_VOP_MUL3, size, a, 0.0, x, y, w=w, z=z, out=z)


def _zdot(a, x, y, size, conj_x):
    torch = nat.require_device()
    size = int(size)
    dx = _dev(torch, x, max(size, 0))
    dy = _dev(torch, y, max(size, 0))
    need = ctypes.c_size_t()
    nat.check(nat.lib().jkpm_zdot_workspace(ctypes.byref(need)))
    ws = torch.empty(int(need.value), dtype=torch.uint8, device="cuda")
    out = torch.empty(1, dtype=torch.complex128, device="cuda")
    nat.check(nat.lib().jkpm_zdot(size, _pair(a), dx.data_ptr(), dy.data_ptr(), conj_x,
                                  out.data_ptr(), ws.data_ptr(), ws.numel(),
                                  nat.stream_ptr(None)))
    return complex(out.cpu().numpy()[0])


def vdot(a, x, y, size, n_threads=None):
    """sum_i y_i a conj(x_i) (FP64, fixed order)."""
    _host("x", x, size)
    _host("y", y, size)
    return _zdot(a, x, y, size, 1)


def dot(a, x, y, size, n_threads=None):
    """sum_i y_i a x_i (FP64, fixed order)."""
    _host("x", x, size)
    _host("y", y, size)
    return _zdot(a, x, y, size, 0)


def aAxby(a, b, x, y, data, indices, len_row, num_columns, n_threads=None):
    """y <- b y + a A x for the row-major ELL (data, indices, len_row) of
    num_columns rows, in place on y (x must not be y)."""
    torch = nat.require_device()
    n, k = int(num_columns), int(len_row)
    _host("x", x, n)
    _host("y", y, n, out=True)
    data = _host("data", data, n * k)
    indices = np.asarray(indices)
    if indices.dtype != np.int32 or indices.ndim != 1 or indices.size < n * k:
        raise ValueError("indices must be a 1-D int32 array of num_columns*len_row entries")
    if x is y or np.shares_memory(x, y):
        raise ValueError("x must not alias y")
    h = nat.DeviceELL(data[:n * k], indices[:n * k], n, k, 1.0, "c128")
    xd = torch.from_numpy(x[:n].reshape(n, 1)).to("cuda")
    yd = torch.from_numpy(y[:n].reshape(n, 1)).to("cuda")
    h.aAxby(complex(a), complex(b), xd, yd)
    y[:n] = yd.cpu().numpy().reshape(n)
