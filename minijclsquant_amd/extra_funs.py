"""Host data-format helpers and coefficient tables of the reference namespace.

The reference's namespace carries these Cython routines of
jclsquant/cython_modules/Extra_funs.pyx through star imports
(ell_matrix.py:9, modifiers.py:10, kernel_and_moments.py:8):

* ``data_indices_c`` / ``data_indices_c_0`` (:81-137, :182-219): CSR ->
  row-major ELL into caller-provided arrays;
* ``diagonal_modifier`` (:352-366): data_new = data + diag on the valid
  diagonal slots;
* ``light_modifier_cython`` (:230-342): the light-mode Peierls factor;
* ``table_G_delta`` (:430-535): the [M, M, 2M] Kubo-Bastin coefficient
  tables.

They keep the reference's names, argument orders and in-place outputs.  The
ELL conversions and the table are one-time host input work, vectorised with
numpy; the light modifier runs on the device (jkpm_ell_peierls).
"""
from __future__ import annotations

import numpy as np

from .ell_matrix import _fill

__all__ = ["data_indices_c", "data_indices_c_0", "diagonal_modifier",
           "light_modifier_cython", "table_G_delta"]


def _out(name, a, n, dtype):
    if not isinstance(a, np.ndarray) or a.dtype != dtype or a.ndim != 1 \
            or not a.flags.c_contiguous or a.size < n:
        raise ValueError(f"{name} must be a 1-D C-contiguous {np.dtype(dtype).name} "
                         f"array of at least {n} entries")
    return a


def data_indices_c_0(len_row, A_shape, A_indptr, A_data, A_indices, indices, data,
                     valid_row_values, n_threads=None):
    """Fills the row-major ELL (indices, data, valid_row_values) of a CSR
    matrix and returns (data, indices, valid_row_values); padding repeats the
    row's last column with value 0, valid 0."""
    tot = int(A_shape) * int(len_row)
    _out("indices", indices, tot, np.int32)
    _out("data", data, tot, np.complex128)
    _out("valid_row_values", valid_row_values, tot, np.int32)
    idx, val, (d,) = _fill(int(len_row), int(A_shape), A_indptr, A_indices, [A_data],
                           [np.complex128])
    indices[:tot], data[:tot], valid_row_values[:tot] = idx, d, val
    return np.asarray(data), np.asarray(indices), np.asarray(valid_row_values)


def data_indices_c(len_row, A_shape, A_indptr, A_data, A_indices, dx_data, dy_data,
                   conjugates_data, indices, data, data_dx, data_dy, data_conjugates,
                   valid_row_values, n_threads=None):
    """data_indices_c_0 plus the bond vectors and conjugate flags on the same
    slots; returns (data, indices, data_dx, data_dy, data_conjugates,
    valid_row_values)."""
    tot = int(A_shape) * int(len_row)
    for name, arr, dt in (("indices", indices, np.int32), ("data", data, np.complex128),
                          ("data_dx", data_dx, np.complex128),
                          ("data_dy", data_dy, np.complex128),
                          ("data_conjugates", data_conjugates, np.int32),
                          ("valid_row_values", valid_row_values, np.int32)):
        _out(name, arr, tot, dt)
    nnz = int(np.asarray(A_indptr)[-1])
    for name, arr in (("dx_data", dx_data), ("dy_data", dy_data),
                      ("conjugates_data", conjugates_data)):
        if np.asarray(arr).size < nnz:
            raise ValueError(f"{name} holds fewer entries than the CSR matrix ({nnz})")
    idx, val, (d, dx, dy, cj) = _fill(
        int(len_row), int(A_shape), A_indptr, A_indices,
        [A_data, np.asarray(dx_data)[:nnz], np.asarray(dy_data)[:nnz],
         np.asarray(conjugates_data)[:nnz]],
        [np.complex128, np.complex128, np.complex128, np.int32])
    indices[:tot], data[:tot], data_dx[:tot], data_dy[:tot] = idx, d, dx, dy
    data_conjugates[:tot], valid_row_values[:tot] = cj, val
    return (np.asarray(data), np.asarray(indices), np.asarray(data_dx), np.asarray(data_dy),
            np.asarray(data_conjugates), np.asarray(valid_row_values))


def diagonal_modifier(data, indices, valid_row_values, data_new, diagonal_array, len_row,
                      len_data, n_threads=None):
    """data_new[s] = diagonal_array[row] + data[s] on the valid diagonal slots
    (row = s // len_row == indices[s]), else data[s], for s < len_data."""
    n = int(len_data)
    _out("data_new", data_new, n, np.complex128)
    data = np.asarray(data)[:n]
    rows = np.arange(n) // int(len_row)
    sel = (np.asarray(indices)[:n] == rows) & (np.asarray(valid_row_values)[:n] == 1)
    out = np.array(data, dtype=np.complex128, copy=True)
    diag = np.asarray(diagonal_array)
    out[sel] = diag[rows[sel]] + data[sel]
    data_new[:n] = out


def light_modifier_cython(a, A0, t, w, Tp, eta, data, dx_vec, dy_vec, data_new, size,
                          n_threads=None):
    """data_new = a * data * exp(i A0 (cos(w t) dx + eta sin(w t) dy) / env)
    on the device, env = cosh((t - 2 Tp)/(0.5673 Tp)) for Tp != 0 else 1;
    eta = 0 selects the linear modes (no dx term).  Complex parameters act as
    complex numbers, as in the reference.  dx_vec/dy_vec must be real-valued
    (the bond vectors)."""
    from ._native import DeviceELL, require_device

    require_device()
    n = int(size)
    _out("data_new", data_new, n, np.complex128)
    A0, t, w, Tp, eta = (complex(v) for v in (A0, t, w, Tp, eta))
    env = 1.0 + 0j if Tp == 0 else np.cosh((t - 2 * Tp) / (0.5673 * Tp))
    ax = 0j if eta == 0 else A0 * np.cos(w * t) / env
    ay = A0 * np.sin(w * t) / env * (1.0 if eta == 0 else eta)
    dx = np.asarray(dx_vec)[:n]
    dy = np.asarray(dy_vec)[:n]
    if np.any(np.imag(dx) != 0) or np.any(np.imag(dy) != 0):
        raise ValueError("dx_vec/dy_vec must be real-valued bond vectors")
    if n == 0:
        return
    d = np.ascontiguousarray(np.asarray(data)[:n], dtype=np.complex128)
    idx = np.zeros(n, dtype=np.int32)  # one row of n slots: the factor is per slot
    src = DeviceELL(d, idx, 1, n, 1.0, "c128")
    bonds = DeviceELL.like(src, np.real(dx) + 1j * np.real(dy), 1.0, "c128")
    dst = DeviceELL.like(src, d, 1.0, "c128")
    dst.peierls(src, bonds, complex(a), ax, ay)
    data_new[:n] = dst.values().cpu().numpy().reshape(n)


def table_G_delta(M, alpha_integral, H_kpm_bounds, n_threads=None):
    """(bigmas, bigmin), each float64 [M, M, 2M], at E = linspace(-alpha,
    alpha, 2M): bigmas[m, n, i] = (G_m g_m / B)(delta_n g_n), bigmin[m, n, i] =
    -(delta_m g_m)(G_n g_n / B) with g = Jackson, G_0 = 0, G_n = -2 sin(n acos
    E)/s, delta_0 = 1/(pi s), delta_n = 2 cos(n acos E)/(pi s), s = sqrt(1 - E^2)
    (a zero divisor, s or pi s, taken as 1e-16).  Built as outer products."""
    from .kernel_and_moments import JacksonKernel

    M = int(M)
    E = np.linspace(-alpha_integral, alpha_integral, 2 * M)
    n = np.arange(M)
    g = JacksonKernel(M)
    s = np.sqrt(1.0 - E * E)
    sg = np.where(s == 0, 1e-16, s)                # guard of G (Extra_funs.pyx:436-438)
    sd = np.where(np.pi * s == 0, 1e-16, np.pi * s)  # guard of delta (:466-468)
    th = np.arccos(E)
    G = -2.0 * np.sin(n[:, None] * th[None, :]) / sg[None, :]
    G[0] = 0.0
    dl = 2.0 * np.cos(n[:, None] * th[None, :]) / sd[None, :]
    dl[0] = 1.0 / sd
    Gg = G * (g / H_kpm_bounds)[:, None]
    dg = dl * g[:, None]
    bigmas = np.ascontiguousarray(Gg[:, None, :] * dg[None, :, :])
    bigmin = np.ascontiguousarray(-dg[:, None, :] * Gg[None, :, :])
    return bigmas, bigmin
