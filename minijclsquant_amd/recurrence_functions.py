"""Chebyshev moment recursions with the reference's names and signatures.

Mirrors jclsquant/recurrence_functions.py:9-111.  The recursion
T_0 = v_r, T_1 = A v_r, T_{m+1} = 2 A T_m - T_{m-1} and the dots
mu_m = <v_l | T_m> run on the GPU as one fused kernel per step
(ELL SpMV + three-term update + FP64 moment partials).

Extensions (keyword-only, defaults give the reference behaviour):
  * a vector argument may be an [R, N] stack: the R recursions share each
    pass over the matrix and the result is [R, M];
  * ``dtype='c64'`` runs the recursion in complex64 (the dots stay FP64);
  * ``doubling=True`` (rec_A_tab, Hermitian A only) produces the M moments
    from M/2 steps via mu_2n = 2<T_n|T_n> - mu_0, mu_2n+1 = 2<T_n|T_n+1> - mu_1.
"""
from __future__ import annotations

import numpy as np

from . import _kpm

__all__ = ["rec_A_tab", "rec_A_tab2v", "rec_A_tab2v_2", "rec_A_vec", "rec_com_A_vec",
           "rec_com_A_vec_tab"]


def rec_A_tab(M, A, v=None, *, dtype="c128", doubling=False, tune=None):
    """mu_m = <v|T_m(A)|v>, m < M (jclsquant/recurrence_functions.py:9-40)."""
    if v is None:
        print('You must introduce a vector.')
        return
    mu = _kpm.moments(A, int(M), v, dtype=dtype, doubling=doubling, tune=tune)[0]
    return _kpm.to_host(mu, _kpm.is_single(v))


def rec_A_tab2v(M, A, v_l=None, v_r=None, *, dtype="c128", tune=None):
    """mu_m = <v_l|T_m(A)|v_r> (jclsquant/recurrence_functions.py:44-74)."""
    if v_r is None or v_l is None:
        print('You must introduce a right and left vector.')
        return
    mu = _kpm.moments(A, int(M), v_r, v_l=v_l, dtype=dtype, tune=tune)[0]
    return _kpm.to_host(mu, _kpm.is_single(v_r))


def rec_A_tab2v_2(M, A, v_l=None, v_r=None, *, dtype="c128", tune=None):
    """(<v_l|T_m v_r>, <v_r|T_m v_r>) (jclsquant/recurrence_functions.py:76-111)."""
    if v_r is None or v_l is None:
        print('You must introduce a right and left vector.')
        return
    mu = _kpm.moments(A, int(M), v_r, v_l=v_l, v_l2=v_r, dtype=dtype, tune=tune)
    single = _kpm.is_single(v_r)
    return _kpm.to_host(mu[0], single), _kpm.to_host(mu[1], single)


def _series(M, A, V, moments_kernel, v, dtype, table):
    import numpy as np

    from . import _native as nat

    n = int(A.shape[0])
    # The reference always forms the m = 0 and m = 1 terms (its loop then
    # runs m = 2 .. M-1), so M < 2 behaves as M = 2 for the sums and the
    # table needs two rows.
    if table and M < 2:
        raise ValueError("the commutator table needs M >= 2")
    M = max(int(M), 2)
    h = A.device(dtype, 1.0)
    vop = None
    if V is not None:
        if not (np.array_equal(np.asarray(V.indices), np.asarray(A.indices))
                and int(V.len_row) == int(A.len_row)):
            raise ValueError("V must have A's ELL slot pattern (same indices and len_row)")
        vop = nat.DeviceELL.like(h, V.data, 1.0, dtype)
    x = _kpm.as_batch(v, n, dtype)
    out, tab = nat.series(h, vop, x, int(M), moments_kernel, want_sum=not table,
                          want_table=table)
    single = _kpm.is_single(v)
    if table:
        t = tab.permute(0, 2, 1).cpu().numpy()  # [M, R, N]
        return t[:, 0, :].copy() if single else t
    o = out.transpose(0, 1).cpu().numpy()  # [R, N]
    return o[0].copy() if single else o


def rec_A_vec(M, A, moments_kernel=None, v=None, *, dtype="c128"):
    """sum_{m<M} moments_kernel[m] T_m(A) v (jclsquant/recurrence_functions.py:114-143)."""
    if moments_kernel is None:
        print('You must introduce the analytical_moments*Kernel which is an array of size M')
        return
    if v is None:
        print('You must introduce a vector.')
        return
    return _series(M, A, None, moments_kernel, v, dtype, False)


def rec_com_A_vec(M, A, V=None, moments_kernel=None, v=None, *, dtype="c128"):
    """sum_m moments_kernel[m] C_m of the commutator recursion
    C_m = 2 A C_{m-1} - C_{m-2} + 2 V T_{m-1}, C_0 = 0, C_1 = V v
    (jclsquant/recurrence_functions.py:146-196).  V shares A's slots."""
    if V is None:
        print('You must introduce kpm renormalized velocity .')
        return
    if moments_kernel is None:
        print('You must introduce the analytical_moments*Kernel which is an array of size M .')
        return
    if v is None:
        print('You must introduce a vector.')
        return
    return _series(M, A, V, moments_kernel, v, dtype, False)


def rec_com_A_vec_tab(M, A, V=None, moments_kernel=None, v=None, *, dtype="c128"):
    """Table [M, N] of moments_kernel[m] C_m (jclsquant/recurrence_functions.py:202-250)."""
    if V is None:
        print('You must introduce kpm renormalized velocity .')
        return
    if v is None:
        print('You must introduce a vector.')
        return
    return _series(M, A, V, moments_kernel, v, dtype, True)
