"""Chebyshev moment recursions with the reference's names and signatures.

Mirrors jclsquant/recurrence_functions.py:9-111.  The recursion
T_0 = v_r, T_1 = A v_r, T_{m+1} = 2 A T_m - T_{m-1} and the dots
mu_m = <v_l | T_m> run on the GPU as one fused kernel per step
(ELL SpMV + three-term update + FP64 moment partials).

Extensions (keyword-only, defaults give the reference behaviour):
  * a vector argument may be an [R, N] stack: the R recursions share each
    pass over the matrix and the result is [R, M];
  * ``dtype='c64'`` runs the recursion in complex64 (the dots stay FP64).
"""
from __future__ import annotations

import numpy as np

from . import _kpm

__all__ = ["rec_A_tab", "rec_A_tab2v", "rec_A_tab2v_2"]


def rec_A_tab(M, A, v=None, *, dtype="c128"):
    """mu_m = <v|T_m(A)|v>, m < M (jclsquant/recurrence_functions.py:9-40)."""
    if v is None:
        print('You must introduce a vector.')
        return
    mu = _kpm.moments(A, int(M), v, dtype=dtype)[0]
    return _kpm.to_host(mu, _kpm.is_single(v))


def rec_A_tab2v(M, A, v_l=None, v_r=None, *, dtype="c128"):
    """mu_m = <v_l|T_m(A)|v_r> (jclsquant/recurrence_functions.py:44-74)."""
    if v_r is None or v_l is None:
        print('You must introduce a right and left vector.')
        return
    mu = _kpm.moments(A, int(M), v_r, v_l=v_l, dtype=dtype)[0]
    return _kpm.to_host(mu, _kpm.is_single(v_r))


def rec_A_tab2v_2(M, A, v_l=None, v_r=None, *, dtype="c128"):
    """(<v_l|T_m v_r>, <v_r|T_m v_r>) (jclsquant/recurrence_functions.py:76-111)."""
    if v_r is None or v_l is None:
        print('You must introduce a right and left vector.')
        return
    mu = _kpm.moments(A, int(M), v_r, v_l=v_l, v_l2=v_r, dtype=dtype)
    single = _kpm.is_single(v_r)
    return _kpm.to_host(mu[0], single), _kpm.to_host(mu[1], single)
