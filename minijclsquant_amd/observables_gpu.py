"""GPU-named KPM entry points of the reference.

Mirrors jclsquant/observables_gpu.py:12-93 (kpm_dos_gpu, kpm_dos_2_gpu) and
the binding functions rec_A_tab_gpu / rec_A_tab_2_gpu of
jclsquant/cuda_cython/obv.pyx:23-49, which fill a caller-owned complex128
array in place through the extern "C" symbols rec_A_tab / rec_A_tab_2.
"""
from __future__ import annotations

import numpy as np

from . import _native as nat
from .observables import kpm_dos, kpm_n_dos_n, random_vector_generator

__all__ = ["kpm_dos_gpu", "kpm_dos_2_gpu", "rec_A_tab_gpu", "rec_A_tab_2_gpu"]


def kpm_dos_gpu(H, M=None, random_vector=None, **kw):
    """Same contract as kpm_dos (jclsquant/observables_gpu.py:12-42)."""
    return kpm_dos(H, M, random_vector, **kw)


def kpm_dos_2_gpu(H, M=None, random_vector_l=None, random_vector_r=None, bounds=True,
                  proyector=None, **kw):
    """Same contract as kpm_n_dos_n (jclsquant/observables_gpu.py:46-93)."""
    return kpm_n_dos_n(H, M, random_vector_l, random_vector_r, bounds, proyector, **kw)


def _check_out(A_vec, M):
    if not (isinstance(A_vec, np.ndarray) and A_vec.dtype == np.complex128
            and A_vec.flags.c_contiguous and A_vec.size >= M):
        raise ValueError("A_vec must be a C-contiguous complex128 array with >= M entries")


def rec_A_tab_gpu(data, indices, len_row, v, A_vec, M):
    """A_vec[:M] = <v|T_m(A)|v> via the extern "C" rec_A_tab symbol."""
    _check_out(A_vec, M)
    A_vec[:M] = nat.rec_A_tab_host(data, indices, int(len_row), v, int(M))


def rec_A_tab_2_gpu(data, indices, len_row, v_l, v_r, A_vec, M):
    """A_vec[:M] = <v_l|T_m(A)|v_r> via the extern "C" rec_A_tab_2 symbol."""
    _check_out(A_vec, M)
    A_vec[:M] = nat.rec_A_tab_host(data, indices, int(len_row), v_l, int(M), v_r=v_r)
