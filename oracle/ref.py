"""The reference's own CPU kernels, built here from its Cython source.

TEST INFRASTRUCTURE ONLY: imported by tests/ (to check the C restatement in
kpm_oracle.c against the reference itself) and by bench.py's cpu_baseline
leg (to time the reference's CPU path as itself).  The product never
imports it.

``make -C oracle ref`` compiles jclsquant/cython_modules/blas_funs.pyx
(where it lies in the reference tree) into oracle/_ref/blas_funs*.so.  This
module loads that extension and restates only the Python driver loop of the
reference's CPU moments, rec_A_tab (jclsquant/recurrence_functions.py:9-40,
with A.dot = aAxby as in jclsquant/ell_matrix.py:209-211), around the
reference's compiled aAxby and vdot (blas_funs.pyx:64-91).
"""
from __future__ import annotations

import importlib.machinery
import importlib.util
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_DIR = HERE / "_ref"

_mod = None


def path():
    hits = sorted(REF_DIR.glob("blas_funs*.so"))
    return hits[0] if hits else None


def available() -> bool:
    return path() is not None


def blas_funs():
    """The reference's compiled blas_funs module (not registered in
    sys.modules)."""
    global _mod
    if _mod is None:
        p = path()
        if p is None:
            raise RuntimeError("oracle/_ref is not built: run make -C oracle ref where the "
                               "reference tree is present")
        loader = importlib.machinery.ExtensionFileLoader("blas_funs", str(p))
        spec = importlib.util.spec_from_file_location("blas_funs", str(p), loader=loader)
        mod = importlib.util.module_from_spec(spec)
        loader.exec_module(mod)
        _mod = mod
    return _mod


def rec_A_tab(M, data, indices, len_row, v, n_threads=8):
    """mu_m = <v|T_m(A)|v>, m < M (M >= 2), by the reference's loop over its
    compiled kernels: T_1 = A v, then T_0 <- 2 A T_1 - T_0 and a swap per
    moment, each moment a vdot against v."""
    bf = blas_funs()
    data = np.ascontiguousarray(data, dtype=np.complex128)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.complex128)
    n, K, thr = v.size, int(len_row), int(n_threads)
    out = np.zeros(M, dtype=np.complex128)
    T0 = np.copy(v, order="C")
    T1 = np.zeros(n, dtype=np.complex128)
    bf.aAxby(1 + 0j, 0 + 0j, v, T1, data, indices, K, n, thr)
    out[0] = bf.vdot(1 + 0j, v, v, n, thr)
    out[1] = bf.vdot(1 + 0j, v, T1, n, thr)
    for m in range(2, M):
        bf.aAxby(2.0 + 0j, -1 + 0j, T1, T0, data, indices, K, n, thr)
        T0, T1 = T1, T0
        out[m] = bf.vdot(1 + 0j, v, T1, n, thr)
    return out
