"""Python face of the CPU oracle (libkpm_oracle.so built from kpm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results or as the timed CPU baseline.  The
product (minijclsquant_amd) never imports it.

Every function restates the reference algorithm in complex128; see the
reference citations in kpm_oracle.c.  How each one is pinned (closed forms,
eigendecomposition, library routines, brute force) is in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libkpm_oracle.so"

_c = ctypes
_vp = _c.c_void_p
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = _c.CDLL(str(LIB))
        i, d, l = _c.c_int, _c.c_double, _c.c_long
        sig = {
            "orc_csr_to_ell": (None, [i, i, _vp, _vp, _vp, _vp, _vp, _vp, i]),
            "orc_scal": (None, [_vp, _vp, l, i]),
            "orc_aAxby": (None, [_vp, _vp, _vp, _vp, _vp, _vp, i, i, i]),
            "orc_vdot": (None, [_vp, _vp, _vp, l, i, _vp]),
            "orc_dot": (None, [_vp, _vp, _vp, l, i, _vp]),
            "orc_vec_mul": (None, [_vp, _vp, _vp, _vp, l, i]),
            "orc_diagonal_modifier": (None, [_vp, _vp, _vp, _vp, _vp, i, l, i]),
            "orc_rec_A_tab": (i, [i, _vp, _vp, i, i, _vp, _vp, i]),
            "orc_rec_A_tab2v": (i, [i, _vp, _vp, i, i, _vp, _vp, _vp, i]),
            "orc_rec_A_tab2v_2": (i, [i, _vp, _vp, i, i, _vp, _vp, _vp, _vp, i]),
            "orc_jackson": (None, [i, _vp]),
            "orc_moments_delta": (None, [d, i, _vp]),
            "orc_dos": (None, [_vp, i, _vp, i, d, d, i, _vp]),
            "orc_mu_nm": (i, [i, _vp, _vp, i, _vp, _vp, i, i, _vp, _vp, i]),
            "orc_grad_values": (i, [i, _vp, _vp, i, i, _vp, _vp, _vp, i]),
            "orc_table_G_delta": (None, [i, _vp, i, d, _vp, _vp]),
            "orc_rec_A_vec": (i, [i, _vp, _vp, i, i, _vp, _vp, _vp, i]),
            "orc_rec_com_A_vec": (i, [i, _vp, _vp, _vp, i, i, _vp, _vp, _vp, _vp, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def threads(n=None) -> int:
    if n is not None:
        return int(n)
    return int(os.environ.get("OMP_NUM_THREADS", min(os.cpu_count() or 1, 16)))


def _z(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return a.ctypes.data


def _cz(x):
    x = complex(x)
    return np.array([x.real, x.imag])


def csr_to_ell(csr, len_row=None, n_threads=None):
    """(data, indices, valid) row-major ELL of a CSR matrix."""
    n = csr.shape[0]
    counts = np.diff(csr.indptr)
    K = int(counts.max()) if len_row is None else int(len_row)
    data = np.zeros(n * K, np.complex128)
    idx = np.zeros(n * K, np.int32)
    val = np.zeros(n * K, np.int32)
    ip, ci, dd = _i(csr.indptr), _i(csr.indices), _z(csr.data)
    lib().orc_csr_to_ell(K, n, _p(ip), _p(dd), _p(ci), _p(idx), _p(data), _p(val),
                         threads(n_threads))
    return data, idx, val, K


def aAxby(a, b, x, y, data, indices, len_row, n_threads=None):
    """In place on y (complex128, C-contiguous)."""
    x, data, indices = _z(x), _z(data), _i(indices)
    assert y.dtype == np.complex128 and y.flags.c_contiguous
    av, bv = _cz(a), _cz(b)
    lib().orc_aAxby(_p(av), _p(bv), _p(x), _p(y), _p(data), _p(indices), int(len_row),
                    y.shape[0], threads(n_threads))


def vdot(a, x, y, n_threads=None):
    x, y = _z(x), _z(y)
    out = np.zeros(2)
    av = _cz(a)
    lib().orc_vdot(_p(av), _p(x), _p(y), x.size, threads(n_threads), _p(out))
    return complex(out[0], out[1])


def dot(a, x, y, n_threads=None):
    x, y = _z(x), _z(y)
    out = np.zeros(2)
    av = _cz(a)
    lib().orc_dot(_p(av), _p(x), _p(y), x.size, threads(n_threads), _p(out))
    return complex(out[0], out[1])


def scal(a, x, n_threads=None):
    assert x.dtype == np.complex128 and x.flags.c_contiguous
    av = _cz(a)
    lib().orc_scal(_p(av), _p(x), x.size, threads(n_threads))


def vec_mul(a, x, y, n_threads=None):
    x, y = _z(x), _z(y)
    z = np.zeros_like(x)
    av = _cz(a)
    lib().orc_vec_mul(_p(av), _p(x), _p(y), _p(z), x.size, threads(n_threads))
    return z


def diagonal_modifier(data, indices, valid, diag, len_row, n_threads=None):
    data, indices, valid, diag = _z(data), _i(indices), _i(valid), _z(diag)
    out = np.zeros_like(data)
    lib().orc_diagonal_modifier(_p(data), _p(indices), _p(valid), _p(out), _p(diag),
                                int(len_row), data.size, threads(n_threads))
    return out


def rec_A_tab(M, data, indices, len_row, v, n_threads=None):
    data, indices, v = _z(data), _i(indices), _z(v)
    mu = np.zeros(M, np.complex128)
    rc = lib().orc_rec_A_tab(int(M), _p(data), _p(indices), int(len_row), v.size, _p(v),
                             _p(mu), threads(n_threads))
    assert rc == 0
    return mu


def rec_A_tab2v(M, data, indices, len_row, v_l, v_r, n_threads=None):
    data, indices, v_l, v_r = _z(data), _i(indices), _z(v_l), _z(v_r)
    mu = np.zeros(M, np.complex128)
    rc = lib().orc_rec_A_tab2v(int(M), _p(data), _p(indices), int(len_row), v_r.size, _p(v_l),
                               _p(v_r), _p(mu), threads(n_threads))
    assert rc == 0
    return mu


def rec_A_tab2v_2(M, data, indices, len_row, v_l, v_r, n_threads=None):
    data, indices, v_l, v_r = _z(data), _i(indices), _z(v_l), _z(v_r)
    mu1 = np.zeros(M, np.complex128)
    mu2 = np.zeros(M, np.complex128)
    rc = lib().orc_rec_A_tab2v_2(int(M), _p(data), _p(indices), int(len_row), v_r.size,
                                 _p(v_l), _p(v_r), _p(mu1), _p(mu2), threads(n_threads))
    assert rc == 0
    return mu1, mu2


def jackson(M):
    g = np.zeros(M)
    lib().orc_jackson(int(M), _p(g))
    return g


def moments_delta(e, M):
    out = np.zeros(M)
    lib().orc_moments_delta(float(e), int(M), _p(out))
    return out


def dos(mu, energies, b0, b1, kernel=True):
    mu, E = _z(mu), np.ascontiguousarray(energies, np.float64)
    out = np.zeros(E.size)
    lib().orc_dos(_p(mu), mu.size, _p(E), E.size, float(b0), float(b1), 1 if kernel else 0,
                  _p(out))
    return out


def mu_nm(M, h_data, h_idx, h_k, v_data, v_idx, v_k, r, n_threads=None):
    """[M, M] mu_mn = <r| v T_m v T_n |r> (Hermitian H and v)."""
    h_data, h_idx, v_data, v_idx, r = _z(h_data), _i(h_idx), _z(v_data), _i(v_idx), _z(r)
    out = np.zeros((M, M), np.complex128)
    rc = lib().orc_mu_nm(int(M), _p(h_data), _p(h_idx), int(h_k), _p(v_data), _p(v_idx),
                         int(v_k), r.size, _p(r), _p(out), threads(n_threads))
    assert rc == 0
    return out


def rec_A_vec(M, data, indices, len_row, c, v, n_threads=None):
    """sum_{m<M} c_m T_m(A) v (complex128)."""
    data, indices, v = _z(data), _i(indices), _z(v)
    cc = _z(np.asarray(c, dtype=np.complex128))
    assert cc.size >= M
    out = np.zeros(v.size, np.complex128)
    rc = lib().orc_rec_A_vec(int(M), _p(data), _p(indices), int(len_row), v.size, _p(cc), _p(v),
                             _p(out), threads(n_threads))
    assert rc == 0
    return out


def rec_com_A_vec(M, a_data, v_data, indices, len_row, c, v, table=False, n_threads=None):
    """sum_m c_m C_m (or the table [M, N] of c_m C_m) of the commutator
    recursion C_m = 2 A C_{m-1} - C_{m-2} + 2 V T_{m-1}, C_0 = 0, C_1 = V v."""
    a_data, v_data, indices, v = _z(a_data), _z(v_data), _i(indices), _z(v)
    cc = _z(np.asarray(c, dtype=np.complex128))
    assert cc.size >= M
    n = v.size
    out = None if table else np.zeros(n, np.complex128)
    tab = np.zeros((M, n), np.complex128) if table else None
    rc = lib().orc_rec_com_A_vec(int(M), _p(a_data), _p(v_data), _p(indices), int(len_row), n,
                                 _p(cc), _p(v), None if out is None else _p(out),
                                 None if tab is None else _p(tab), threads(n_threads))
    assert rc == 0
    return tab if table else out


def table_G_delta(M, E, bound):
    """(bigmas, bigmin), each float64 [M, M, len(E)], at rescaled energies E."""
    E = np.ascontiguousarray(E, np.float64)
    bm = np.zeros((M, M, E.size))
    bn = np.zeros((M, M, E.size))
    lib().orc_table_G_delta(int(M), _p(E), E.size, float(bound), _p(bm), _p(bn))
    return bm, bn


def kubo_sigma(mu, E, bound, kind="greenwood"):
    """sigma(E_i) from double moments mu [M, M] (complex).  'greenwood':
    sum_mn a_m a_n mu_mn with a = Jackson * moments_delta; 'g_delta':
    sum_mn (bigmas + bigmin)[m, n, i] mu_mn with the materialised table."""
    mu = np.asarray(mu, np.complex128)
    M = mu.shape[0]
    E = np.asarray(E, np.float64)
    if kind == "greenwood":
        g = jackson(M)
        out = np.zeros(E.size, np.complex128)
        for i, e in enumerate(E):
            a = moments_delta(e, M) * g
            out[i] = a @ mu @ a
        return out
    bm, bn = table_G_delta(M, E, bound)
    return np.einsum("mni,mn->i", bm + bn, mu)


def grad_values(M, data, indices, len_row, r, c, n_threads=None):
    """G = dJ/dRe(a) + i dJ/dIm(a) for J = Re sum_n c_n <r|T_n|r>, per ELL slot."""
    data, indices, r = _z(data), _i(indices), _z(r)
    cc = _z(np.asarray(c, dtype=np.complex128))
    assert cc.size >= M
    out = np.zeros(data.size, np.complex128)
    rc = lib().orc_grad_values(int(M), _p(data), _p(indices), int(len_row), r.size, _p(r),
                               _p(cc), _p(out), threads(n_threads))
    assert rc == 0
    return out


def peierls(data, dx, dy, a, ax, ay):
    """a * data * exp(i (ax dx + ay dy)) slot by slot: the Peierls factor of
    the reference's light modifiers (jclsquant/modifiers.py:133-181 over
    vec_mul; GPU form jclsquant/cuda_cython/cuda/modifiers.cu:25-68 with
    alpha = i ax, beta = i ay)."""
    dx = np.real(np.asarray(dx))
    dy = np.real(np.asarray(dy))
    return a * np.asarray(data, dtype=np.complex128) * np.exp(1j * (ax * dx + ay * dy))


def rho_neq(data, indices, len_row, dx, dy, t_vec, v, c_fd, c_u, phases, n_threads=None):
    """(U, U_F) of the reference's time loop (jclsquant/cuda_cython/cuda/
    obv.cu:262-361): U_F = sum_m c_fd[m] T_m(H(t_0)) v, then for every later
    t_i: U_F <- sum_m c_u[m] T_m(H(t_i)) U_F and U <- the same of U, with
    H(t) = peierls(data, dx, dy, 1, *phases(t)) (data already rescaled)."""
    d0 = peierls(data, dx, dy, 1.0, *phases(t_vec[0]))
    uf = rec_A_vec(len(c_fd), d0, indices, len_row, c_fd, v, n_threads)
    u = np.array(v, dtype=np.complex128)
    for t in t_vec[1:]:
        d = peierls(data, dx, dy, 1.0, *phases(t))
        uf = rec_A_vec(len(c_u), d, indices, len_row, c_u, uf, n_threads)
        u = rec_A_vec(len(c_u), d, indices, len_row, c_u, u, n_threads)
    return u, uf
