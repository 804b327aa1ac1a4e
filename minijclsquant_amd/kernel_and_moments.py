"""Damping kernel and Chebyshev coefficient tables of common functions.

Mirrors jclsquant/kernel_and_moments.py: JacksonKernel (:19-24),
moments_delta (:72-78) and the coefficient tables the time-evolution and
response calculations expand in Chebyshev polynomials of the rescaled energy
x in (-1, 1):

  moments_FD_T0   step theta(Ef - x), Fermi-Dirac at T = 0   (:27-31)
  moments_FD_T    Fermi-Dirac at reduced temperature T_mod    (:87-104)
  moments_FD      a tabulated occupation f(x)                 (:107-115)
  moments_Gmas    retarded Green's function 1/(Ef - x + i0)   (:47-54)
  moments_Gmin    advanced Green's function 1/(Ef - x - i0)   (:56-62)
  moments_Gsum    Gmas + Gmin                                 (:64-69)
  moments_Gmas_2  1/(z - x), z = Ef + i broadening            (:33-41)
  moments_delta_2 -Im(moments_Gmas_2)/pi (Lorentzian delta)   (:43-44)
  moments_U       exp(-i x t) (Jacobi-Anger, Bessel J_n)      (:82-85)
  moments_by_matrix  sum_i c_i matrix[i, :]                   (:11-15)

These are host-side tables (microseconds of work).  The DOS is
reconstructed on the device from the same delta/Jackson formulas
(``dos_kernel`` in csrc/dos.hip); the series tables feed jkpm_series.
"""
from __future__ import annotations

import warnings
from math import pi, sqrt

import numpy as np
import scipy.integrate
import scipy.special

__all__ = ["JacksonKernel", "moments_delta", "moments_FD_T0", "moments_FD_T", "moments_FD",
           "moments_Gmas", "moments_Gmin", "moments_Gsum", "moments_Gmas_2", "moments_delta_2",
           "moments_U", "moments_by_matrix"]


def moments_by_matrix(moments_vec, matrix):
    """sum_i moments_vec[i] * matrix[i, :] (complex) over len(moments_vec) rows."""
    m = np.asarray(matrix)
    out = np.zeros(m.shape[1], dtype=complex)
    for i, c in enumerate(moments_vec):
        out += c * m[i, :]
    return out


def JacksonKernel(N):
    """g_n = [(N-n+1) cos(pi n/(N+1)) + sin(pi n/(N+1)) cot(pi/(N+1))]/(N+1),
    n = 0..N-1, with g_0 = 1 (not halved)."""
    n = np.arange(0, N)
    g = (N - n + 1) * np.cos(np.pi * n / (N + 1)) + np.sin(np.pi * n / (N + 1)) * (
        1 / np.tan(np.pi / (N + 1)))
    return g / (N + 1)


def moments_delta(Ef, N):
    """[1, 2cos(n acos Ef), ...] / (pi sqrt(1 - Ef^2)) for a rescaled scalar
    energy |Ef| < 1 (math.sqrt raises outside, as in the reference)."""
    n = np.arange(1, N)
    n = 2 * np.cos(n * np.arccos(Ef))
    n = np.concatenate((np.array([1]), n))
    return n / (pi * sqrt(1 - Ef ** 2))


def moments_FD_T0(Ef, N):
    """Coefficients of theta(Ef - x): c_0 = 1 - acos(Ef)/pi,
    c_n = -2 sin(n acos Ef)/(n pi)."""
    th = np.arccos(Ef)
    n = np.arange(1, N)
    return np.concatenate((np.array([1 - th / pi]), -2 * np.sin(n * th) / (n * pi)))


def moments_Gmas(Ef, N):
    """Retarded Green's function 1/(Ef - x + i0): c_0 = -i/s,
    c_n = -2i exp(-i n acos Ef)/s, s = sqrt(1 - Ef^2)."""
    n = np.arange(1, N)
    s = sqrt(1 - Ef ** 2)
    return np.concatenate((np.array([-1j / s]), -2j * np.exp(-1j * n * np.arccos(Ef)) / s))


def moments_Gmin(Ef, N):
    """Advanced Green's function 1/(Ef - x - i0): the complex conjugate of
    moments_Gmas."""
    n = np.arange(1, N)
    s = sqrt(1 - Ef ** 2)
    return np.concatenate((np.array([1j / s]), 2j * np.exp(1j * n * np.arccos(Ef)) / s))


def moments_Gsum(Ef, N):
    """moments_Gmas + moments_Gmin: c_0 = 0, c_n = -4 sin(n acos Ef)/s."""
    n = np.arange(1, N)
    return np.concatenate((np.array([0]), -4 * np.sin(n * np.arccos(Ef)) / sqrt(1 - Ef ** 2)))


def moments_Gmas_2(Ef, broadening, N):
    """1/(z - x) with z = Ef + i*broadening: c_0 = -i/w,
    c_m = 2 (z - i w)^m/(i w), w = sqrt(1 - z^2) (principal branch)."""
    z = Ef + 1j * broadening
    w = np.sqrt(1 - z ** 2)
    m = np.arange(1, N)
    return np.concatenate((np.array([-1j / w]), 2 * (z - 1j * w) ** m / (1j * w)))


def moments_delta_2(Ef, broadening, M):
    """Lorentzian-broadened delta: -Im(moments_Gmas_2)/pi."""
    return -np.imag(moments_Gmas_2(Ef, broadening, M)) / np.pi


def moments_U(t, N):
    """exp(-i x t) = J_0(t) + 2 sum_n (-i)^n J_n(t) T_n(x) (Jacobi-Anger);
    t is the time in units of hbar / (half the spectral width)."""
    n = np.arange(1, N)
    return np.concatenate((np.array([scipy.special.jv(0, t)]),
                           2 * (-1j) ** n * scipy.special.jv(n, t)))


def _cheb_projection(f, N, E):
    """c_m = (2 - [m==0]) * simpson(f(E) T_m(E) / (pi sqrt(1 - E^2))) on the
    grid E (scipy's composite Simpson rule, as the reference uses)."""
    out = np.zeros(N)
    w = 1 / (pi * np.sqrt(1 - E ** 2))
    th = np.arccos(E)
    dx = np.abs(E[1] - E[0])
    for m in range(N):
        y = f * np.cos(m * th) * w
        s = scipy.integrate.simpson(y, x=E, dx=dx)
        out[m] = s if m == 0 else 2 * s
    return out


def moments_FD_T(Ef, T_mod, N):
    """Fermi-Dirac occupation 1/(1 + exp((x - Ef)/T_mod)) projected on
    T_0..T_{N-1} by Simpson's rule on 2N points of [-1+1e-3, 1-1e-3]
    (reduced temperature T_mod = k_B T / half width)."""
    E = np.linspace(-1 + 1e-3, 1 - 1e-3, 2 * N)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)  # exp overflow is a clean 0
        f = 1 / (1 + np.exp((E - Ef) / T_mod))
    return _cheb_projection(f, N, E)


def moments_FD(N, FD):
    """A tabulated occupation FD on len(FD) points of [-1+1e-3, 1-1e-3]
    projected on T_0..T_{N-1}; the m = 0 term carries weight 1."""
    FD = np.asarray(FD, dtype=np.float64)
    E = np.linspace(-1 + 1e-3, 1 - 1e-3, len(FD))
    return _cheb_projection(FD, N, E)
