"""Damping kernel and delta-function Chebyshev coefficients.

Mirrors jclsquant/kernel_and_moments.py:19-24 (JacksonKernel) and :72-78
(moments_delta).  These are the host-side forms of the API; the DOS of the
observables is reconstructed on the device from the same formulas
(``dos_kernel`` in csrc/dos.hip).
"""
from __future__ import annotations

from math import pi, sqrt

import numpy as np

__all__ = ["JacksonKernel", "moments_delta"]


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
