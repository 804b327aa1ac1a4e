"""Kubo double Chebyshev moments and the conductivity they give.

The reference has no Kubo-Bastin moment contraction: its nearest pieces are
the double-expansion operator sum position_operator
(jclsquant/cuda_cython/cuda/obv.cu:562-644, O(M^2) SpMVs per vector) and the
coefficient tensor table_G_delta (jclsquant/cython_modules/Extra_funs.pyx:
479-535, imported but never called).  Here

    mu_mn = (1/R) sum_r <r| v T_m(H~) v T_n(H~) |r>

is computed as one complex GEMM of two stacks of Chebyshev vectors on FP32
MFMA, and sigma(E) is contracted on the device without materialising the
[M, M, E] coefficient table (each energy is a rank-1/rank-2 bilinear form).

H~ = H/B must be Hermitian and v = -i H .* dx (jclsquant/modifiers.py:56-79)
lives on H's slots, as the reference's ell_matrix keeps dx_vec/dy_vec there.
"""
from __future__ import annotations

import numpy as np

from . import _kpm
from . import _native as nat
from .modifiers import modifier_velocity

__all__ = ["kpm_kubo_moments", "kpm_sigma", "velocity_handle"]

KINDS = {"greenwood": 0, "g_delta": 1}


def _scale_of(H):
    s = _kpm.bounds_scale(H.bounds)
    if s is None:
        raise ValueError("the Kubo path needs symmetric bounds [-B, B]")
    return s


def velocity_handle(H, direction="x", v_data=None):
    """Device handles (H~ = H/B, v) in complex64 on one slot pattern."""
    s = _scale_of(H)
    Hd = H.device("c64", s)
    if v_data is None:
        if H.dx_vec is None:
            raise ValueError("H carries no bond vectors: build it with DX/DY "
                             "(e.g. ell_matrix(H, Dx, Dy)) or pass v_data")
        v_data = modifier_velocity(H, direction)
    Vd = nat.DeviceELL.like(Hd, v_data, 1.0, "c64")
    return Hd, Vd


def kpm_kubo_moments(H, M, random_vector, direction="x", v_data=None):
    """mu_mn as a device tensor complex128 [M, M] (mean over the R vectors
    of an [R, N] stack, or one vector [N]).  v_data overrides the velocity
    values (row-major, on H's slots)."""
    Hd, Vd = velocity_handle(H, direction, v_data)
    r = _kpm.as_batch(random_vector, int(H.shape[0]), "c64")
    mu = nat.kubo_mu_nm(Hd, Vd, r, int(M))
    nat.check_finite(mu, "double moments [m, n]", _kpm._NONFINITE_HINT)
    return mu


def kpm_sigma(H, M=None, random_vector=None, direction="x", energies=None, kind="greenwood",
              mu=None):
    """[E, 2] table of [energy, Re sigma(E)].

    kind 'greenwood': sum_mn a_m(E~) a_n(E~) mu_mn with a = Jackson * delta,
    i.e. B^2 Tr[v delta(E-H) v delta(E-H)] estimated from the R vectors (the
    T = 0 Kubo-Greenwood sigma_xx up to the factor pi hbar e^2 / (Omega B^2)).
    kind 'g_delta': the contraction of the reference's table_G_delta with mu
    (bigmas + bigmin summed over m, n) at the same energies.
    """
    if M is None:
        M = int(np.sqrt(H.shape[0]))
    B = H.bounds[1]
    if energies is None:
        energies = np.linspace(H.bounds[0] + 1e-3, H.bounds[1] - 1e-3, 2 * M + 1)
    energies = np.asarray(energies, dtype=np.float64)
    if mu is None:
        if random_vector is None:
            from .observables import random_vector_generator

            random_vector = random_vector_generator(H.shape[0])
        mu = kpm_kubo_moments(H, M, random_vector, direction)
    es = _kpm.rescale_energies(energies, H.bounds)
    sig = nat.kubo_sigma(mu, es, KINDS[kind], float(B))
    out = np.zeros((energies.size, 2))
    out[:, 0] = energies
    out[:, 1] = sig.real.cpu().numpy()
    return out
