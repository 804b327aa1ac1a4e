"""KPM density of states with the reference's entry points.

Mirrors jclsquant/observables.py:19-120 (random_vector_generator, kpm_dos,
kpm_n_dos_n).  In the reference these are the CPU path (Cython + OpenMP);
here both the CPU-named and the GPU-named entry points run the HIP path:
the recursion and the Jackson-damped DOS reconstruction are device kernels.

Keyword-only extensions (defaults reproduce the reference):
  energies  explicit energy grid instead of linspace(b0+1e-3, b1-1e-3, 2M+1)
  dtype     'c128' (default) or 'c64' recursion (moments stay FP64)
  kernel    'jackson' (default) or None
  doubling  two moments per Chebyshev step (Hermitian H only, kpm_dos)
A random_vector may be an [R, N] stack; the DOS is then that of the mean
moments over the R vectors.
"""
from __future__ import annotations

import numpy as np

from . import _kpm
from .modifiers import modifier_bounds

__all__ = ["random_vector_generator", "kpm_dos", "kpm_n_dos_n", "kpm_moments"]


def random_vector_generator(shape_H, seed=None):
    """Random-phase vector exp(i U[0, 2pi)) from numpy's default_rng(seed)
    (jclsquant/observables.py:19-21)."""
    rng = np.random.default_rng(seed)
    return np.exp(1j * rng.uniform(0, 2 * np.pi, shape_H))


def _energy_grid(H, M, energies):
    if energies is not None:
        return np.asarray(energies, dtype=np.float64)
    return np.linspace(H.bounds[0] + 1e-3, H.bounds[1] - 1e-3, 2 * M + 1)


def _rescaled(H):
    """(matrix, scale) for the spectral rescale of modifier_bounds: symmetric
    bounds fold 1/B into the device upload; asymmetric ones go through the
    host modifier (diagonal shift, then scale)."""
    s = _kpm.bounds_scale(H.bounds)
    if s is None:
        return H.modifier(modifier_bounds), 1.0
    return H, s


def kpm_moments(H, M, random_vector, *, dtype="c128", bounds=True, doubling=False, tune=None):
    """Moments mu_m = <r|T_m(H~)|r> as a device tensor complex128 [R, M].
    tune: None applies a launch shape the tuner already has for this shape,
    True tunes it (once per shape and machine), False keeps the defaults."""
    Hk, s = _rescaled(H) if bounds else (H, 1.0)
    return _kpm.moments(Hk, int(M), random_vector, dtype=dtype, scale=s, doubling=doubling,
                        tune=tune)[0]


def _dos_table(H, Ef_vec, mu_rm, kernel):
    delta_mat = np.zeros((len(Ef_vec), 2))
    delta_mat[:, 0] = Ef_vec
    es = _kpm.rescale_energies(Ef_vec, H.bounds)
    dos = _kpm.dos_device(mu_rm.transpose(0, 1), es, kernel)
    delta_mat[:, 1] = dos.cpu().numpy()
    return delta_mat


def kpm_dos(H, M=None, random_vector=None, *, energies=None, dtype="c128", kernel="jackson",
            doubling=False, tune=None):
    """DOS [E, 2] = [energies, unnormalised DOS] (jclsquant/observables.py:26-54).
    tune: as in kpm_moments."""
    if M is None:
        M = int(np.sqrt(H.shape[0]))
    Ef_vec = _energy_grid(H, M, energies)
    if random_vector is None:
        random_vector = random_vector_generator(H.shape[0])
    mu = kpm_moments(H, M, random_vector, dtype=dtype, doubling=doubling, tune=tune)
    return _dos_table(H, Ef_vec, mu, kernel)


def kpm_n_dos_n(H, M=None, random_vector_l=None, random_vector_r=None, bounds=True,
                proyector=None, *, energies=None, dtype="c128", kernel="jackson", tune=None):
    """DOS from mu_m = <v_l|T_m|v_r>, optionally with v_l projected by an ELL
    matrix first (jclsquant/observables.py:59-120)."""
    if M is None:
        M = int(np.sqrt(H.shape[0]))
    Ef_vec = _energy_grid(H, M, energies)
    if random_vector_l is None and random_vector_r is None:
        random_vector_r = random_vector_generator(H.shape[0])
        random_vector_l = np.copy(random_vector_r)
    if proyector is not None:
        random_vector_l = _project(proyector, random_vector_l)
    Hk, s = _rescaled(H) if bounds else (H, 1.0)
    mu = _kpm.moments(Hk, int(M), random_vector_r, v_l=random_vector_l, dtype=dtype, scale=s,
                      tune=tune)[0]
    return _dos_table(H, Ef_vec, mu, kernel)


def _project(P, v):
    """P v on the device (aAxby with a=1, b=0), for one vector or a stack."""
    from . import _native as nat

    torch = nat.require_device()
    n = int(P.shape[0])
    x = _kpm.as_batch(v, n, "c128")
    y = torch.zeros_like(x)
    P.device("c128", 1.0).aAxby(1.0, 0.0, x, y)
    out = y.transpose(0, 1).cpu().numpy()
    return out[0].copy() if np.ndim(v) == 1 else out.copy()
