"""Device-level drivers shared by the reference-named API modules.

Everything here moves data to the device once, runs the HIP kernels through
the C-ABI (``_native``) and brings results back.  Batches of R vectors are
kept row-interleaved on the device ([N, R]) so one pass over the matrix serves
all R recursions.
"""
from __future__ import annotations

import numpy as np

from . import _native as nat


def _is_torch(x) -> bool:
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


def as_batch(v, n: int, dtype) -> "torch.Tensor":
    """Vector(s) -> contiguous device tensor [N, R] of complex dtype.

    Accepts a length-N vector or an [R, N] stack (numpy or torch)."""
    torch = nat.require_device()
    cd = nat.torch_cdtype(nat.vtype_of(dtype))
    t = v if _is_torch(v) else torch.from_numpy(np.ascontiguousarray(v))
    if t.dim() == 1:
        t = t.reshape(1, -1)
    if t.dim() != 2 or t.shape[1] != n:
        raise ValueError(f"vector(s) must have length N={n} (shape [N] or [R, N])")
    return t.to(device="cuda", dtype=cd).transpose(0, 1).contiguous()


def apply_launch(h, vr, doubling: bool, tune):
    """Launch shape of a fresh handle: tune None applies the tuner's table
    entry for this shape when there is one; True tunes it (once per shape
    and machine, then from the table); False keeps the library defaults."""
    if tune is False:
        return
    from . import autotune

    if tune:
        autotune.tuned(h, vr, doubling=doubling)
    else:
        autotune.apply_cached(h, vr, doubling)


def moments(A, M: int, v_r, v_l=None, v_l2=None, dtype="c128", scale: float = 1.0,
            doubling: bool = False, tune=None):
    """mu[l, r, m] = <v_l{l}[r] | T_m(scale*A) | v_r[r]> as a device tensor
    complex128 [n_left, R, M]; A is an ``ell_matrix``.  doubling (v_l = v_r,
    Hermitian A): two moments per Chebyshev step (jkpm_moments_ex).  tune: see
    apply_launch."""
    n = int(A.shape[0])
    h = A.device(dtype, scale)
    vr = as_batch(v_r, n, dtype)
    apply_launch(h, vr, doubling, tune)
    if doubling:
        if v_l is not None or v_l2 is not None:
            raise ValueError("the doubling form needs v_l = v_r")
        mu = nat.moments_ex(h, None, vr, M, True)
        nat.check_finite(mu, "moments [m, r]", _NONFINITE_HINT)
        return mu.transpose(0, 1)[None]
    vl = vr if v_l is None else as_batch(v_l, n, dtype)
    vl2 = None if v_l2 is None else as_batch(v_l2, n, dtype)
    for w in (vl, vl2):
        if w is not None and w.shape != vr.shape:
            raise ValueError("left and right vectors must have the same shape")
    mu = h.moments(vr, M, vl, vl2)  # [n_left, M, R]
    nat.check_finite(mu, "moments [l, m, r]", _NONFINITE_HINT)
    return mu.transpose(1, 2)


_NONFINITE_HINT = ("; the inputs hold NaN/Inf, or the spectral bound is below the largest "
                   "|eigenvalue| so that T_n(H/B) grows without bound (set H.bounds wider)")


def to_host(mu, single: bool):
    """[R, M] device moments -> numpy ([M] when a single vector was given)."""
    out = mu.cpu().numpy()
    return out[0].copy() if single else out.copy()


def is_single(v) -> bool:
    return (v.dim() if _is_torch(v) else np.ndim(v)) == 1


def rescale_energies(E, b) -> np.ndarray:
    """E~ = E/h - c/h with h = (b1-b0)/2, c = (b1+b0)/2
    (jclsquant/observables.py:52)."""
    E = np.asarray(E, dtype=np.float64)
    h = 0.5 * (b[1] - b[0])
    c = 0.5 * (b[1] + b[0])
    return E / h - c / h


def dos_device(mu_mr, e_scaled, kernel="jackson"):
    """DOS on device; mu_mr complex128 [M, R] device tensor (mean over R)."""
    torch = nat.require_device()
    k = nat.JKPM_KERNEL_JACKSON if kernel in ("jackson", True, 1) else nat.JKPM_KERNEL_NONE
    es = torch.as_tensor(np.asarray(e_scaled, dtype=np.float64), device=mu_mr.device)
    if es.numel() and not bool((es.abs() < 1.0).all()):
        raise ValueError("math domain error: rescaled energies must lie inside (-1, 1)")
    return nat.dos(mu_mr.contiguous(), es, k)


def bounds_scale(b) -> float | None:
    """Scale of the symmetric rescale H~ = H / ((b1-b0)/2)
    (jclsquant/modifiers.py:24-26), or None for asymmetric bounds."""
    if abs(b[1]) == abs(b[0]):
        return 1.0 / ((b[1] - b[0]) / 2)
    return None
