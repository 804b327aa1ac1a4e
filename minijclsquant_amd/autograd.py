"""Differentiable KPM: torch.autograd.Functions over the HIP kernels.

``chebyshev_moments(values, indices, r, M, scale)`` returns the moments
mu[r, m] = <r|T_m(scale * A)|r> of the ELL matrix whose row-major complex128
values are ``values`` (a device tensor that may require grad).  Its backward
is the hand-written adjoint Chebyshev recursion of the HIP library
(jkpm_moments_grad: sqrt(M) checkpoints + recomputation), not torch autograd
through a loop.  ``kpm_dos_torch`` maps moments to the Jackson DOS with the
device DOS kernel and its transpose.

``lattice_values`` builds ELL values from real onsite energies and per-bond
hoppings by a gather, so gradients reach those parameters through torch.

Gradients follow torch's convention for complex tensors:
dL/dRe(z) + i dL/dIm(z).  The matrix must be Hermitian at the point of
evaluation (the adjoint uses A^H = A; a non-Hermitian one raises
ValueError); the spectral bound is a constant.
The reference computes no gradients; this extends its rec_A_tab
(jclsquant/recurrence_functions.py:9-40) and kpm_dos
(jclsquant/observables.py:26-54).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat

__all__ = ["chebyshev_moments", "kpm_dos_torch", "lattice_values", "param_slots"]


class _Moments(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, indices, r, M, scale, seg, keep, tune):
        N = r.shape[0]
        K = values.numel() // N
        h = nat.DeviceELL(values.detach().contiguous(), indices, N, K, scale, "c128")
        if values.requires_grad:
            # the adjoint runs A in place of A^H: refuse before the forward pass
            err, mag = h.hermiticity()
            if not err <= nat.HERMITIAN_RTOL * mag:
                raise ValueError(
                    f"chebyshev_moments: the gradient needs a Hermitian matrix, but "
                    f"max |A_ij - conj(A_ji)| = {err:.3e} exceeds "
                    f"{nat.HERMITIAN_RTOL:.0e} * max |A_ij| = {nat.HERMITIAN_RTOL * mag:.3e}")
        if tune:
            # launch shape from the persistent table (timed once per shape)
            from . import autotune

            autotune.tuned(h, r, passes=not keep)
        ctx.stack = None
        if keep:
            # the forward keeps every T_n: the backward runs no forward pass
            mu, ctx.stack = h.moments_keep(r, M)
        else:
            mu = h.moments(r, M)[0]  # [M, R]
        ctx.h = h
        ctx.M = int(M)
        ctx.seg = int(seg)
        ctx.save_for_backward(r)
        return mu.transpose(0, 1).contiguous()

    @staticmethod
    def backward(ctx, grad):
        (r,) = ctx.saved_tensors
        g = grad.transpose(0, 1).contiguous()
        if ctx.stack is not None:
            gv = ctx.h.grad_values_kept(r, ctx.M, g, ctx.stack)
            ctx.stack = None
            return gv, None, None, None, None, None, None, None
        seg = ctx.seg
        if seg == _AUTO:
            # keep every T_n when that fits in half the free device memory
            R = r.shape[1]
            need = ctx.h.grad_workspace_bytes(R, ctx.M, -1)
            free, _ = torch.cuda.mem_get_info(r.device)
            seg = -1 if need <= free // 2 else 0
        gv = ctx.h.grad_values(r, ctx.M, g, seg)
        return gv, None, None, None, None, None, None, None


_AUTO = -(2**31)


def chebyshev_moments(values, indices, r, M, scale=1.0, seg=None, tune=True):
    """mu [R, M] complex128 of the ELL (values, indices) times scale.

    values: device complex128 [N*K] (row-major ELL slots); indices: device
    int32 [N*K]; r: device complex128 [N, R] (row-interleaved vectors).
    seg (backward): checkpoints every seg steps with recomputation (0:
    sqrt(M-1)), every T_n kept by the backward's own forward pass (< 0), or
    None: when a gradient will be taken and max(M,2) vector batches fit in
    half the free device memory, the forward itself keeps every T_n and the
    backward runs no forward pass; otherwise sqrt(M-1) checkpoints.
    tune: run the forward with the autotuner's launch shape for this
    (N, K, R) (timed on the first call, then read from its table)."""
    if r.dtype != torch.complex128:
        raise ValueError("the differentiable path runs in complex128")
    keep = False
    if seg is None and torch.is_grad_enabled() and values.requires_grad:
        need = max(int(M), 2) * r.numel() * r.element_size()
        free, _ = torch.cuda.mem_get_info(r.device)
        keep = need <= free // 2
    return _Moments.apply(values, indices, r, int(M), float(scale),
                          _AUTO if seg is None else int(seg), keep, bool(tune))


class _Dos(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mu_rm, e_scaled, kernel):
        mu_mr = mu_rm.transpose(0, 1).contiguous()
        ctx.save_for_backward(e_scaled)
        ctx.shape = mu_mr.shape
        ctx.kernel = kernel
        return nat.dos(mu_mr, e_scaled, kernel)

    @staticmethod
    def backward(ctx, gdos):
        (e,) = ctx.saved_tensors
        M, R = ctx.shape
        g = nat.dos_grad(gdos, M, R, e, ctx.kernel)
        return g.transpose(0, 1).contiguous(), None, None


def kpm_dos_torch(mu, e_scaled, kernel="jackson"):
    """DOS [E] (float64) of the mean moments of mu [R, M] at rescaled
    energies (device float64), differentiable in mu."""
    k = nat.JKPM_KERNEL_JACKSON if kernel in ("jackson", True, 1) else nat.JKPM_KERNEL_NONE
    e = torch.as_tensor(e_scaled, dtype=torch.float64, device=mu.device)
    return _Dos.apply(mu, e, k)


def param_slots(A, n_bonds_of=None):
    """Map every ELL slot of a Hermitian lattice matrix to a real parameter:
    slot (i, i) -> onsite i (index i); slots (i, j) and (j, i) of one bond ->
    the same bond index N + b, bonds numbered in the order of the upper
    triangle (i < j) of the ELL; padding slots -> -1.  Returns
    (slot_param int64 [N*K], bond_pairs int64 [n_bonds, 2])."""
    N, K = int(A.shape[0]), int(A.len_row)
    rows = np.arange(N * K) // K
    cols = np.asarray(A.indices, dtype=np.int64)
    valid = np.asarray(A.valid_row_values) == 1
    out = np.full(N * K, -1, dtype=np.int64)
    diag = valid & (rows == cols)
    out[diag] = rows[diag]
    upper = valid & (rows < cols)
    pairs = np.stack([rows[upper], cols[upper]], 1)
    bond_id = np.arange(pairs.shape[0])
    out[upper] = N + bond_id
    lower = valid & (rows > cols)
    # the (j, i) slot of bond (i, j): look the pair up by key i*N + j
    key_upper = pairs[:, 0] * N + pairs[:, 1]
    order = np.argsort(key_upper)
    key_lower = cols[lower] * N + rows[lower]
    pos = np.searchsorted(key_upper[order], key_lower)
    if np.any(pos >= key_upper.size) or np.any(key_upper[order][np.minimum(pos, key_upper.size - 1)] != key_lower):
        raise ValueError("the slot pattern is not symmetric")
    out[lower] = N + bond_id[order][pos]
    return out, pairs


def lattice_values(onsite, hoppings, slot_param):
    """Row-major complex128 ELL values from real onsite energies [N] and bond
    hoppings [n_bonds] (torch tensors, may require grad): a gather, so the
    gradient of a slot flows to its parameter (both slots of a bond add)."""
    p = torch.cat([onsite, hoppings]).to(torch.complex128)
    idx = torch.as_tensor(slot_param, device=p.device)
    safe = idx.clamp(min=0)
    vals = p[safe]
    return torch.where(idx >= 0, vals, torch.zeros_like(vals))
