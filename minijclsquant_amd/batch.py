"""Many disorder realizations (of uneven sizes) in one batch.

The reference computes one realization and one vector at a time (callers
loop over ``kpm_dos``, jclsquant/observables.py:26-54).  Here G
realizations are stacked into one block-diagonal ELL (rows of realization g
at [offsets[g], offsets[g+1]), its columns shifted by offsets[g], its values
pre-scaled by its own 1/B_g), so every Chebyshev step of all G realizations
and all R vectors is one kernel launch; moments come out per realization.
The whole M-step recursion can be captured into one HIP graph.

``square_lattice_ell`` builds the row-major ELL of a periodic square lattice
directly (no CSR), with the same slots, order and values as
``ell_matrix(square_lattice(...)[0])``.
"""
from __future__ import annotations

import numpy as np

from . import _kpm
from . import _native as nat

__all__ = ["square_lattice_ell", "block_diagonal", "stack_csr", "lattice_stride", "BatchKPM",
           "kpm_dos_batch"]


def square_lattice_ell(L, t=-1.0, W=0.0, seed=None):
    """(data c128 [N*5], indices int32 [N*5]) of the periodic L x L square
    lattice with onsite U[-W/2, W/2] from default_rng(seed); columns sorted
    ascending per row, as in the CSR -> ELL path.  L >= 3."""
    L = int(L)
    if L < 3:
        raise ValueError("L >= 3 (smaller periodic lattices have coinciding bonds)")
    N = L * L
    rng = np.random.default_rng(seed)
    eps = (-W / 2 + W * rng.random(N)) if W else np.zeros(N)
    i = np.arange(N, dtype=np.int64)
    x, y = i % L, i // L
    cols = np.stack([((y - 1) % L) * L + x, y * L + (x - 1) % L, i,
                     y * L + (x + 1) % L, ((y + 1) % L) * L + x], 1)
    vals = np.full((N, 5), complex(t))
    vals[:, 2] = eps
    order = np.argsort(cols, axis=1, kind="stable")
    cols = np.take_along_axis(cols, order, 1)
    vals = np.take_along_axis(vals, order, 1)
    return vals.reshape(-1), cols.reshape(-1).astype(np.int32)


def block_diagonal(blocks):
    """Stack (data, indices, n_rows, len_row, scale) blocks into one
    row-major ELL of width K = max len_row; returns (data, indices, offsets, K).
    Narrower rows are padded like the reference (last column repeated, 0)."""
    K = max(int(b[3]) for b in blocks)
    offsets = np.zeros(len(blocks) + 1, dtype=np.int64)
    for g, b in enumerate(blocks):
        offsets[g + 1] = offsets[g] + int(b[2])
    data = np.zeros(int(offsets[-1]) * K, dtype=np.complex128)
    idx = np.zeros(int(offsets[-1]) * K, dtype=np.int32)
    for g, (d, ix, n, k, s) in enumerate(blocks):
        n, k = int(n), int(k)
        d2 = np.asarray(d).reshape(n, k) * s
        i2 = np.asarray(ix, dtype=np.int64).reshape(n, k) + offsets[g]
        if k < K:
            d2 = np.concatenate([d2, np.zeros((n, K - k), d2.dtype)], 1)
            i2 = np.concatenate([i2, np.repeat(i2[:, -1:], K - k, 1)], 1)
        o0, o1 = int(offsets[g]) * K, int(offsets[g + 1]) * K
        data[o0:o1] = d2.reshape(-1)
        idx[o0:o1] = i2.reshape(-1)
    return data, idx, offsets, K


def lattice_stride(rows, cols, n_rows, lo=64):
    """The row stride of a lattice-like matrix: the most common column
    distance |j - i| in (lo, n_rows/2] over its entries (rows[e], cols[e]),
    e.g. L for a row-major L x L square lattice; 0 when no such distance
    holds a quarter of the rows (no dominant long-range neighbour)."""
    d = np.abs(np.asarray(cols, dtype=np.int64) - np.asarray(rows, dtype=np.int64))
    d = d[(d > lo) & (d <= n_rows // 2)]
    if d.size < max(1, n_rows // 4):
        return 0
    counts = np.bincount(d)
    s = int(np.argmax(counts))
    return s if counts[s] >= max(1, n_rows // 4) else 0


def _ell_stride(indices, n, k):
    n, k = int(n), int(k)
    return lattice_stride(np.repeat(np.arange(n, dtype=np.int64), k), indices, n)


def _csr_stride(m):
    n = int(m.shape[0])
    return lattice_stride(np.repeat(np.arange(n, dtype=np.int64), np.diff(m.indptr)),
                          m.indices[:int(m.indptr[-1])], n)


def stack_csr(mats, scales):
    """Block-diagonal CSR of square CSR blocks, block g's values times
    scales[g]: (indptr int32, indices int32, data complex128, row offsets)."""
    sizes = [int(m.shape[0]) for m in mats]
    offsets = np.zeros(len(mats) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(sizes)
    nnz = np.zeros(len(mats) + 1, dtype=np.int64)
    nnz[1:] = np.cumsum([int(m.indptr[-1]) for m in mats])
    if offsets[-1] >= 2**31 or nnz[-1] >= 2**31:
        raise ValueError("the stacked CSR exceeds 2^31-1 rows or entries")
    indptr = np.empty(int(offsets[-1]) + 1, dtype=np.int32)
    indices = np.empty(int(nnz[-1]), dtype=np.int32)
    data = np.empty(int(nnz[-1]), dtype=np.complex128)
    indptr[0] = 0
    for g, (m, s) in enumerate(zip(mats, scales)):
        r0, r1, e0, e1 = int(offsets[g]), int(offsets[g + 1]), int(nnz[g]), int(nnz[g + 1])
        indptr[r0 + 1:r1 + 1] = np.asarray(m.indptr[1:], dtype=np.int64) + e0
        indices[e0:e1] = np.asarray(m.indices[:e1 - e0], dtype=np.int64) + r0
        np.multiply(np.asarray(m.data[:e1 - e0], dtype=np.complex128), s, out=data[e0:e1])
    return indptr, indices, data, offsets


class BatchKPM:
    """Device state for the moments of G realizations x R vectors.

    blocks: list of (data, indices, n_rows, len_row, scale) -- row-major ELL
    of each realization and the scale 1/B_g of its spectral rescale.
    vectors: list of [R, N_g] arrays (same R for every realization).
    tune: pick the step kernel's launch shapes (plain and doubling) with the
    autotuner before the batch's work-group chunks are laid out.
    stride_chunks: cut each realization whose matrix has a dominant row
    stride (``lattice_stride``, e.g. L for an L x L square lattice) into
    chunks of that many rows; False keeps the proportional cut."""

    def __init__(self, blocks, vectors, M, dtype="c64", doubling=False, tune=False,
                 stride_chunks=True):
        nat.require_device()
        data, idx, offsets, K = block_diagonal(blocks)
        H = nat.DeviceELL(data, idx, int(offsets[-1]), K, 1.0, dtype)
        del data, idx
        strides = [_ell_stride(b[1], b[2], b[3]) for b in blocks] if stride_chunks else None
        self._setup(H, offsets, vectors, [int(b[2]) for b in blocks], M, dtype, doubling, tune,
                    strides)

    @classmethod
    def from_csr(cls, mats, scales, vectors, M, dtype="c64", doubling=False, tune=False,
                 stride_chunks=True):
        """The batch from scipy CSR realizations (square, each with its scale
        1/B_g): their CSR arrays are stacked block-diagonally on the host and
        converted to ELL on the device (jkpm_ell_create_csr), so no host ELL
        is built.  The ELL width is the longest row of any realization, and
        the slots follow CSR order with the reference padding, as
        ``BatchKPM(blocks, ...)`` over ``ell_matrix`` blocks."""
        nat.require_device()
        indptr, indices, data, offsets = stack_csr(mats, scales)
        H = nat.DeviceELL.from_csr(indptr, indices, data, int(offsets[-1]), 1.0, dtype)
        del indptr, indices, data
        self = cls.__new__(cls)
        strides = [_csr_stride(m) for m in mats] if stride_chunks else None
        self._setup(H, offsets, vectors, [int(m.shape[0]) for m in mats], M, dtype, doubling,
                    tune, strides)
        return self

    def _setup(self, H, offsets, vectors, sizes, M, dtype, doubling, tune, strides=None):
        torch = nat.require_device()
        self.offsets = offsets
        self.M = int(M)
        self.doubling = bool(doubling)
        self.H = H
        v = np.concatenate([np.asarray(vg).reshape(-1, n) for vg, n in zip(vectors, sizes)],
                           axis=1)
        self.R = v.shape[0]
        self.vr = _kpm.as_batch(v, int(offsets[-1]), dtype)
        del v
        if tune:
            # the batch's chunk tables follow the handle's launch settings
            from . import autotune

            autotune.tuned(self.H, self.vr)
            autotune.tuned(self.H, self.vr, doubling=True)
        # Chunks of one lattice stride: the work-groups of consecutive chunks
        # run side by side on one XCD and read each other's vertical
        # neighbour rows at the same time, so T_n is fetched from HBM once.
        self.strides = strides
        self.batch = nat.Batch(self.H, offsets,
                               None if not strides or not any(strides) else strides)
        if tune and strides and any(strides):
            # the work-group count now comes from the strides: pick the
            # variants on the batch itself
            from . import autotune

            autotune.tuned_batch(self)
        self.ws = None
        self.mu = torch.empty((self.batch.G, self.M, self.R), dtype=torch.complex128, device="cuda")
        self.graph = None
        self.stream = None

    def _launch(self, stream):
        if self.ws is None:
            import torch

            need = nat.moments_ex_workspace_bytes(None, self.batch, self.R, self.M, self.doubling)
            self.ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        return nat.moments_ex(None, self.batch, self.vr, self.M, self.doubling,
                              workspace=self.ws, out=self.mu, stream=stream)

    def run(self, stream=None):
        """Eager launch of the whole batch; returns mu [G, M, R] (device)."""
        return self._launch(stream)

    def capture(self):
        """Captures the recursion into a HIP graph on a private stream."""
        torch = nat.require_device()
        self.stream = torch.cuda.Stream()
        self.stream.wait_stream(torch.cuda.current_stream())
        self._launch(torch.cuda.current_stream())  # allocates the workspace outside capture
        torch.cuda.synchronize()
        with nat.Graph.capture(self.stream) as g:
            self._launch(self.stream)
        self.graph = g
        return g

    def replay(self):
        """Replays the captured graph; returns mu [G, M, R] after the replay
        has finished on the capture stream."""
        torch = nat.require_device()
        if self.graph is None:
            self.capture()
        self.stream.wait_stream(torch.cuda.current_stream())
        self.graph.replay(self.stream)
        torch.cuda.current_stream().wait_stream(self.stream)
        return self.mu


def kpm_dos_batch(mats, M, random_vectors, energies=None, dtype="c64", kernel="jackson",
                  graph=True, doubling=False):
    """DOS tables [E, 2] for a list of ell_matrix realizations, computed in
    one batch (random_vectors: list of [R, N_g] or [N_g] arrays)."""
    blocks = []
    for A in mats:
        s = _kpm.bounds_scale(A.bounds)
        if s is None:
            raise ValueError("batched realizations need symmetric bounds")
        blocks.append((A.data, A.indices, int(A.shape[0]), int(A.len_row), s))
    vecs = [np.atleast_2d(v) for v in random_vectors]
    bk = BatchKPM(blocks, vecs, M, dtype, doubling)
    mu = bk.replay() if graph else bk.run()
    out = []
    for g, A in enumerate(mats):
        E = (np.linspace(A.bounds[0] + 1e-3, A.bounds[1] - 1e-3, 2 * M + 1)
             if energies is None else np.asarray(energies, dtype=np.float64))
        d = _kpm.dos_device(mu[g], _kpm.rescale_energies(E, A.bounds), kernel)
        out.append(np.stack([E, d.cpu().numpy()], 1))
    return out
