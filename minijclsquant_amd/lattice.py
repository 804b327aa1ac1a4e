"""Lattice Hamiltonian builders for the KPM workloads.

The reference has one builder, the honeycomb ``lattice_hexagonal`` +
``H_graphene`` (jclsquant/lattice.py:6-47, jclsquant/hams.py:63-168), with an
O(B^2) Python border pairing, unseeded disorder and a 1e-6 cross-perturbation
that makes H non-Hermitian.  Parity with the reference is taken at the ELL
boundary (same data/indices/vectors -> same moments), so these builders are
this package's own: vectorised, seeded and exactly Hermitian.

* ``square_lattice``    L x L, nearest-neighbour hopping t, onsite energies,
                        K = 5 (4 bonds + the diagonal slot, always present).
* ``honeycomb_lattice`` Lx x Ly two-atom cells (N = 2 Lx Ly), K = 4
                        (3 bonds + diagonal), with the bond vectors Dx, Dy for
                        the velocity operator (entry (i, j) = x_j - x_i, the
                        reference's convention, jclsquant/hams.py:96-104).
* ``lattice_hexagonal`` the reference's 4-atom-cell position builder.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

__all__ = [
    "square_lattice",
    "honeycomb_lattice",
    "lattice_hexagonal",
    "A_CC",
]

A_LAT = 0.24595          # nm, graphene lattice constant (jclsquant/hams.py:81)
A_CC = A_LAT / np.sqrt(3)  # nm, carbon-carbon distance


def _uniform(rng, n, lo, hi):
    return lo + (hi - lo) * rng.random(n)


def square_lattice(L, t=-1.0, W=0.0, seed=None, periodic=True, onsite=None,
                   bond_noise=0.0, Ly=None):
    """Square lattice, site i = y*Lx + x.

    onsite: explicit array of N onsite energies, else U[-W/2, W/2] from
    default_rng(seed) (W = 0 gives zeros).  bond_noise: per-bond hopping
    t_ij = t + N(0, bond_noise^2) (same rng, drawn after the onsite terms),
    applied symmetrically so H stays Hermitian.  Returns a CSR matrix
    (complex128, sorted indices, explicit diagonal) and the per-bond
    hopping array [2, N] (bonds to +x and +y of each site)."""
    Lx = int(L)
    Ly = Lx if Ly is None else int(Ly)
    N = Lx * Ly
    rng = np.random.default_rng(seed)
    if onsite is None:
        eps = _uniform(rng, N, -W / 2, W / 2) if W else np.zeros(N)
    else:
        eps = np.asarray(onsite, dtype=np.float64)
    hop = np.full((2, N), float(t))
    if bond_noise:
        hop = hop + rng.normal(0.0, bond_noise, size=(2, N))
    idx = np.arange(N)
    x, y = idx % Lx, idx // Lx
    rows = [idx]
    cols = [idx]
    vals = [eps.astype(np.complex128)]
    for d, (nx, ny, inside) in enumerate((((x + 1) % Lx, y, x + 1 < Lx),
                                          (x, (y + 1) % Ly, y + 1 < Ly))):
        j = ny * Lx + nx
        keep = np.ones(N, bool) if periodic else inside
        i_k, j_k, h_k = idx[keep], j[keep], hop[d][keep]
        rows += [i_k, j_k]
        cols += [j_k, i_k]
        vals += [h_k.astype(np.complex128), h_k.astype(np.complex128)]
    H = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N)).tocsr()
    H.sum_duplicates()
    H.sort_indices()
    return H, hop


def honeycomb_lattice(Lx, Ly=None, t=-2.7, W=0.0, seed=None, periodic=True, onsite=None):
    """Honeycomb lattice of Lx x Ly two-atom cells, N = 2 Lx Ly.

    Cell (x, y) at R = x a1 + y a2 (a1 = a(1, 0), a2 = a(1/2, sqrt(3)/2));
    site A = 2(y Lx + x) at R, site B = A + 1 at R + (a1 + a2)/3.  Bonds:
    A(x,y)-B(x,y), A(x,y)-B(x-1,y), A(x,y)-B(x,y-1), wrapped periodically.
    Onsite energies U[-W/2, W/2] from default_rng(seed) unless `onsite` is
    given.  Returns (H, Dx, Dy) as CSR matrices with one pattern (Dx/Dy hold
    the minimum-image bond vectors x_j - x_i, y_j - y_i in nm; diagonal 0)."""
    Lx = int(Lx)
    Ly = Lx if Ly is None else int(Ly)
    N = 2 * Lx * Ly
    rng = np.random.default_rng(seed)
    if onsite is None:
        eps = _uniform(rng, N, -W / 2, W / 2) if W else np.zeros(N)
    else:
        eps = np.asarray(onsite, dtype=np.float64)
    a1 = np.array([A_LAT, 0.0])
    a2 = np.array([A_LAT / 2, A_LAT * np.sqrt(3) / 2])
    delta = (a1 + a2) / 3
    c = np.arange(Lx * Ly)
    x, y = c % Lx, c // Lx
    A = 2 * c
    rows, cols, vals, dxs, dys = [np.arange(N)], [np.arange(N)], [eps.astype(np.complex128)], \
        [np.zeros(N)], [np.zeros(N)]
    bond_types = (
        (x, y, delta, np.ones(Lx * Ly, bool)),
        ((x - 1) % Lx, y, delta - a1, x >= 1),
        (x, (y - 1) % Ly, delta - a2, y >= 1),
    )
    for bx, by, vec, inside in bond_types:
        keep = np.ones(Lx * Ly, bool) if periodic else inside
        B = 2 * (by * Lx + bx) + 1
        a_k, b_k = A[keep], B[keep]
        rows += [a_k, b_k]
        cols += [b_k, a_k]
        vals += [np.full(a_k.size, t, dtype=np.complex128)] * 2
        dxs += [np.full(a_k.size, vec[0]), np.full(a_k.size, -vec[0])]
        dys += [np.full(a_k.size, vec[1]), np.full(a_k.size, -vec[1])]
    r = np.concatenate(rows)
    cc = np.concatenate(cols)

    def mk(v):
        M = sp.coo_matrix((np.concatenate(v).astype(np.complex128), (r, cc)), shape=(N, N)).tocsr()
        M.sort_indices()
        return M

    H = mk(vals)
    Dx = mk(dxs)
    Dy = mk(dys)
    if H.nnz != Dx.nnz:  # duplicates merged only in tiny periodic lattices
        raise ValueError("lattice too small: bonds coincide under periodic wrapping")
    return H, Dx, Dy


def lattice_hexagonal(N):
    """Positions of a honeycomb flake built from 4-atom rectangular cells
    (jclsquant/lattice.py:6-47); N should be 4 k^2."""
    a_cc = A_CC
    N_tot = N // 4
    if int(np.sqrt(N_tot)) / np.sqrt(N_tot) != 1.0:
        print('The number of total atoms wont be the one inserted if you want otherwise '
              'insert an even power of 2')
    a1 = np.array([np.sqrt(3) * a_cc / 2, a_cc / 2])
    a2 = np.array([0, a_cc])
    S_0 = np.array([[0.0, 0.0], a2, a2 + a1, a2 + a1 + a2])
    ncx = int(np.sqrt(N_tot))
    ncy = ncx
    ix = np.arange(ncx)
    Sx = (S_0[None, :, :] + np.stack([ix * np.sqrt(3) * a_cc, np.zeros(ncx)], 1)[:, None, :])
    Sx = Sx.reshape(-1, 2)
    iy = np.arange(ncy)
    S = Sx[None, :, :] + np.stack([np.zeros(ncy), iy * 3 * a_cc], 1)[:, None, :]
    return S.reshape(-1, 2)
