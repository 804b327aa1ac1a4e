"""ELL value modifiers used by the KPM path.

Mirrors jclsquant/modifiers.py:14-36 (modifier_bounds), :56-79
(modifier_velocity), :218-239 (modifier_diagonal) and :287-309
(modifier_diagonal_anderson).  Each returns a new value array for
``ell_matrix.modifier``; they are one-time input transforms on the host
arrays, as in the reference.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "modifier_bounds",
    "modifier_bounds_inverse",
    "modifier_velocity",
    "modifier_diagonal",
    "modifier_diagonal_anderson",
    "diagonal_modifier",
]


def diagonal_modifier(data, indices, valid_row_values, diagonal_array, len_row):
    """data_new[s] = diag[row] + data[s] on valid diagonal slots, else data[s]
    (jclsquant/cython_modules/Extra_funs.pyx:352-361)."""
    data = np.asarray(data)
    out = np.array(data, dtype=np.complex128, copy=True)
    rows = np.arange(data.size) // int(len_row)
    sel = (np.asarray(indices) == rows) & (np.asarray(valid_row_values) == 1)
    diag = np.asarray(diagonal_array)
    out[sel] = diag[rows[sel]] + data[sel]
    return out


def modifier_bounds(A):
    """H~ = H / ((b1-b0)/2) for symmetric bounds; otherwise shift the valid
    diagonal slots by -(b1+b0)/2 first."""
    b = A.bounds
    if np.abs(b[1]) == np.abs(b[0]):
        return np.copy(A.data) * (1 / ((b[1] - b[0]) / 2) + 0j)
    diag = (-0.5 * (b[1] + b[0])) * np.ones(int(A.shape[0]), dtype=np.complex128)
    data_new = diagonal_modifier(A.data, A.indices, A.valid_row_values, diag, A.len_row)
    return data_new * (1 / (0.5 * (b[1] - b[0])) + 0j)


def modifier_bounds_inverse(A):
    return np.copy(A.data) * (((A.bounds[1] - A.bounds[0]) / 2) + 0j)


def modifier_velocity(A, direction=None):
    """v = -i * H (elementwise) * bond vector along `direction` (0/'x', 1/'y',
    2/'z'), i.e. vec_mul(-1j, data, d_vec) = (-1j * d) * data."""
    if direction is None:
        print(' You must introduce a desired direction: 0,1,2 or x,y,z ')
        return
    if direction in (0, 'x'):
        d = A.dx_vec
    elif direction in (1, 'y'):
        d = A.dy_vec
    elif direction in (2, 'z'):
        d = A.dz_vec
    else:
        print(' You must introduce a valid direction: 0,1,2 or x,y,z ')
        return
    return (-1j * np.asarray(d, dtype=np.complex128)) * np.asarray(A.data)


def modifier_diagonal(A, id=None, diagonal_array=None):
    if diagonal_array is None:
        print('You must introduce valid modification of the diagonal .')
        return
    if len(diagonal_array) != A.shape[0]:
        print('You must introduce an array that is of the same shape as the matrix.')
    return diagonal_modifier(A.data, A.indices, A.valid_row_values, diagonal_array, A.len_row)


def modifier_diagonal_anderson(A, W=None, seed=None):
    """Onsite disorder W*(2*U[0,1) - 1) from default_rng(seed), added to the
    existing diagonal slots."""
    if W is None:
        print('You must introduce a disorder strength to modify the system. ')
    rng = np.random.default_rng(seed)
    diagonal_array = (W) * (2 * rng.random(int(A.shape[0])) - 1)
    return diagonal_modifier(A.data, A.indices, A.valid_row_values, diagonal_array, A.len_row)
