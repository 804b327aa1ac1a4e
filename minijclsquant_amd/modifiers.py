"""ELL value modifiers used by the KPM path.

Mirrors jclsquant/modifiers.py:14-36 (modifier_bounds), :56-79
(modifier_velocity), :218-239 (modifier_diagonal) and :287-309
(modifier_diagonal_anderson).  Each returns a new value array for
``ell_matrix.modifier``; they are one-time input transforms on the host
arrays, as in the reference.

The light (Peierls) modifiers of :133-181 (modifier_light) and :362-443
(modifier_hoppings_c, real space, light modes) run on the device:
``light_phases`` turns a light mode into the factor exp(i (ax dx + ay dy))
and jkpm_ell_peierls applies it to every slot in one pass.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "modifier_bounds",
    "modifier_bounds_inverse",
    "modifier_velocity",
    "modifier_diagonal",
    "modifier_diagonal_anderson",
    "light_phases",
    "modifier_light",
    "modifier_hoppings_c",
]

_LIGHT_MSG = (" You must introduce a valid id : light, circle,light_packed or circle packed "
              "with irs valid parameters .")


def _diagonal_values(data, indices, valid_row_values, diagonal_array, len_row):
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
    data_new = _diagonal_values(A.data, A.indices, A.valid_row_values, diag, A.len_row)
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
    return _diagonal_values(A.data, A.indices, A.valid_row_values, diagonal_array, A.len_row)


def modifier_diagonal_anderson(A, W=None, seed=None):
    """Onsite disorder W*(2*U[0,1) - 1) from default_rng(seed), added to the
    existing diagonal slots."""
    if W is None:
        print('You must introduce a disorder strength to modify the system. ')
    rng = np.random.default_rng(seed)
    diagonal_array = (W) * (2 * rng.random(int(A.shape[0])) - 1)
    return _diagonal_values(A.data, A.indices, A.valid_row_values, diagonal_array, A.len_row)


def light_phases(modifier_id, modifier_params, t):
    """(ax, ay) of the Peierls factor exp(i (ax dx + ay dy)) at time t, or
    None (after the reference's message) for an unknown mode/polarisation:

      linear        (A0, w):          ax = 0,             ay = A0 sin(wt)
      circle        (A0, w, pol):     ax = A0 cos(wt),    ay = +-A0 sin(wt)
      linear_packed (A0, w, Tp):      as linear / cosh((t - 2Tp)/(0.5673 Tp))
      circle_packed (A0, w, pol, Tp): as circle / cosh((t - 2Tp)/(0.5673 Tp))

    pol 'r' takes +, 'l' takes -.  Complex parameters are kept complex (an
    imaginary phase scales the magnitude), as the reference's complex
    arithmetic does (jclsquant/modifiers.py:133-181,362-443)."""
    if modifier_id is None or modifier_params is None:
        print(_LIGHT_MSG)
        return None
    if modifier_id == "linear":
        A0, w = modifier_params
        return 0j, A0 * np.sin(w * t) + 0j
    if modifier_id == "linear_packed":
        A0, w, Tp = modifier_params
        env = np.cosh((t - 2 * Tp) / (0.5673 * Tp))
        return 0j, A0 * np.sin(w * t) / env + 0j
    if modifier_id in ("circle", "circle_packed"):
        if modifier_id == "circle":
            A0, w, pol = modifier_params
            env = 1.0
        else:
            A0, w, pol, Tp = modifier_params
            env = np.cosh((t - 2 * Tp) / (0.5673 * Tp))
        if pol not in ("r", "l"):
            print('You must introduce a valid polarization')
            return None
        sgn = 1.0 if pol == "r" else -1.0
        return A0 * np.cos(w * t) / env + 0j, sgn * A0 * np.sin(w * t) / env + 0j
    print(_LIGHT_MSG)
    return None


def _bond_values(A):
    dx = np.asarray(A.dx_vec)
    dy = np.asarray(A.dy_vec)
    if np.any(np.imag(dx) != 0) or np.any(np.imag(dy) != 0):
        raise ValueError("bond vectors dx_vec/dy_vec must be real")
    return np.real(dx) + 1j * np.real(dy)


def peierls_handles(A, dtype="c128", scale=1.0):
    """Device handles (H, bonds, H(t)) of A for repeated Peierls updates:
    H holds A.data * scale, bonds holds dx + i dy on the same slots and
    H(t) is the output buffer.  Cached on A until its data changes."""
    from ._native import DeviceELL

    if A.dx_vec is None or A.dy_vec is None:
        raise ValueError("the light modifiers need bond vectors: build the matrix with DX/DY")
    key = ("peierls", dtype, float(scale))
    hit = A._dev.get(key)
    if hit is None:
        h = A.device(dtype, scale)
        bonds = DeviceELL.like(h, _bond_values(A), 1.0, "c128")
        out = DeviceELL.like(h, A.data, scale, dtype)
        hit = (h, bonds, out)
        A._dev[key] = hit
    return hit


def modifier_light(A, data_new, modifier_id=None, modifier_params=None, t=None):
    """data_new[:] = A.data * exp(i (ax dx + ay dy)) for the light mode at
    time t (jclsquant/modifiers.py:133-181), computed on the device; returns
    data_new (None after the reference's message for an invalid mode)."""
    ph = light_phases(modifier_id, modifier_params, t)
    if ph is None:
        return None
    h, bonds, out = peierls_handles(A)
    out.peierls(h, bonds, 1.0, ph[0], ph[1])
    data_new[...] = out.values().cpu().numpy().reshape(np.shape(data_new))
    return data_new


def modifier_hoppings_c(A, data_new, dx_vec_new, dy_vec_new, modifier_id=None,
                        modifier_params=None, t=None):
    """Real-space hopping modifier of the light modes
    (jclsquant/modifiers.py:362-443); the phonon mode and k-space matrices
    are not part of this package."""
    if getattr(A, "space", "r") == "k":
        raise NotImplementedError("k-space hopping modifiers are not part of this package")
    if modifier_id == "phonons_packed":
        raise NotImplementedError("the phonon hopping modifier is not part of this package")
    return modifier_light(A, data_new, modifier_id, modifier_params, t)
