"""GPU-named KPM entry points of the reference.

Mirrors jclsquant/observables_gpu.py:12-93 (kpm_dos_gpu, kpm_dos_2_gpu) and
the binding functions rec_A_tab_gpu / rec_A_tab_2_gpu of
jclsquant/cuda_cython/obv.pyx:23-49, which fill a caller-owned complex128
array in place through the extern "C" symbols rec_A_tab / rec_A_tab_2.
"""
from __future__ import annotations

import numpy as np

from . import _native as nat
from .observables import kpm_dos, kpm_n_dos_n, random_vector_generator

__all__ = ["kpm_dos_gpu", "kpm_dos_2_gpu", "rec_A_tab_gpu", "rec_A_tab_2_gpu",
           "kpm_rho_neq_gpu", "rho_neq_propagate"]


def kpm_dos_gpu(H, M=None, random_vector=None, **kw):
    """Same contract as kpm_dos (jclsquant/observables_gpu.py:12-42)."""
    return kpm_dos(H, M, random_vector, **kw)


def kpm_dos_2_gpu(H, M=None, random_vector_l=None, random_vector_r=None, bounds=True,
                  proyector=None, **kw):
    """Same contract as kpm_n_dos_n (jclsquant/observables_gpu.py:46-93)."""
    return kpm_n_dos_n(H, M, random_vector_l, random_vector_r, bounds, proyector, **kw)


def _check_out(A_vec, M):
    if not (isinstance(A_vec, np.ndarray) and A_vec.dtype == np.complex128
            and A_vec.flags.c_contiguous and A_vec.size >= M):
        raise ValueError("A_vec must be a C-contiguous complex128 array with >= M entries")


def rec_A_tab_gpu(data, indices, len_row, v, A_vec, M):
    """A_vec[:M] = <v|T_m(A)|v> via the extern "C" rec_A_tab symbol."""
    _check_out(A_vec, M)
    A_vec[:M] = nat.rec_A_tab_host(data, indices, int(len_row), v, int(M))


def rec_A_tab_2_gpu(data, indices, len_row, v_l, v_r, A_vec, M):
    """A_vec[:M] = <v_l|T_m(A)|v_r> via the extern "C" rec_A_tab_2 symbol."""
    _check_out(A_vec, M)
    A_vec[:M] = nat.rec_A_tab_host(data, indices, int(len_row), v_l, int(M), v_r=v_r)


def _time_moments(t_vec, H, M):
    """Time-step coefficients exp(-i x dt) truncated where |c_m| <= 1e-15
    (jclsquant/observables_gpu.py:138-147); at least the two terms the
    series always forms."""
    from .constants import hbar_fs
    from .kernel_and_moments import moments_U

    delta_t = t_vec[1]
    delta_t_kpm = delta_t * (0.5 * (H.bounds[1] - H.bounds[0])) / hbar_fs
    c = moments_U(delta_t_kpm, M)
    m = 0
    while m < c.size and np.abs(c[m]) > 1e-15:
        m += 1
    if m >= c.size:
        raise ValueError("the time step needs more than M Chebyshev terms; raise M")
    M2 = max(m, 2)
    return moments_U(delta_t_kpm, M2), M2


def rho_neq_propagate(H_kpm, t_vec, v, moments_FD, moments_Uv, modifier_id, modifier_params,
                      dtype="c128"):
    """Device part of kpm_rho_neq_gpu (jclsquant/cuda_cython/cuda/obv.cu:262-361):
    H(t_0) = Peierls(H_kpm, t_0), U_F = sum_m c^FD_m T_m(H(t_0)) v, then for
    every later t_i both U and U_F advance by sum_m c^U_m T_m(H(t_i)).  The
    two vectors share each pass over H(t_i) (one R = 2 batch) and the
    Peierls update is one device pass per time step.  Returns the device
    tensor [N, 2] = (U, U_F) of the given dtype."""
    from . import _kpm
    from . import _native as nat
    from .modifiers import light_phases, peierls_handles

    torch = nat.require_device()
    h0, bonds, ht = peierls_handles(H_kpm, dtype)
    n = int(H_kpm.shape[0])

    def phases(t):
        ph = light_phases(modifier_id, modifier_params, t)
        if ph is None:
            raise ValueError(f"invalid light modifier {modifier_id!r} / {modifier_params!r}")
        return ph

    x = _kpm.as_batch(v, n, dtype)
    ht.peierls(h0, bonds, 1.0, *phases(t_vec[0]))
    uf, _ = nat.series(ht, None, x, len(moments_FD), moments_FD)
    X = torch.cat([x, uf], dim=1).contiguous()
    for t in t_vec[1:]:
        ht.peierls(h0, bonds, 1.0, *phases(t))
        X, _ = nat.series(ht, None, X, len(moments_Uv), moments_Uv)
    return X


def kpm_rho_neq_gpu(H, t_vec=None, tau=0, modifier_id=None, modifier_params=None, Temp=None,
                    Ef=None, observale_list=None, M=None, random_vector=None, proyector=None):
    """Time evolution of the occupied part of a random vector under light
    (jclsquant/observables_gpu.py:96-265), for real-space matrices and the
    light modes 'circle_packed' / 'linear_packed'.  With the observable
    ['n', n_meass, M_n] it returns (n-weighted DOS, DOS) of H(t_end) from
    <U|T_m|U_F> and <U|T_m|U>, as the reference does; without observables
    it returns None.  Temp None is taken as 0 K (the reference's T = 0
    branch)."""
    from .constants import hbar_fs, kb
    from .kernel_and_moments import JacksonKernel, moments_FD_T, moments_FD_T0
    from .modifiers import modifier_bounds, modifier_hoppings_c

    if getattr(H, "space", "r") == "k":
        raise NotImplementedError("k-space time evolution is not part of this package")
    if M is None:
        M = int(np.sqrt(H.shape[0]))
    if Ef is None:
        print('No Fermi energy has been included by default is set to 0 .')
        Ef = 0
    if random_vector is None:
        random_vector = random_vector_generator(H.shape[0])
    if t_vec is None:
        print('No time vector has been inserted by default is set from 0 to 5*tau_phi with '
              'tau_phi=hbar*M/(pi*bounds[1]) with tau_phi/100 steps')
        t_vec = np.linspace(0, 5 * hbar_fs * M / (np.pi * (0.5 * (H.bounds[1] - H.bounds[0]))),
                            500)
    if Temp is None:
        print('No temperature inserted by default is set to zero.')
        Temp = 0
    if modifier_params is None:
        print('You must introduce an array that contains the elements necesary to apply the '
              'time evolution to the hamiltonian .')
        return None
    if modifier_id is None:
        print('You must introduce an string with one of the following : linear , circle , '
              'linear packed , circle packed .')
        return None
    if modifier_id not in ("circle_packed", "linear_packed"):
        raise NotImplementedError(f"modifier {modifier_id!r} is not part of this package")
    t_vec = np.asarray(t_vec)
    moments_Uv, _ = _time_moments(t_vec, H, M)
    H_kpm = H.modifier(modifier_bounds)
    half = 0.5 * (H.bounds[1] - H.bounds[0])
    e_red = (Ef - 0.5 * (H.bounds[1] + H.bounds[0])) / half
    if Temp == 0:
        moments_kernel_FD = moments_FD_T0(e_red, M) * JacksonKernel(M)
    else:
        moments_kernel_FD = moments_FD_T(e_red, (Temp * kb) / half, M) * JacksonKernel(M)
    k_n = -1
    k_sigma = -1
    M_n = None
    for obs in (observale_list or []):
        if obs[0] == 'n':
            M_n = obs[2]
            k_n = 0
        elif obs[0] == 'sigma_nequil':
            k_sigma = 0
        else:
            print('There is no observable selected the function will return F(t=0) , U(t=0) , '
                  'F(t=end) and U(t=end) .')
    X = rho_neq_propagate(H_kpm, t_vec, random_vector, moments_kernel_FD.astype(np.complex128),
                          moments_Uv, modifier_id, modifier_params)
    UU = X.transpose(0, 1).cpu().numpy()
    U, U_F = UU[0].copy(), UU[1].copy()
    H_time = H_kpm.deep_copy()
    H_time.modifier_nocop(H_kpm, modifier_hoppings_c, modifier_id, modifier_params, t_vec[-1])
    if k_n >= 0 and k_sigma < 0:
        dos_n_vec = kpm_dos_2_gpu(H_time, M_n, U, U_F, False, proyector)
        dos = kpm_dos_2_gpu(H_time, M_n, U, U, False, proyector)
        return dos_n_vec, dos
    if k_n >= 0 and k_sigma >= 0:
        return kpm_dos_2_gpu(H_time, M_n, U, U_F, False, proyector)
    return None
