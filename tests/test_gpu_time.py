"""GPU parity of the device Peierls substitution, the value readback and the
light-driven time evolution (kpm_rho_neq_gpu) against the oracle.

Tolerances: complex128 Peierls factor <= 1e-14 of max |value| (FP64 sincos
and exp against numpy's); complex64 <= 1e-6; the time loop (series of up to
M terms per step, complex128) <= 1e-10 of max |entry|; DOS tables from the
evolved vectors <= 1e-9 of max |DOS|.
"""
import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu

B = (3 * 2.7 + 0.5) * 1.01


def lattice(L=6, seed=1):
    H, Dx, Dy = jcl.honeycomb_lattice(L, t=-2.7, W=1.0, seed=seed)
    A = jcl.ell_matrix(H, Dx, Dy)
    A.bounds = np.array([-B, B])
    return A


def rel(got, ref):
    return np.max(np.abs(got - ref)) / np.max(np.abs(ref))


@pytest.mark.parametrize("dtype,tol", [("c128", 0.0), ("c64", 1e-7)])
def test_values_round_trip(torch_gpu, dtype, tol):
    A = lattice(5, 2)
    h = nat.DeviceELL(A.data, A.indices, A.shape[0], A.len_row, 1 / B, dtype)
    got = h.values().cpu().numpy()
    assert rel(got, A.data) <= max(tol, 1e-15)


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-14), ("c64", 1e-6)])
def test_peierls_vs_oracle(torch_gpu, dtype, tol):
    A = lattice(6, 3)
    N, K = A.shape[0], A.len_row
    h = nat.DeviceELL(A.data, A.indices, N, K, 1 / B, dtype)
    bonds = nat.DeviceELL.like(h, np.real(A.dx_vec) + 1j * np.real(A.dy_vec), 1.0, "c128")
    out = nat.DeviceELL.like(h, A.data, 1 / B, dtype)
    for a, ax, ay in [(1.0, 3.0, -2.0), (0.5 - 0.25j, 1.5 + 0.2j, 0.7 - 0.1j)]:
        out.peierls(h, bonds, a, ax, ay)
        want = O.peierls(A.data, A.dx_vec, A.dy_vec, a, ax, ay)
        assert rel(out.values().cpu().numpy(), want) <= tol
    # in place: dst = src applies the factor on top of the stored values
    h2 = nat.DeviceELL(A.data, A.indices, N, K, 1 / B, dtype)
    b2 = nat.DeviceELL.like(h2, np.real(A.dx_vec) + 1j * np.real(A.dy_vec), 1.0, "c128")
    h2.peierls(h2, b2, 1.0, 2.0, 1.0)
    h2.peierls(h2, b2, 1.0, -2.0, -1.0)
    assert rel(h2.values().cpu().numpy(), A.data) <= max(tol, 1e-14)


def test_peierls_rejects_foreign_pattern(torch_gpu):
    A, A2 = lattice(4, 1), lattice(4, 2)
    h = nat.DeviceELL(A.data, A.indices, A.shape[0], A.len_row)
    other = nat.DeviceELL(A2.data, A2.indices, A2.shape[0], A2.len_row)
    with pytest.raises(ValueError):
        other.peierls(h, h, 1.0, 1.0, 1.0)


@pytest.mark.parametrize("mode,params", [("linear", (4.0, 0.8)), ("circle", (4.0, 0.8, "l")),
                                         ("linear_packed", (4.0, 0.8, 2.0)),
                                         ("circle_packed", (4.0, 0.8, "r", 2.0))])
def test_light_modifiers_vs_oracle(torch_gpu, mode, params):
    A = lattice(5, 4)
    t = 1.7
    ax, ay = jcl.light_phases(mode, params, t)
    want = O.peierls(A.data, A.dx_vec, A.dy_vec, 1.0, ax, ay)
    out = np.zeros_like(A.data)
    assert jcl.modifier_light(A, out, mode, params, t) is out
    assert rel(out, want) <= 1e-14
    out2 = np.zeros_like(A.data)
    jcl.modifier_hoppings_c(A, out2, A.dx_vec.copy(), A.dy_vec.copy(), mode, params, t)
    assert rel(out2, want) <= 1e-14
    assert jcl.modifier_light(A, out, "bogus", params, t) is None


def test_rho_neq_propagate_vs_oracle(torch_gpu):
    from minijclsquant_amd.observables_gpu import rho_neq_propagate

    A = lattice(6, 5)
    Hk = A.modifier(jcl.modifier_bounds)
    N, K = A.shape[0], A.len_row
    v = jcl.random_vector_generator(N, 9)
    t_vec = np.linspace(0.0, 2.0, 6)
    params = (6.0, 0.9, "r", 1.0)
    from minijclsquant_amd.constants import hbar_fs

    c_u = jcl.moments_U(t_vec[1] * B / hbar_fs, 40)
    c_fd = (jcl.moments_FD_T0(0.05, 64) * jcl.JacksonKernel(64)).astype(np.complex128)
    X = rho_neq_propagate(Hk, t_vec, v, c_fd, c_u, "circle_packed", params).cpu().numpy()

    def phases(t):
        return jcl.light_phases("circle_packed", params, t)

    u, uf = O.rho_neq(Hk.data, Hk.indices, K, Hk.dx_vec, Hk.dy_vec, t_vec, v, c_fd, c_u, phases)
    assert rel(X[:, 0], u) <= 1e-10
    assert rel(X[:, 1], uf) <= 1e-10


def test_kpm_rho_neq_gpu_vs_oracle(torch_gpu, capsys):
    """The whole driver: evolved vectors, H(t_end) and the two DOS tables."""
    from minijclsquant_amd.constants import hbar_fs
    from minijclsquant_amd.observables_gpu import _time_moments

    A = lattice(6, 6)
    N, K = A.shape[0], A.len_row
    v = jcl.random_vector_generator(N, 11)
    t_vec = np.linspace(0.0, 1.5, 4)
    M, M_n = 64, 48
    params = (5.0, 1.1, 0.8)
    dos_n, dos = jcl.kpm_rho_neq_gpu(A, t_vec, 0, "linear_packed", params, Temp=0, Ef=0.3,
                                     observale_list=[["n", 1, M_n]], M=M, random_vector=v)
    # oracle pipeline
    Hk = A.modifier(jcl.modifier_bounds)
    c_u, M2 = _time_moments(t_vec, A, M)
    assert np.abs(jcl.moments_U(t_vec[1] * B / hbar_fs, M2 + 1)[M2]) <= 1e-15
    c_fd = jcl.moments_FD_T0(0.3 / B, M) * jcl.JacksonKernel(M)

    def phases(t):
        return jcl.light_phases("linear_packed", params, t)

    u, uf = O.rho_neq(Hk.data, Hk.indices, K, Hk.dx_vec, Hk.dy_vec, t_vec, v, c_fd, c_u, phases)
    d_end = O.peierls(Hk.data, Hk.dx_vec, Hk.dy_vec, 1.0, *phases(t_vec[-1]))
    E = np.linspace(-B + 1e-3, B - 1e-3, 2 * M_n + 1)
    for got, vl in ((dos_n, uf), (dos, u)):
        mu = O.rec_A_tab2v(M_n, d_end, Hk.indices, K, u, vl)
        want = O.dos(mu, E, -B, B)
        assert np.array_equal(got[:, 0], E)
        assert rel(got[:, 1], want) <= 1e-9
    # no observable: nothing returned, as in the reference
    assert jcl.kpm_rho_neq_gpu(A, t_vec, 0, "linear_packed", params, Temp=0, Ef=0.0,
                               observale_list=[], M=M, random_vector=v) is None
