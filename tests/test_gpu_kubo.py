"""GPU parity of the Kubo double moments mu_mn (complex64 stacks, FP32 MFMA
GEMM, FP64 combine) and of the sigma(E) contraction.

Tolerances: mu_mn against the complex128 oracle, max |dmu| <= 2e-5 * max |mu|
(complex64 recursion + FP32 accumulation over N*R terms); sigma kernel
against the oracle's FP64 contraction of the same mu <= 1e-10 relative.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from numpy.polynomial import chebyshev as C

import minijclsquant_amd as jcl
from minijclsquant_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL_MU = 2e-5


def lattice(L, W=1.0, seed=0):
    H, Dx, Dy = jcl.honeycomb_lattice(L, t=-2.7, W=W, seed=seed)
    A = jcl.ell_matrix(H, Dx, Dy)
    B = (3 * 2.7 + W / 2) * 1.01
    A.bounds = np.array([-B, B])
    return A, B


def vectors(N, R, seed0=0):
    return np.stack([jcl.random_vector_generator(N, seed0 + s) for s in range(R)])


def oracle_mu(A, B, v_data, r, M):
    hd = A.data * (1 / B)
    return np.mean([O.mu_nm(M, hd, A.indices, A.len_row, v_data, A.indices, A.len_row, v, )
                    for v in r], axis=0)


@pytest.mark.parametrize("L,M,R", [(8, 32, 2), (6, 40, 1), (10, 130, 3), (16, 64, 4)])
def test_mu_nm_vs_oracle(torch_gpu, L, M, R):
    A, B = lattice(L, seed=L)
    r = vectors(A.shape[0], R, 10 * L)
    mu = jcl.kpm_kubo_moments(A, M, r, "x").cpu().numpy()
    ref = oracle_mu(A, B, jcl.modifier_velocity(A, "x"), r, M)
    err = np.max(np.abs(mu - ref)) / np.max(np.abs(ref))
    assert err <= TOL_MU, err


def test_packed_values_match_complex_values(torch_gpu):
    """A real H and its velocity -i d H (purely imaginary) are detected at
    upload; the Kubo step then reads one float per slot of each.  With the
    packed copies switched off the same moments come from the complex
    values: equal to the complex64 bar, and both against the oracle."""
    from minijclsquant_amd.kubo import velocity_handle
    from minijclsquant_amd import _kpm

    A, B = lattice(10, seed=4)
    r = vectors(A.shape[0], 3, 77)
    Hd, Vd = velocity_handle(A, "x")
    assert Hd.is_real() and not Hd.is_imag()
    assert Vd.is_imag() and not Vd.is_real()
    rb = _kpm.as_batch(r, int(A.shape[0]), "c64")
    M = 48
    packed = nat.kubo_mu_nm(Hd, Vd, rb, M).cpu().numpy()
    Hd.use_real_values(False)
    Vd.use_real_values(False)
    cplx = nat.kubo_mu_nm(Hd, Vd, rb, M).cpu().numpy()
    ref = oracle_mu(A, B, jcl.modifier_velocity(A, "x"), r, M)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(packed - cplx)) <= TOL_MU * scale
    for mu in (packed, cplx):
        assert np.max(np.abs(mu - ref)) <= TOL_MU * scale
    # a value upload that is neither real nor imaginary clears both kinds
    Vd.set_values(jcl.modifier_velocity(A, "x") + 0.5, 1.0)
    assert not Vd.is_imag() and not Vd.is_real()


def test_mu_nm_y_direction_and_single_vector(torch_gpu):
    A, B = lattice(8, seed=3)
    r = jcl.random_vector_generator(A.shape[0], 5)
    mu = jcl.kpm_kubo_moments(A, 24, r, "y").cpu().numpy()
    ref = oracle_mu(A, B, jcl.modifier_velocity(A, "y"), r[None, :], 24)
    assert np.max(np.abs(mu - ref)) / np.max(np.abs(ref)) <= TOL_MU


def identity_velocity(A):
    rows = np.arange(A.data.size) // A.len_row
    return np.where((A.indices == rows) & (A.valid_row_values == 1), 1.0 + 0j, 0.0 + 0j)


@pytest.mark.parametrize("L,M,R", [(32, 96, 4), (512, 512, 4)])
def test_mu_nm_identity_velocity(torch_gpu, L, M, R):
    """v = I: mu_mn = (mu_{m+n} + mu_{|m-n|}) / 2 from the single moments.
    (512, 512, 4) is the Kubo workload's full size: honeycomb 512 x 512 cells,
    N = 524,288, M = 512, R = 4."""
    A, B = lattice(L, seed=L)
    N = A.shape[0]
    r = vectors(N, R, 77)
    mu2 = jcl.kpm_kubo_moments(A, M, r, v_data=identity_velocity(A)).cpu().numpy()
    single = jcl.kpm_moments(A, 2 * M, r, dtype="c128").cpu().numpy().mean(axis=0)
    m = np.arange(M)
    want = 0.5 * (single[m[:, None] + m[None, :]] + single[np.abs(m[:, None] - m[None, :])])
    err = np.max(np.abs(mu2 - want)) / N
    assert err <= TOL_MU, err


@pytest.mark.parametrize("kind", ["greenwood", "g_delta"])
def test_sigma_kernel_vs_oracle(torch_gpu, kind):
    torch = torch_gpu
    rng = np.random.default_rng(1)
    M, B = 96, 3.3
    mu = rng.normal(size=(M, M)) + 1j * rng.normal(size=(M, M))
    E = np.linspace(-0.98, 0.98, 37)
    got = nat.kubo_sigma(torch.from_numpy(mu).cuda(), E, jcl.kubo.KINDS[kind], B).cpu().numpy()
    want = O.kubo_sigma(mu, E, B, kind)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) <= 1e-10


def test_kpm_sigma_end_to_end(torch_gpu):
    A, B = lattice(12, seed=9)
    r = vectors(A.shape[0], 2, 4)
    M = 64
    E = np.linspace(-B + 0.05, B - 0.05, 101)
    s = jcl.kpm_sigma(A, M, r, "x", energies=E)
    ref_mu = oracle_mu(A, B, jcl.modifier_velocity(A, "x"), r, M)
    want = O.kubo_sigma(ref_mu, E / B, B, "greenwood").real
    assert s.shape == (E.size, 2) and np.array_equal(s[:, 0], E)
    assert np.max(np.abs(s[:, 1] - want)) / np.max(np.abs(want)) <= 1e-4
    s2 = jcl.kpm_sigma(A, M, r, "x", energies=E, kind="g_delta")
    want2 = O.kubo_sigma(ref_mu, E / B, B, "g_delta").real
    assert np.max(np.abs(s2[:, 1] - want2)) / np.max(np.abs(want2)) <= 1e-4


def test_kpm_sigma_trace_vs_exact_diagonalisation(torch_gpu):
    """Every basis vector as a random vector (R = N = 32): kpm_sigma at
    physical energies E equals (1/N) sum_ab |<a|v|b>|^2 d_a d_b with
    d_a = sum_m g_m delta_m(E/B) T_m(eps_a) from dense eigh of H/B, i.e.
    (1/N) Tr[v delta_K(E/B - H~) v delta_K(E/B - H~)].  Catches energies not
    rescaled by B and a kernel applied on one side only."""
    A, B = lattice(4, seed=5)
    N, K = A.shape[0], A.len_row
    rows = np.arange(A.data.size) // K
    Ht = sp.csr_matrix((A.data / B, (rows, A.indices)), shape=(N, N))
    V = sp.csr_matrix((jcl.modifier_velocity(A, "x"), (rows, A.indices)), shape=(N, N))
    M = 40
    E = np.linspace(-B + 0.3, B - 0.3, 23)
    s = jcl.kpm_sigma(A, M, np.eye(N, dtype=np.complex128), "x", energies=E)
    lam, U = np.linalg.eigh(Ht.toarray())
    vab2 = np.abs(U.conj().T @ V.toarray() @ U) ** 2
    g = O.jackson(M)
    want = []
    for e in E / B:
        d = C.chebval(lam, g * O.moments_delta(e, M))
        want.append(d @ vab2 @ d / N)
    want = np.array(want)
    assert np.max(np.abs(s[:, 1] - want)) <= 1e-4 * np.max(np.abs(want))


def test_kubo_argument_errors(torch_gpu):
    torch = torch_gpu
    A, B = lattice(4)
    H128 = A.device("c128", 1 / B)
    H64 = A.device("c64", 1 / B)
    V64 = nat.DeviceELL.like(H64, jcl.modifier_velocity(A, "x"), 1.0, "c64")
    other = nat.DeviceELL(A.data, A.indices, A.shape[0], A.len_row, 1 / B, "c64")
    r = torch.zeros((A.shape[0], 1), dtype=torch.complex64, device="cuda")
    V128 = nat.DeviceELL.like(H128, jcl.modifier_velocity(A, "x"), 1.0, "c128")
    with pytest.raises(ValueError):
        nat.kubo_mu_nm(H128, V128, r, 8)  # complex128 handles are refused
    with pytest.raises(ValueError):
        nat.kubo_mu_nm(H64, other, r, 8)  # v not on H's slot pattern
    with pytest.raises(ValueError):
        nat.kubo_sigma(torch.zeros((4, 4), dtype=torch.complex128, device="cuda"),
                       np.array([1.5]), 0, B)
    out = nat.kubo_mu_nm(H64, V64, r + 1, 8)
    assert out.shape == (8, 8)


@pytest.mark.parametrize("L,R,M", [(5, 1, 40), (7, 3, 33)])
def test_mu_nm_odd_k(torch_gpu, L, R, M):
    """K = N*R odd (unaligned operand rows) and M not a multiple of the tile:
    square lattices with the v = I identity against the single moments."""
    H, _ = jcl.square_lattice(L, W=1.0, seed=L)
    A = jcl.ell_matrix(H)
    B = 4.5 * 1.01
    A.bounds = np.array([-B, B])
    N = A.shape[0]
    assert (N * R) % 2 == 1
    r = vectors(N, R, 5)
    mu2 = jcl.kpm_kubo_moments(A, M, r, v_data=identity_velocity(A)).cpu().numpy()
    single = jcl.kpm_moments(A, 2 * M, r, dtype="c128").cpu().numpy().mean(axis=0)
    m = np.arange(M)
    want = 0.5 * (single[m[:, None] + m[None, :]] + single[np.abs(m[:, None] - m[None, :])])
    assert np.max(np.abs(mu2 - want)) / N <= TOL_MU


@pytest.mark.parametrize("L,R,M,odd", [(24, 4, 256, False), (7, 3, 33, True)])
def test_gemm_three_and_four_products_agree(torch_gpu, monkeypatch, L, R, M, odd):
    """The default GEMM (three real products per complex one) against the
    four-product kernel (JKPM_KUBO_GEMM=4) on the same stacks: both sum the
    same FP32 products in a different form, so they agree to FP32 rounding
    over N*R terms, 2e-6 of max |mu|; aligned and unaligned (odd K) rows."""
    if odd:
        H, _ = jcl.square_lattice(L, W=1.0, seed=L)
        A = jcl.ell_matrix(H)
        A.bounds = np.array([-4.5 * 1.01, 4.5 * 1.01])
        assert (A.shape[0] * R) % 2 == 1
        v = identity_velocity(A)
    else:
        A, _ = lattice(L, seed=L)
        v = jcl.modifier_velocity(A, "x")
    r = vectors(A.shape[0], R, 31)
    monkeypatch.delenv("JKPM_KUBO_GEMM", raising=False)
    mu3 = jcl.kpm_kubo_moments(A, M, r, v_data=v).cpu().numpy()
    monkeypatch.setenv("JKPM_KUBO_GEMM", "4")
    mu4 = jcl.kpm_kubo_moments(A, M, r, v_data=v).cpu().numpy()
    assert np.max(np.abs(mu3 - mu4)) / np.max(np.abs(mu4)) <= 2e-6


def test_mu_nm_independent_of_launch_widths(torch_gpu, monkeypatch):
    """The stack steps' work-group count (JKPM_KUBO_CHUNK_MULT) leaves every
    stack vector, and so mu_nm, bitwise unchanged; the GEMM's split-K slice
    count (JKPM_KUBO_SLICES) only regroups FP32 partial sums: 2e-6 of
    max |mu|."""
    A, _ = lattice(24, seed=24)
    r = vectors(A.shape[0], 4, 41)
    v = jcl.modifier_velocity(A, "x")
    for name in ("JKPM_KUBO_CHUNK_MULT", "JKPM_KUBO_SLICES", "JKPM_KUBO_GEMM"):
        monkeypatch.delenv(name, raising=False)
    ref = jcl.kpm_kubo_moments(A, 128, r, v_data=v).cpu().numpy()
    for mult in ("2", "4", "32"):
        monkeypatch.setenv("JKPM_KUBO_CHUNK_MULT", mult)
        assert np.array_equal(jcl.kpm_kubo_moments(A, 128, r, v_data=v).cpu().numpy(), ref), mult
    monkeypatch.delenv("JKPM_KUBO_CHUNK_MULT")
    for slices in ("8", "64", "1024"):
        monkeypatch.setenv("JKPM_KUBO_SLICES", slices)
        got = jcl.kpm_kubo_moments(A, 128, r, v_data=v).cpu().numpy()
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) <= 2e-6, slices


def _mu_nm_fp64(torch, A, B, v_data, r, M, chunk=1 << 17):
    """(1/R) sum_r <r| v T_m v T_n |r> from complex128 stacks of the complex128
    step path (jkpm_moments_keep; pinned to the oracle by the moment tests),
    Rk_n = v T_n r by jkpm_aAxby, contracted as a complex128 GEMM."""
    from minijclsquant_amd import _kpm

    N, K, R = A.shape[0], A.len_row, r.shape[0]
    H = nat.DeviceELL(A.data, A.indices, N, K, 1 / B, "c128")
    V = nat.DeviceELL.like(H, v_data, 1.0, "c128")
    _, T = H.moments_keep(_kpm.as_batch(r, N, "c128"), M)
    Rk = torch.zeros((M, N, R), dtype=torch.complex128, device="cuda")
    for n in range(M):
        V.aAxby(1.0, 0.0, T[n], Rk[n])
    del T
    _, L = H.moments_keep(Rk[0].clone(), M)
    L2, R2 = L[:M].reshape(M, N * R), Rk.reshape(M, N * R)
    mu = torch.zeros((M, M), dtype=torch.complex128, device="cuda")
    for k0 in range(0, N * R, chunk):
        k1 = min(N * R, k0 + chunk)
        mu += torch.mm(L2[:, k0:k1].conj().contiguous(),
                       R2[:, k0:k1].transpose(0, 1).contiguous())
    return mu / R


def test_fp64_reference_vs_oracle(torch_gpu):
    """The FP64 reference used at full size equals the oracle's mu_nm on a
    small lattice (so it is pinned to the same brute-force checks)."""
    A, B = lattice(8, seed=8)
    r = vectors(A.shape[0], 3, 61)
    v = jcl.modifier_velocity(A, "x")
    ref = _mu_nm_fp64(torch_gpu, A, B, v, r, 40).cpu().numpy()
    want = oracle_mu(A, B, v, r, 40)
    assert np.max(np.abs(ref - want)) <= 1e-12 * np.max(np.abs(want))


def test_kubo_workload_size_real_velocity_vs_fp64(torch_gpu):
    """The Kubo workload at full size with the real v_x (honeycomb 512 x 512
    cells, N = 524,288, M = 512, R = 4): complex64 stacks + FP32 MFMA mu_nm
    against complex128 stacks contracted in FP64, <= 2e-5 of max |mu|; the
    Kubo-Greenwood sigma on 512 energies <= 1e-4 of its peak."""
    torch = torch_gpu
    A, B = lattice(512, seed=2024)
    r = vectors(A.shape[0], 4, 700)
    M = 512
    mu = jcl.kpm_kubo_moments(A, M, r, "x")
    es = np.linspace(-0.99, 0.99, 512)
    sig = nat.kubo_sigma(mu, es, 0, B).cpu().numpy().real
    mu = mu.cpu().numpy()
    torch.cuda.empty_cache()
    ref = _mu_nm_fp64(torch, A, B, jcl.modifier_velocity(A, "x"), r, M)
    sig_ref = nat.kubo_sigma(ref, es, 0, B).cpu().numpy().real
    ref = ref.cpu().numpy()
    torch.cuda.empty_cache()
    err = np.max(np.abs(mu - ref)) / np.max(np.abs(ref))
    serr = np.max(np.abs(sig - sig_ref)) / np.max(np.abs(sig_ref))
    print(f"mu_nm err {err:.3e}, sigma err {serr:.3e}")
    assert err <= TOL_MU, err
    assert serr <= 1e-4, serr
