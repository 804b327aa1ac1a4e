"""GPU parity of the two-steps-per-launch path (rows renumbered into LDS
patches, jkpm_tiles_*) against the one-step path and the CPU oracle.

Tolerances: complex128 max_m |dmu_m| / mu_0 <= 1e-10 against the oracle and
<= 1e-12 against the one-step kernel (same arithmetic per row, moments summed
in another order); complex64 (FP64 dots) <= 1e-5 against the oracle; the
doubling form as the plain one.  Small max_rows force many patches with
halos on small lattices.
"""
import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _kpm
from minijclsquant_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def honeycomb(L, W=1.0, seed=0):
    H, Dx, Dy = jcl.honeycomb_lattice(L, t=-2.7, W=W, seed=seed)
    A = jcl.ell_matrix(H, Dx, Dy)
    B = (3 * 2.7 + W / 2) * 1.01
    return A, B


def vectors(N, R, seed0=0):
    return np.stack([jcl.random_vector_generator(N, seed0 + s) for s in range(R)])


def rel(mu, ref):
    return float(np.max(np.abs(mu - ref)) / np.max(np.abs(ref[..., 0])))


@pytest.mark.parametrize("M", [1, 2, 3, 4, 5, 64, 131])
@pytest.mark.parametrize("max_rows", [16, 40, 0])
def test_tiles_match_one_step_c128(torch_gpu, M, max_rows):
    A, B = honeycomb(10, seed=M)
    N = A.shape[0]
    h = nat.DeviceELL(A.data, A.indices, N, A.len_row, 1 / B, "c128")
    r = vectors(N, 3, 7 * M)
    vr = _kpm.as_batch(r, N, "c128")
    plan = nat.TilePlan(h, 3, max_rows)
    info = plan.info()
    if max_rows:
        assert info["max_ext"] <= max_rows and info["npatch"] > 1 and info["halo_rows"] > 0
    got = plan.moments(vr, M).cpu().numpy()  # [M, R]
    ref = h.moments(vr, M)[0].cpu().numpy()
    assert rel(got.T, ref.T) <= 1e-12
    want = O.rec_A_tab(M, A.data / B, A.indices, A.len_row, r[1])
    assert np.max(np.abs(got[:, 1] - want)) / abs(want[0]) <= 1e-10


@pytest.mark.parametrize("R", [1, 5, 16, 20])
def test_tiles_c64_vs_oracle(torch_gpu, R):
    A, B = honeycomb(12, seed=R)
    N, M = A.shape[0], 200
    h = nat.DeviceELL(A.data, A.indices, N, A.len_row, 1 / B, "c64")
    r = vectors(N, R, 100)
    vr = _kpm.as_batch(r, N, "c64")
    plan = nat.TilePlan(h, R, 48)
    got = plan.moments(vr, M).cpu().numpy()
    for k in (0, R - 1):
        want = O.rec_A_tab(M, A.data / B, A.indices, A.len_row, r[k])
        assert np.max(np.abs(got[:, k] - want)) / abs(want[0]) <= 1e-5


@pytest.mark.parametrize("M", [1, 2, 3, 7, 64, 200])
def test_tiles_doubling_c128(torch_gpu, M):
    A, B = honeycomb(9, seed=M + 1)
    N = A.shape[0]
    h = nat.DeviceELL(A.data, A.indices, N, A.len_row, 1 / B, "c128")
    r = vectors(N, 4, 3)
    vr = _kpm.as_batch(r, N, "c128")
    plan = nat.TilePlan(h, 4, 24)
    got = plan.moments(vr, M, doubling=True).cpu().numpy()
    for k in range(4):
        want = O.rec_A_tab(M, A.data / B, A.indices, A.len_row, r[k])
        assert np.max(np.abs(got[:, k] - want)) / abs(want[0]) <= 1e-10, k


def test_tiles_square_k5_and_nonhermitian(torch_gpu):
    """K = 5 and a ragged non-Hermitian ELL (the plain form needs no symmetry)."""
    H, _ = jcl.square_lattice(14, W=1.0, seed=2)
    A = jcl.ell_matrix(H)
    B = 4.5 * 1.01
    N = A.shape[0]
    h = nat.DeviceELL(A.data, A.indices, N, A.len_row, 1 / B, "c128")
    r = vectors(N, 2, 9)
    plan = nat.TilePlan(h, 2, 30)
    got = plan.moments(_kpm.as_batch(r, N, "c128"), 90).cpu().numpy()
    want = O.rec_A_tab(90, A.data / B, A.indices, A.len_row, r[0])
    assert np.max(np.abs(got[:, 0] - want)) / abs(want[0]) <= 1e-10
    rng = np.random.default_rng(5)
    n, k = 150, 6
    idx = rng.integers(0, n, size=(n, k)).astype(np.int32)
    val = 0.1 * (rng.normal(size=n * k) + 1j * rng.normal(size=n * k))
    h2 = nat.DeviceELL(val, idx.ravel(), n, k, 1.0, "c128")
    v = vectors(n, 1, 4)
    plan2 = nat.TilePlan(h2, 1, 20)
    got2 = plan2.moments(_kpm.as_batch(v, n, "c128"), 40).cpu().numpy()[:, 0]
    want2 = O.rec_A_tab(40, val, idx.ravel(), k, v[0])
    assert np.max(np.abs(got2 - want2)) / np.max(np.abs(want2)) <= 1e-10


def test_tiles_refresh_after_new_values(torch_gpu):
    A, B = honeycomb(8, seed=1)
    N = A.shape[0]
    h = nat.DeviceELL(A.data, A.indices, N, A.len_row, 1 / B, "c128")
    r = vectors(N, 1, 2)
    vr = _kpm.as_batch(r, N, "c128")
    plan = nat.TilePlan(h, 1, 32)
    A2, _ = honeycomb(8, W=2.0, seed=5)
    h.set_values(A2.data)
    plan.refresh()
    got = plan.moments(vr, 50).cpu().numpy()[:, 0]
    want = O.rec_A_tab(50, A2.data / B, A.indices, A.len_row, r[0])
    assert np.max(np.abs(got - want)) / abs(want[0]) <= 1e-10


def test_tiles_full_size_precision_form(torch_gpu):
    """Honeycomb 1024 x 1024 cells, K = 4, M = 2048, R = 16, complex64 with
    FP64 dots, the LDS-sized plan: moments agree with the one-step kernel
    and the DOS with the complex128 path to 1e-5."""
    torch = torch_gpu
    A, B = honeycomb(1024, seed=2024)
    N, M, R = A.shape[0], 2048, 16
    r = vectors(N, R, 1000)
    h = nat.DeviceELL(A.data, A.indices, N, A.len_row, 1 / B, "c64")
    vr = _kpm.as_batch(r, N, "c64")
    plan = nat.TilePlan(h, R)
    info = plan.info()
    assert info["max_ext"] <= 512 and info["halo_rows"] < 0.25 * N
    mu_t = plan.moments(vr, M)
    mu_1 = h.moments(vr, M)[0]
    d = (mu_t - mu_1).abs().max().item() / mu_1[0].abs().max().item()
    assert d <= 1e-5, d
    h128 = nat.DeviceELL(A.data, A.indices, N, A.len_row, 1 / B, "c128")
    mu128 = h128.moments(_kpm.as_batch(r, N, "c128"), M)[0]
    es = torch.as_tensor(np.linspace(-0.99, 0.99, 4096), device="cuda")
    d64 = nat.dos(mu_t, es).cpu().numpy()
    d128 = nat.dos(mu128, es).cpu().numpy()
    assert np.max(np.abs(d64 - d128)) / np.max(np.abs(d128)) <= 1e-5
    # the doubling form through the same plan
    mu_d = plan.moments(vr, M, doubling=True)
    dd = nat.dos(mu_d, es).cpu().numpy()
    assert np.max(np.abs(dd - d128)) / np.max(np.abs(d128)) <= 1e-5
