"""GPU parity of the wavefront path (two Chebyshev steps per launch; the
work-groups of one XCD hand T_{n+1} to each other through the L2 behind
progress flags, jkpm_wave_*) against the one-step path and the CPU oracle.

Tolerances: against the one-step kernel on the same handle (the same
complex64 arithmetic per row up to the compiler's choice of fused
multiply-adds, moments summed in another order) max_m |dmu_m| / mu_0 <= 1e-6;
against the complex128 oracle <= 1e-5, the bar of the one-step complex64
path.  The lattices use 4 work-groups per CU (1024 chunks on MI355X), the
largest grid whose work-groups are all resident at once.
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
    """max |mu - ref| / |mu_0| over [M, R] arrays."""
    return float(np.max(np.abs(mu - ref)) / np.max(np.abs(ref[0])))


def device_handle(A, B, doubling):
    N = A.shape[0]
    h = nat.DeviceELL(A.data, A.indices, N, A.len_row, 1 / B, "c64")
    v, _ = h.get_launch(doubling)
    h.set_launch(v, 4, doubling)
    return h


@pytest.mark.parametrize("doubling", [False, True])
@pytest.mark.parametrize("R", [16, 8])
@pytest.mark.parametrize("M", [2, 3, 4, 5, 6, 64, 129, 130, 259])
def test_wave_matches_one_step(torch_gpu, doubling, R, M):
    print(f"wave vs one-step: honeycomb 256, R={R}, M={M}, doubling={doubling}", flush=True)
    A, B = honeycomb(256, seed=M)
    N = A.shape[0]
    h = device_handle(A, B, doubling)
    r = vectors(N, R, 11 * M)
    vr = _kpm.as_batch(r, N, "c64")
    plan = nat.WavePlan(h, R, doubling)
    info = plan.info()
    assert info["lag"] >= 2 and info["fixup_rows"] > 0
    got = plan.moments(vr, M).cpu().numpy()
    ref = nat.moments_ex(h, None, vr, M, doubling).cpu().numpy()
    assert rel(got, ref) <= 1e-6
    # a second call (next flag epoch) gives the same moments bit for bit
    again = plan.moments(vr, M).cpu().numpy()
    assert np.array_equal(got, again)


@pytest.mark.parametrize("doubling", [False, True])
def test_wave_vs_oracle(torch_gpu, doubling):
    M, R = 200, 16
    print(f"wave vs oracle: honeycomb 256, R={R}, M={M}, doubling={doubling}", flush=True)
    A, B = honeycomb(256, seed=5)
    N = A.shape[0]
    h = device_handle(A, B, doubling)
    r = vectors(N, R, 300)
    vr = _kpm.as_batch(r, N, "c64")
    got = nat.WavePlan(h, R, doubling).moments(vr, M).cpu().numpy()
    for k in (0, 9):
        want = O.rec_A_tab(M, A.data / B, A.indices, A.len_row, r[k])
        assert np.max(np.abs(got[:, k] - want)) / abs(want[0]) <= 1e-5


@pytest.mark.parametrize("doubling", [False, True])
def test_wave_headline_lattice(torch_gpu, doubling):
    """Honeycomb 1024 x 1024 cells (N = 2,097,152), R = 16: one chunk per
    lattice row of cells; the periodic wrap of each lattice row is finished
    after the main loop (1016 tail rows), the rows next to the XCD boundaries
    by the fix-up launch (16384 rows)."""
    M, R = 256, 16
    print(f"wave headline lattice: R={R}, M={M}, doubling={doubling}", flush=True)
    A, B = honeycomb(1024, seed=2024)
    N = A.shape[0]
    h = device_handle(A, B, doubling)
    r = vectors(N, R, 1000)
    vr = _kpm.as_batch(r, N, "c64")
    plan = nat.WavePlan(h, R, doubling)
    info = plan.info()
    assert info["nchunks"] == 1024 and info["lag"] >= 2
    assert info["tail_rows"] == 1016 and info["fixup_rows"] == 16384
    got = plan.moments(vr, M).cpu().numpy()
    ref = nat.moments_ex(h, None, vr, M, doubling).cpu().numpy()
    assert rel(got, ref) <= 1e-6


def test_wave_refusals(torch_gpu):
    A, B = honeycomb(64, seed=1)
    N = A.shape[0]
    h64 = nat.DeviceELL(A.data, A.indices, N, A.len_row, 1 / B, "c64")
    with pytest.raises(ValueError, match="R = 8 or 16"):
        nat.WavePlan(h64, 4)
    h128 = nat.DeviceELL(A.data, A.indices, N, A.len_row, 1 / B, "c128")
    with pytest.raises(ValueError, match="complex64"):
        nat.WavePlan(h128, 16)
    # random columns: every block reads far more than four chunks
    N2, K = 131072, 4
    rng = np.random.default_rng(3)
    cols = rng.integers(0, N2, size=(N2, K)).astype(np.int32)
    vals = (rng.standard_normal((N2, K)) + 0j).astype(np.complex128)
    hr = nat.DeviceELL(vals.ravel(), cols.ravel(), N2, K, 0.1, "c64")
    v, _ = hr.get_launch(False)
    hr.set_launch(v, 4, False)
    with pytest.raises(ValueError, match="wavefront plan"):
        nat.WavePlan(hr, 16)
