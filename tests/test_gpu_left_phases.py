"""Left vectors as 32-bit phases (complex64 plain recursions, one left set).

* jkpm_phase_pack: the phases match numpy's angle of the same complex64
  numbers (to one unit of 2 pi / 2^32), and the check accepts unit complex
  numbers, including the axis points, and rejects an element off the unit
  circle or a NaN.
* Moments with the phases against the oracle's complex128 moments
  (<= 1e-5 mu_0, the complex64 bar), and against the same handle with the
  phases turned off: equal to 1e-6 mu_0 but not bitwise, so the phase loads
  did run.
* Left vectors that are not unit complex numbers take the complex loads: with
  v_l = 2 v_r the moments are exactly twice those of v_l = v_r on the complex
  loads (a power-of-two scale is exact in every operation).
* The choice is made on the device: a captured batch graph, replayed after
  its vectors are doubled in place, gives exactly four times the complex-load
  moments of the original vectors.
"""
import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _kpm
from minijclsquant_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_phase_pack(torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(5)
    phi = rng.uniform(0, 2 * np.pi, 100_003)
    axes = np.array([0.0, 0.5 * np.pi, np.pi, 1.5 * np.pi, -0.5 * np.pi, 2 * np.pi - 1e-9])
    z = np.exp(1j * np.concatenate([phi, axes])).astype(np.complex64)
    v = torch.as_tensor(z, device="cuda")
    ph, ok = nat.phase_pack(v)
    assert ok
    got = ph.cpu().numpy().view(np.uint32).astype(np.int64)
    want = np.rint(np.angle(z.astype(np.complex128)) / (2 * np.pi) * 2.0**32).astype(np.int64)
    d = (got - want) % (1 << 32)
    assert np.all((d <= 1) | (d >= (1 << 32) - 1))
    for bad in (np.complex64(1.001 * np.exp(0.7j)), np.complex64(complex("nan+0j")),
                np.complex64(0.5), np.complex64(0)):
        zb = z.copy()
        zb[777] = bad
        _, ok = nat.phase_pack(torch.as_tensor(zb, device="cuda"))
        assert not ok, bad
    _, ok = nat.phase_pack(torch.empty(0, dtype=torch.complex64, device="cuda"))
    assert ok


def honeycomb(L, seed):
    H, Dx, Dy = jcl.honeycomb_lattice(L, t=-2.7, W=1.0, seed=seed)
    A = jcl.ell_matrix(H, Dx, Dy)
    return A.modifier(jcl.modifier_bounds)


@pytest.mark.parametrize("R", [8, 16, 5])
def test_moments_with_phases(torch_gpu, R):
    torch = torch_gpu
    Ht = honeycomb(24, 3)
    N = Ht.shape[0]
    r = np.stack([jcl.random_vector_generator(N, 40 + s) for s in range(R)])
    vr = _kpm.as_batch(r, N, "c64")
    h = nat.DeviceELL(Ht.data, Ht.indices, N, Ht.len_row, 1.0, "c64")
    M = 150
    mu_on = h.moments(vr, M)[0].cpu().numpy()
    for k in range(R):
        o = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
        assert np.max(np.abs(mu_on[:, k] - o)) <= 1e-5 * N, k
    h.use_left_phases(False)
    mu_off = h.moments(vr, M)[0].cpu().numpy()
    h.use_left_phases(True)
    assert np.max(np.abs(mu_on - mu_off)) <= 1e-6 * N
    assert np.any(mu_on != mu_off)
    # every launch variant of the plain form reads the phases the same way
    for v in range(nat.JKPM_STEP_VARIANTS):
        h.set_launch(v, 4)
        mu_v = h.moments(vr, M)[0].cpu().numpy()
        assert np.max(np.abs(mu_v - mu_on)) <= 1e-6 * N, v
    del torch


def test_non_unit_left_vectors_read_as_given(torch_gpu):
    torch = torch_gpu
    Ht = honeycomb(16, 8)
    N, R, M = Ht.shape[0], 8, 120
    r = np.stack([jcl.random_vector_generator(N, 90 + s) for s in range(R)])
    vr = _kpm.as_batch(r, N, "c64")
    h = nat.DeviceELL(Ht.data, Ht.indices, N, Ht.len_row, 1.0, "c64")
    h.use_left_phases(False)
    base = h.moments(vr, M)[0]
    h.use_left_phases(True)
    got = h.moments(vr, M, v_l0=2 * vr)[0]
    assert torch.equal(got, 2 * base)


def test_graph_follows_the_vectors(torch_gpu):
    torch = torch_gpu
    blocks, vecs = [], []
    for g, L in enumerate((12, 17)):
        d, i = jcl.square_lattice_ell(L, t=-1.0, W=1.0, seed=30 + g)
        blocks.append((d, i, L * L, 5, 1.0 / 4.6))
        vecs.append(np.stack([jcl.random_vector_generator(L * L, 7 * g + s) for s in range(8)]))
    bk = jcl.BatchKPM(blocks, vecs, 60, "c64")
    eager = bk.run().clone()
    torch.cuda.synchronize()
    got = bk.replay().clone()
    torch.cuda.synchronize()
    assert torch.equal(got, eager)
    bk.vr.mul_(2)
    doubled = bk.replay().clone()
    torch.cuda.synchronize()
    bk.vr.div_(2)
    bk.H.use_left_phases(False)
    ref = bk.run().clone()
    torch.cuda.synchronize()
    bk.H.use_left_phases(True)
    assert torch.equal(doubled, 4 * ref)
    assert not torch.equal(eager, ref)
    assert (eager - ref).abs().max().item() <= 1e-6 * max(b[2] for b in blocks)
