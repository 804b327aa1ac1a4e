"""Vectors per recursion pass (jkpm_ell_set_pass): a call with R vectors run
as passes of g < R vectors (the last pass takes the rest) gives the moments
of the single pass.

* complex128, plain form, one and two left sets, R = 10 in passes of 4
  (4, 4, 2) and 3: against the single pass <= 1e-12 mu_0 (the passes run
  other lane shapes, so the FP64 partials are summed in another order) and
  against the oracle <= 1e-10 mu_0.
* Doubling form, complex128 and complex64: the same bars, complex64 against
  its own single pass.  Passes of an even number of vectors keep the
  16-byte-lane kernels of the single pass, so a vector's complex64 recursion
  is the same and only the FP64 sums move (1e-12 mu_0); passes of 3 or 1
  vectors run the 8-byte-lane kernels, whose complex64 rounding differs
  (the complex64 bar, 1e-5 mu_0).
* complex64 with the left phases, R = 16 in passes of 8, 4 and 5: the same
  two bars.
* A realisation batch with passes, the no-allocation entry point with
  passes (bitwise equal to jkpm_moments with the same setting), the
  moments-with-stack form (runs at once), and argument errors.
"""
import ctypes

import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _kpm
from minijclsquant_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def square(L, seed):
    H, _ = jcl.square_lattice(L, W=1.0, seed=seed)
    A = jcl.ell_matrix(H)
    return A.modifier(jcl.modifier_bounds)


def vectors(N, R, seed):
    return np.stack([jcl.random_vector_generator(N, seed + s) for s in range(R)])


@pytest.mark.parametrize("g", [4, 3])
def test_plain_passes_c128(torch_gpu, g):
    torch = torch_gpu
    Ht = square(21, 4)
    N, R, M = Ht.shape[0], 10, 97
    r = vectors(N, R, 100)
    vr = _kpm.as_batch(r, N, "c128")
    vl1 = torch.flip(vr, dims=[0]).contiguous()
    h = nat.DeviceELL(Ht.data, Ht.indices, N, Ht.len_row, 1.0, "c128")
    one = h.moments(vr, M, v_l1=vl1).cpu().numpy()
    ws_one = h.workspace_bytes(R, 2)
    h.set_pass(g)
    assert h.get_pass() == g and h.get_pass(doubling=True) == 0
    assert h.workspace_bytes(R, 2) >= ws_one
    got = h.moments(vr, M, v_l1=vl1).cpu().numpy()
    assert np.max(np.abs(got - one)) <= 1e-12 * N
    for k in range(R):
        o = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
        assert np.max(np.abs(got[0, :, k] - o)) <= 1e-10 * N, k
    # a left set other than v_r is copied per pass too
    vl0 = (2 * vr).contiguous()
    got2 = h.moments(vr, M, v_l0=vl0)[0]
    h.set_pass(0)
    want2 = h.moments(vr, M, v_l0=vl0)[0]
    assert (got2 - want2).abs().max().item() <= 1e-12 * N


def bar(dtype, g):
    return 1e-12 if dtype == "c128" or g % 2 == 0 else 1e-5


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_doubling_passes(torch_gpu, dtype):
    Ht = square(24, 9)
    N, R, M = Ht.shape[0], 8, 120
    r = vectors(N, R, 300)
    vr = _kpm.as_batch(r, N, dtype)
    h = nat.DeviceELL(Ht.data, Ht.indices, N, Ht.len_row, 1.0, dtype)
    one = nat.moments_ex(h, None, vr, M, True).cpu().numpy()
    for g in (4, 3, 1):
        h.set_pass(g, doubling=True)
        got = nat.moments_ex(h, None, vr, M, True).cpu().numpy()
        assert np.max(np.abs(got - one)) <= bar(dtype, g) * N, g
    if dtype == "c128":
        for k in range(R):
            o = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
            assert np.max(np.abs(got[:, k] - o)) <= 1e-10 * N, k


def test_c64_phases_in_passes(torch_gpu):
    H, Dx, Dy = jcl.honeycomb_lattice(20, t=-2.7, W=1.0, seed=2)
    Ht = jcl.ell_matrix(H, Dx, Dy).modifier(jcl.modifier_bounds)
    N, R, M = Ht.shape[0], 16, 150
    vr = _kpm.as_batch(vectors(N, R, 500), N, "c64")
    h = nat.DeviceELL(Ht.data, Ht.indices, N, Ht.len_row, 1.0, "c64")
    one = h.moments(vr, M)[0].cpu().numpy()
    for g in (8, 4, 5):
        h.set_pass(g)
        got = h.moments(vr, M)[0].cpu().numpy()
        assert np.max(np.abs(got - one)) <= bar("c64", g) * N, g


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("doubling", [False, True])
def test_batch_passes(torch_gpu, doubling, dtype):
    """A realisation batch in passes of 4 vectors, eager and replayed from a
    captured graph (the pass column copies, and for complex64 the left-phase
    packing, are captured too), against the batch in one pass."""
    torch = torch_gpu
    blocks, vecs = [], []
    for g, L in enumerate((10, 15, 7)):
        d, i = jcl.square_lattice_ell(L, t=-1.0, W=1.0, seed=60 + g)
        blocks.append((d, i, L * L, 5, 1.0 / 4.6))
        vecs.append(vectors(L * L, 8, 20 * g))
    bk = jcl.BatchKPM(blocks, vecs, 50, dtype, doubling=doubling)
    one = bk.run().clone()
    bk.H.set_pass(4, doubling=doubling)
    bk.ws = None  # the workspace size follows the setting
    got = bk.run().clone()
    torch.cuda.synchronize()
    assert (got - one).abs().max().item() <= 1e-12 * 225
    for _ in range(3):  # replays with eager work between them
        rep = bk.replay().clone()
        torch.cuda.synchronize()
        assert torch.equal(rep, got)
        bk.run()
        torch.cuda.synchronize()


def test_noalloc_and_stack_with_passes(torch_gpu):
    torch = torch_gpu
    Ht = square(19, 12)
    N, R, M = Ht.shape[0], 12, 64
    vr = _kpm.as_batch(vectors(N, R, 700), N, "c64")
    h = nat.DeviceELL(Ht.data, Ht.indices, N, Ht.len_row, 1.0, "c64")
    stack_one, _ = h.moments_keep(vr, M)
    h.set_pass(4)
    want = h.moments(vr, M)
    mu = torch.empty((1, M, R), dtype=torch.complex128, device="cuda")
    rc = nat.lib().jkpm_moments_noalloc(h.handle, vr.data_ptr(), vr.data_ptr(), None, R, M,
                                        mu.data_ptr(), nat.stream_ptr(None))
    assert rc == nat.JKPM_OK, nat.lib().jkpm_last_error()
    torch.cuda.synchronize()
    assert torch.equal(mu, want)
    stack_g, _ = h.moments_keep(vr, M)
    assert torch.equal(stack_g, stack_one)
    with pytest.raises(ValueError):
        h.set_pass(-1)
    need = ctypes.c_size_t()
    assert nat.lib().jkpm_ell_set_pass(h.handle, 2, 4) == nat.JKPM_ERR_ARG
    assert nat.lib().jkpm_moments_workspace(h.handle, R, 1, ctypes.byref(need)) == nat.JKPM_OK


@pytest.mark.parametrize("L,M", [(4, 1), (4, 2), (6, 3), (3, 5)])
@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_short_recursions_and_tiny_lattices_in_passes(torch_gpu, L, M, dtype):
    """M = 1..5 on lattices of 9..36 rows: passes of 4 (and the doubling form)
    give the single pass's moments, and complex128 gives the oracle's."""
    Ht = square(L, 30 + L)
    N, R = Ht.shape[0], 8
    r = vectors(N, R, 900 + M)
    vr = _kpm.as_batch(r, N, dtype)
    h = nat.DeviceELL(Ht.data, Ht.indices, N, Ht.len_row, 1.0, dtype)
    one = h.moments(vr, M)[0].cpu().numpy()
    one_d = nat.moments_ex(h, None, vr, M, True).cpu().numpy()
    h.set_pass(4)
    h.set_pass(4, doubling=True)
    got = h.moments(vr, M)[0].cpu().numpy()
    got_d = nat.moments_ex(h, None, vr, M, True).cpu().numpy()
    assert np.max(np.abs(got - one)) <= 1e-12 * N
    assert np.max(np.abs(got_d - one_d)) <= 1e-12 * N
    if dtype == "c128":
        for k in range(R):
            o = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
            assert np.max(np.abs(got[:, k] - o)) <= 1e-10 * N, k
            assert np.max(np.abs(got_d[:, k] - o)) <= 1e-10 * N, k
