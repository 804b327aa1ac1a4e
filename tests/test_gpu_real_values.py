"""Real-valued matrices: the step kernels read a packed copy of the real
parts (4 or 8 bytes per slot instead of 8 or 16).

* Detection: a handle is real when every stored imaginary part is exactly 0;
  set_values re-checks, a Peierls update makes the values complex.
* The real-value kernels against the complex-value kernels on the same
  matrix: the complex handle carries one imaginary part of 1e-300 (complex128)
  or 1e-30 (complex64), far below the rounding of any moment, so the two paths
  differ only by rounding: <= 1e-13 mu_0 (complex128), <= 1e-5 mu_0
  (complex64), every non-temporal-matrix-free variant, both forms.
* complex128 real-value moments against the oracle: <= 1e-10 mu_0.
"""
import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _kpm
from minijclsquant_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def lattice(kind, L, seed):
    if kind == "honeycomb":
        H, Dx, Dy = jcl.honeycomb_lattice(L, t=-2.7, W=1.0, seed=seed)
        return jcl.ell_matrix(H, Dx, Dy)
    H, _ = jcl.square_lattice(L, W=1.0, seed=seed)
    return jcl.ell_matrix(H)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_real_detection_and_transitions(torch_gpu, dtype):
    torch = torch_gpu
    A = lattice("honeycomb", 8, 1)
    s = 1.0 / A.bounds[1]
    h = nat.DeviceELL(A.data, A.indices, A.shape[0], A.len_row, s, dtype)
    assert h.is_real()
    cplx = A.data * np.exp(0.3j)
    h.set_values(cplx, s)
    assert not h.is_real()
    h.set_values(A.data.copy(), s)
    assert h.is_real()
    # a Peierls update (jkpm_ell_peierls) makes the destination complex
    bonds = nat.DeviceELL.like(h, (A.dx_vec + 1j * A.dy_vec).astype(np.complex128), 1.0, dtype)
    dst = nat.DeviceELL.like(h, A.data.copy(), s, dtype)
    assert dst.is_real() and not bonds.is_real()
    dst.peierls(h, bonds, 1.0, 0.2, 0.1)
    assert not dst.is_real()
    del torch


@pytest.mark.parametrize("kind,L", [("honeycomb", 24), ("square", 37)])
@pytest.mark.parametrize("dtype,R", [("c128", 4), ("c128", 8), ("c64", 16), ("c64", 8)])
def test_real_kernels_match_complex_kernels(torch_gpu, kind, L, dtype, R):
    A = lattice(kind, L, 5)
    Ht = A.modifier(jcl.modifier_bounds)
    N = A.shape[0]
    r = np.stack([jcl.random_vector_generator(N, s) for s in range(R)])
    vr = _kpm.as_batch(r, N, dtype)
    hr = nat.DeviceELL(Ht.data, Ht.indices, N, A.len_row, 1.0, dtype)
    tiny = Ht.data.copy()
    j = int(np.flatnonzero(tiny != 0)[0])
    tiny[j] += 1j * (1e-300 if dtype == "c128" else 1e-30)
    hc = nat.DeviceELL(tiny, Ht.indices, N, A.len_row, 1.0, dtype)
    assert hr.is_real() and not hc.is_real()
    M = 131
    tol = 1e-13 if dtype == "c128" else 1e-5
    for v in range(24):
        for dbl in (False, True):
            hr.set_launch(v, 4, dbl)
            hc.set_launch(v, 4, dbl)
            a = nat.moments_ex(hr, None, vr, M, dbl).cpu().numpy()
            b = nat.moments_ex(hc, None, vr, M, dbl).cpu().numpy()
            assert np.max(np.abs(a - b)) <= tol * N, (v, dbl)
    if dtype == "c128":
        hr.set_launch(14, 4)
        mu = hr.moments(vr, M)[0].cpu().numpy()
        for k in range(R):
            o = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
            assert np.max(np.abs(mu[:, k] - o)) / abs(o[0]) <= 1e-10


@pytest.mark.parametrize("dtype,R,tol", [("c128", 4, 1e-13), ("c64", 16, 1e-5)])
def test_use_real_values_switch(torch_gpu, dtype, R, tol):
    """jkpm_ell_use_real_values(h, 0) puts a real handle back on the complex
    kernels (the reference's complex128 storage); both give the same moments
    to rounding, and the switch leaves detection alone."""
    A = lattice("square", 29, 3)
    Ht = A.modifier(jcl.modifier_bounds)
    N = A.shape[0]
    r = np.stack([jcl.random_vector_generator(N, 70 + s) for s in range(R)])
    vr = _kpm.as_batch(r, N, dtype)
    h = nat.DeviceELL(Ht.data, Ht.indices, N, A.len_row, 1.0, dtype)
    for dbl in (False, True):
        h.use_real_values(True)
        a = nat.moments_ex(h, None, vr, 200, dbl).cpu().numpy()
        h.use_real_values(False)
        assert h.is_real()
        b = nat.moments_ex(h, None, vr, 200, dbl).cpu().numpy()
        assert np.max(np.abs(a - b)) <= tol * N, dbl
    h.use_real_values(True)


def test_stale_graph_after_values_turn_complex(torch_gpu):
    """A graph captured while the batch's values were real replays the
    real-value kernels; once set_values makes them complex, the packed copy
    holds NaN, so such a replay yields NaN moments (never the moments of
    stale real parts).  A new capture runs the complex kernels and equals
    the eager run."""
    torch = torch_gpu
    blocks, vecs = [], []
    for g, L in enumerate((9, 14)):
        d, i = jcl.square_lattice_ell(L, t=-1.0, W=1.0, seed=20 + g)
        blocks.append((d, i, L * L, 5, 1.0 / 4.6))
        vecs.append(np.stack([jcl.random_vector_generator(L * L, 50 + 10 * g + s) for s in range(4)]))
    bk = jcl.BatchKPM(blocks, vecs, 40, "c128")
    assert bk.H.is_real()
    eager = bk.run().clone()
    torch.cuda.synchronize()
    got = bk.replay().clone()
    torch.cuda.synchronize()
    assert torch.equal(got, eager)
    vals = bk.H.values() * complex(np.exp(0.4j))
    bk.H.set_values(vals)
    assert not bk.H.is_real()
    stale = bk.replay().clone()
    torch.cuda.synchronize()
    assert torch.isnan(stale[:, 1:].real).all()
    fresh_eager = bk.run().clone()
    torch.cuda.synchronize()
    assert torch.isfinite(fresh_eager.real).all()
    bk.capture()
    again = bk.replay().clone()
    torch.cuda.synchronize()
    assert torch.equal(again, fresh_eager)
