"""The doubling form (two moments per Chebyshev step, Hermitian matrices).

mu_2k = 2<T_k|T_k> - mu_0 and mu_2k+1 = 2<T_k|T_k+1> - mu_1 reproduce the
plain recursion's moments: complex128 against the oracle <= 1e-10 * mu_0,
complex64 <= 1e-5 * mu_0; batches against the unbatched doubling path;
the full-size precision workload's DOS against complex128 <= 1e-5.
"""
import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _kpm, _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def herm(L, W=1.0, seed=0):
    H, Dx, Dy = jcl.honeycomb_lattice(L, t=-2.7, W=W, seed=seed)
    return jcl.ell_matrix(H)


@pytest.mark.parametrize("M", [1, 2, 3, 4, 5, 63, 64, 65, 130, 257])
@pytest.mark.parametrize("R", [1, 3, 16])
def test_doubling_vs_oracle_c128(torch_gpu, M, R):
    A = herm(10, seed=M + R)
    Ht = A.modifier(jcl.modifier_bounds)
    r = np.stack([jcl.random_vector_generator(A.shape[0], 40 + s) for s in range(R)])
    mu = jcl.rec_A_tab(M, Ht, r, doubling=True)
    for k in range(R):
        ref = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
        assert np.max(np.abs(mu[k] - ref)) / abs(ref[0]) <= 1e-10, (M, R, k)


def test_doubling_c64_and_square_lattice(torch_gpu):
    H, _ = jcl.square_lattice(64, W=2.0, seed=1)
    A = jcl.ell_matrix(H)
    Ht = A.modifier(jcl.modifier_bounds)
    r = np.stack([jcl.random_vector_generator(A.shape[0], s) for s in range(8)])
    mu = jcl.rec_A_tab(600, Ht, r, dtype="c64", doubling=True)
    for k in range(8):
        ref = O.rec_A_tab(600, Ht.data, Ht.indices, Ht.len_row, r[k])
        assert np.max(np.abs(mu[k] - ref)) / abs(ref[0]) <= 1e-5


def test_doubling_rejects_two_vectors(torch_gpu):
    A = herm(4)
    v = jcl.random_vector_generator(A.shape[0], 1)
    with pytest.raises(ValueError):
        from minijclsquant_amd import _kpm as k

        k.moments(A, 10, v, v_l=v, doubling=True)


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_doubling_batch_matches_single(torch_gpu, dtype):
    Ls, Ws = [5, 17, 9, 33], [0.5, 1.0, 3.0, 2.2]
    blocks, vecs = [], []
    for g, (L, W) in enumerate(zip(Ls, Ws)):
        d, i = jcl.square_lattice_ell(L, W=W, seed=g)
        B = (4 + W / 2) * 1.01
        blocks.append((d, i, L * L, 5, 1.0 / B))
        vecs.append(np.stack([jcl.random_vector_generator(L * L, 100 * g + s) for s in range(4)]))
    bk = jcl.BatchKPM(blocks, vecs, 201, dtype, doubling=True)
    mu = bk.replay().cpu().numpy()
    for g, (blk, v) in enumerate(zip(blocks, vecs)):
        h = nat.DeviceELL(blk[0], blk[1], blk[2], 5, blk[4], dtype)
        ref = nat.moments_ex(h, None, _kpm.as_batch(v, blk[2], dtype), 201, True).cpu().numpy()
        assert np.max(np.abs(mu[g] - ref)) / np.max(np.abs(ref)) <= 1e-12
        if dtype == "c128":
            o = O.rec_A_tab(201, blk[0] * blk[4], blk[1], 5, v[0])
            assert np.max(np.abs(mu[g][:, 0] - o)) / abs(o[0]) <= 1e-10


def test_doubling_precision_form_full_size(torch_gpu):
    """Honeycomb 1024 x 1024 cells, M = 2048, R = 16, complex64 doubling vs
    the complex128 plain recursion: DOS error on 4096 energies <= 1e-5."""
    A = herm(1024, W=1.0, seed=2024)
    B = (3 * 2.7 + 0.5) * 1.01
    A.bounds = np.array([-B, B])
    N, M, R = A.shape[0], 2048, 16
    r = np.stack([jcl.random_vector_generator(N, 1000 + s) for s in range(R)])
    E = np.linspace(-B + 1e-3, B - 1e-3, 4096)
    d64 = jcl.kpm_dos(A, M, r, energies=E, dtype="c64", doubling=True)[:, 1]
    d128 = jcl.kpm_dos(A, M, r, energies=E, dtype="c128")[:, 1]
    assert np.max(np.abs(d64 - d128)) / np.max(np.abs(d128)) <= 1e-5
