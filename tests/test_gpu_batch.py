"""Batches of disorder realizations and HIP graph replay.

* Each realization of a block-diagonal batch gives the moments of the
  unbatched path on the same matrix and vectors (same per-row arithmetic;
  only the FP64 order of the dot partials differs): <= 1e-12 relative.
* complex128 batches against the oracle: <= 1e-10.
* Graph replays equal the eager run bitwise, replay after replay, with eager
  library work and synchronisations in between.
* The batching workload at full size: 64 square lattices, L ~ U{128..1024},
  W ~ U[0.5, 3], t = -1, R = 8, M = 1024, complex64, one graph.
"""
import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def realizations(Ls, Ws, R, seed):
    blocks, vecs, mats = [], [], []
    for g, (L, W) in enumerate(zip(Ls, Ws)):
        d, i = jcl.square_lattice_ell(L, t=-1.0, W=W, seed=seed + g)
        B = (4 + W / 2) * 1.01
        blocks.append((d, i, L * L, 5, 1.0 / B))
        vecs.append(np.stack([jcl.random_vector_generator(L * L, 10_000 * g + s) for s in range(R)]))
    return blocks, vecs


def unbatched(block, v, M, dtype):
    d, i, n, k, s = block
    h = nat.DeviceELL(d, i, n, k, s, dtype)
    from minijclsquant_amd import _kpm

    return h.moments(_kpm.as_batch(v, n, dtype), M)[0].cpu().numpy()  # [M, R]


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_batch_equals_unbatched(torch_gpu, dtype):
    Ls, Ws = [5, 17, 9, 33, 3, 24], [0.5, 1.0, 3.0, 2.2, 1.1, 0.7]
    R, M = 4, 150
    blocks, vecs = realizations(Ls, Ws, R, 1)
    bk = jcl.BatchKPM(blocks, vecs, M, dtype)
    mu = bk.run().cpu().numpy()
    assert mu.shape == (len(Ls), M, R)
    for g, (blk, v) in enumerate(zip(blocks, vecs)):
        ref = unbatched(blk, v, M, dtype)
        err = np.max(np.abs(mu[g] - ref)) / np.max(np.abs(ref))
        assert err <= 1e-12, (g, err)
        if dtype == "c128":
            for r in range(R):
                o = O.rec_A_tab(M, blk[0] * blk[4], blk[1], 5, v[r])
                assert np.max(np.abs(mu[g][:, r] - o)) / abs(o[0]) <= 1e-10


def test_graph_replay_bitwise(torch_gpu):
    torch = torch_gpu
    Ls, Ws = [40, 12, 65], [1.0, 2.0, 0.5]
    blocks, vecs = realizations(Ls, Ws, 8, 7)
    bk = jcl.BatchKPM(blocks, vecs, 300, "c64")
    eager = bk.run().clone()
    torch.cuda.synchronize()
    for _ in range(3):
        got = bk.replay().clone()
        torch.cuda.synchronize()
        assert torch.equal(got, eager)
        # unrelated eager library work between replays
        jcl.rec_A_tab(20, jcl.ell_matrix(jcl.square_lattice(6, W=1.0, seed=1)[0]),
                      jcl.random_vector_generator(36, 1))
        torch.cuda.synchronize()


def test_kpm_dos_batch_matches_single(torch_gpu):
    mats = [jcl.ell_matrix(jcl.square_lattice(L, W=1.0, seed=L)[0]) for L in (8, 13)]
    vs = [jcl.random_vector_generator(A.shape[0], 3) for A in mats]
    out = jcl.kpm_dos_batch(mats, 64, vs, dtype="c128")
    for A, v, d in zip(mats, vs, out):
        ref = jcl.kpm_dos(A, 64, v)
        assert np.allclose(d[:, 0], ref[:, 0])
        assert np.max(np.abs(d[:, 1] - ref[:, 1])) / np.max(np.abs(ref[:, 1])) <= 1e-12


def test_batching_workload_full_size(torch_gpu):
    torch = torch_gpu
    rng = np.random.default_rng(64)
    G, R, M = 64, 8, 1024
    Ls = rng.integers(128, 1025, G)
    Ws = rng.uniform(0.5, 3.0, G)
    blocks, vecs = realizations(Ls, Ws, R, 500)
    bk = jcl.BatchKPM(blocks, vecs, M, "c64")
    mu = bk.replay().cpu().numpy()
    N = np.array([L * L for L in Ls])
    # mu_0 = sum |r|^2 = N_g per vector; Hermitian: real moments; |mu| <= mu_0
    assert np.max(np.abs(mu[:, 0, :].real - N[:, None]) / N[:, None]) <= 1e-6
    assert np.max(np.abs(mu.imag) / N[:, None, None]) <= 1e-5
    assert np.max(np.abs(mu) / N[:, None, None]) <= 1 + 1e-5
    # spot-check three realizations against the unbatched path
    for g in (int(np.argmin(Ls)), int(np.argmax(Ls)), 17):
        ref = unbatched(blocks[g], vecs[g], M, "c64")
        assert np.max(np.abs(mu[g] - ref)) / np.max(np.abs(ref)) <= 1e-12, g
    del torch


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_stride_chunks_equal_proportional_chunks(torch_gpu, dtype):
    """Chunks of one lattice stride (L rows) give the moments of the
    proportional cut up to the FP64 order of the dot partials, and the
    oracle's in complex128."""
    Ls, Ws = [70, 129, 100, 40], [0.5, 2.0, 1.3, 3.0]
    R, M = 3, 120
    blocks, vecs = realizations(Ls, Ws, R, 7)
    bs = jcl.BatchKPM(blocks, vecs, M, dtype)
    assert bs.strides == [70, 129, 100, 0]
    bp = jcl.BatchKPM(blocks, vecs, M, dtype, stride_chunks=False)
    ms, mp = bs.run().cpu().numpy(), bp.run().cpu().numpy()
    for g in range(len(Ls)):
        assert np.max(np.abs(ms[g] - mp[g])) / np.max(np.abs(mp[g])) <= 1e-12, g
    if dtype == "c128":
        for g in (0, 1):
            for r in range(R):
                o = O.rec_A_tab(M, blocks[g][0] * blocks[g][4], blocks[g][1], 5, vecs[g][r])
                assert np.max(np.abs(ms[g][:, r] - o)) / abs(o[0]) <= 1e-10
    # the doubling form on the stride tables as well
    bd = jcl.BatchKPM(blocks, vecs, M, dtype, doubling=True)
    md = bd.run().cpu().numpy()
    tol = 1e-10 if dtype == "c128" else 1e-5
    for g in range(len(Ls)):
        assert np.max(np.abs(md[g] - mp[g])) / np.max(np.abs(mp[g])) <= tol, g
