"""Neighbour-exchange step kernels (launch variants 48..51) against their
gathering twins (variants 0, 2, 12, 14) and the oracle.

An NX kernel loads T_n of each lane group's own row and takes the T_n of a
slot whose column is that row, or the row of the lane group next to it, from
registers through a cross-lane permute; it gathers only the other slots.  The
per-row arithmetic and the partial-sum order are those of the twin, so every
comparison with the twin is bitwise.  The cases: lattices whose rows have
both near (i +- 1) and far (i +- stride) columns, a ragged non-Hermitian
matrix with random columns (near columns at random, padding slots), chunk
counts that leave partial wave passes, realisation batches whose stride
chunks are not multiples of a wave's rows, both sweep directions (every
recursion alternates them), both forms, both value types.  complex128
recursions are also checked against the oracle (<= 1e-10 of mu_0).
"""
import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _kpm
from minijclsquant_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TWINS = {48: 0, 49: 2, 50: 12, 51: 14}


def moments(h, vr, M, v, m, doubling=False):
    h.set_launch(v, m, doubling)
    return nat.moments_ex(h, None, vr, M, doubling).cpu().numpy()


def lattice(kind, L, seed):
    if kind == "honeycomb":
        H, Dx, Dy = jcl.honeycomb_lattice(L, t=-2.7, W=1.0, seed=seed)
        return jcl.ell_matrix(H, Dx, Dy)
    H, _ = jcl.square_lattice(L, W=1.0, seed=seed)
    return jcl.ell_matrix(H)


@pytest.mark.parametrize("kind,L", [("honeycomb", 24), ("square", 37)])
@pytest.mark.parametrize("dtype,R", [("c64", 16), ("c64", 8), ("c128", 4), ("c128", 8)])
def test_nx_equals_twin_on_lattices(torch_gpu, kind, L, dtype, R):
    A = lattice(kind, L, seed=5)
    Ht = A.modifier(jcl.modifier_bounds)
    N = A.shape[0]
    r = np.stack([jcl.random_vector_generator(N, s) for s in range(R)])
    h = nat.DeviceELL(Ht.data, Ht.indices, N, A.len_row, 1.0, dtype)
    vr = _kpm.as_batch(r, N, dtype)
    M = 131  # crosses the 64-step partial ring twice
    for m in (1, 4, 32):  # 32 work-groups per CU: chunks far shorter than a pass
        for v, base in TWINS.items():
            for dbl in (False, True):
                got = moments(h, vr, M, v, m, dbl)
                ref = moments(h, vr, M, base, m, dbl)
                assert np.array_equal(got, ref), (v, m, dbl)
    if dtype == "c128":
        mu = moments(h, vr, M, 51, 4)
        for k in range(R):
            o = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
            assert np.max(np.abs(mu[:, k] - o)) / abs(o[0]) <= 1e-10


def random_hermitian(N, K, seed):
    """Hermitian ELL of width K with random structure: a ring (i, i + 1),
    K - 3 random perfect matchings (the last on half the rows), and a
    diagonal on half the rows, so rows hold 2..K entries (padded the
    reference way) and near and far columns sit in any slot."""
    import scipy.sparse as sp

    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []

    def pair(p, q):
        z = complex(rng.normal(), rng.normal()) * 0.3
        rows.extend([p, q])
        cols.extend([q, p])
        vals.extend([z, np.conj(z)])

    for i in range(N):
        pair(i, (i + 1) % N)
    for k in range(K - 3):
        perm = rng.permutation(N)
        n_pairs = N // 2 if k < K - 4 else N // 4
        for t in range(n_pairs):
            pair(int(perm[2 * t]), int(perm[2 * t + 1]))
    for i in np.flatnonzero(rng.random(N) < 0.5):
        rows.append(int(i))
        cols.append(int(i))
        vals.append(complex(rng.normal(), 0.0))
    H = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    H.sum_duplicates()
    A = jcl.ell_matrix(H)
    assert A.len_row <= K
    return A


@pytest.mark.parametrize("dtype,R", [("c64", 16), ("c64", 8), ("c128", 4)])
@pytest.mark.parametrize("K", [4, 5])
def test_nx_random_structure(torch_gpu, dtype, R, K):
    """Random Hermitian structure, N = 3001 (no chunk is a multiple of a
    wave's rows)."""
    N = 3001
    A = random_hermitian(N, K, seed=100 + K)
    Ht = A.modifier(jcl.modifier_bounds)
    r = np.stack([jcl.random_vector_generator(N, 40 + s) for s in range(R)])
    h = nat.DeviceELL(Ht.data, Ht.indices, N, A.len_row, 1.0, dtype)
    vr = _kpm.as_batch(r, N, dtype)
    M = 70
    for m in (1, 8):
        for v, base in TWINS.items():
            for dbl in (False, True):
                got = moments(h, vr, M, v, m, dbl)
                assert np.array_equal(got, moments(h, vr, M, base, m, dbl)), (v, m, dbl)
    if dtype == "c128":
        mu = moments(h, vr, M, 48, 8)
        for k in range(R):
            o = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
            assert np.max(np.abs(mu[:, k] - o)) / abs(o[0]) <= 1e-10


@pytest.mark.parametrize("dtype,R", [("c64", 8), ("c128", 4)])
def test_nx_batch_stride_chunks(torch_gpu, dtype, R):
    """A realisation batch cut into one-stride chunks of 3..130 rows: most
    chunks end inside a wave's pass."""
    Ls, Ws = [5, 17, 9, 33, 3, 24, 130, 64], [0.5, 1.0, 3.0, 2.2, 1.1, 0.7, 1.5, 2.0]
    blocks, vecs = [], []
    for g, (L, W) in enumerate(zip(Ls, Ws)):
        d, i = jcl.square_lattice_ell(L, t=-1.0, W=W, seed=3 + g)
        blocks.append((d, i, L * L, 5, 1.0 / ((4 + W / 2) * 1.01)))
        vecs.append(np.stack([jcl.random_vector_generator(L * L, 1000 * g + s) for s in range(R)]))
    bk = jcl.BatchKPM(blocks, vecs, 97, dtype)
    assert bk.strides and any(bk.strides)
    for dbl in (False, True):
        _, m = bk.H.get_launch(dbl)
        out = {}
        for v in list(TWINS) + list(TWINS.values()):
            bk.H.set_launch(v, m, dbl)
            bk.doubling = dbl
            bk.ws = None  # the two forms need different workspaces
            out[v] = bk.run().cpu().numpy().copy()
        for v, base in TWINS.items():
            assert np.array_equal(out[v], out[base]), (v, dbl)
    if dtype == "c128":
        mu = out[48]  # doubling form last: compare it with the oracle's moments
        for g, (blk, vg) in enumerate(zip(blocks, vecs)):
            for k in range(R):
                o = O.rec_A_tab(97, blk[0] * blk[4], blk[1], 5, vg[k])
                assert np.max(np.abs(mu[g][:, k] - o)) / abs(o[0]) <= 1e-10
