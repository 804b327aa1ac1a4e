"""Device-side CSR -> ELL ingest (jkpm_ell_create_csr) against the oracle.

The oracle's csr_to_ell restates data_indices_c_0
(jclsquant/cython_modules/Extra_funs.pyx:182-219) with the row width of
jclsquant/ell_matrix.py:187-189, and is pinned by its CSR round trip and the
padding rule.  Index and value layouts are integer / copy work, so the bar is
bit-exact: the handle's row-major columns equal the oracle's, its complex128
values equal the oracle's values times the scale divided back by it (the same
arithmetic as the row-major upload), and the moments equal those of the
handle built from the row-major ELL bit for bit.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import minijclsquant_amd as jcl
from minijclsquant_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def ragged_csr(n, seed, max_row=7, empty_every=0):
    """Random CSR with rows of 0..max_row entries in unsorted column order
    (CSR order is what the slots follow) and optional empty rows."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(1, min(max_row, n) + 1, size=n)
    if empty_every:
        counts[::empty_every] = 0
    indptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    nnz = int(indptr[-1])
    cols = np.empty(nnz, np.int32)
    for i in range(n):
        cols[indptr[i]:indptr[i + 1]] = rng.choice(n, size=counts[i], replace=False)
    data = rng.normal(size=nnz) + 1j * rng.normal(size=nnz)
    return indptr, cols, data


def as_csr(indptr, cols, data, n):
    # keeps the given in-row order (no canonicalisation)
    m = sp.csr_matrix((data, cols, indptr), shape=(n, n))
    m.has_sorted_indices = False
    return m


def host_values(h):
    return h.values().cpu().numpy()


def host_indices(h):
    return h.indices().cpu().numpy()


@pytest.mark.parametrize("n,seed,empty_every", [(1, 0, 0), (37, 1, 0), (1000, 2, 7),
                                                (4099, 3, 0)])
def test_ingest_matches_oracle_c128(torch_gpu, n, seed, empty_every):
    indptr, cols, data = ragged_csr(n, seed, empty_every=empty_every)
    print(f"ingest c128 n={n} seed={seed} empty_every={empty_every}", flush=True)
    d_ref, i_ref, _, K = O.csr_to_ell(as_csr(indptr, cols, data, n))
    scale = 1 / 4.7
    h = nat.DeviceELL.from_csr(indptr, cols, data, n, scale=scale, dtype="c128")
    assert (h.n, h.k) == (n, K)
    assert np.array_equal(host_indices(h), i_ref)
    ref = nat.DeviceELL(d_ref, i_ref, n, K, scale=scale, dtype="c128")
    assert np.array_equal(host_values(h), host_values(ref))
    # the unscaled values round-trip exactly
    h1 = nat.DeviceELL.from_csr(indptr, cols, data, n, scale=1.0, dtype="c128")
    assert np.array_equal(host_values(h1), d_ref)


def test_ingest_c64_and_device_source(torch_gpu):
    torch = torch_gpu
    n = 3001
    indptr, cols, data = ragged_csr(n, 5, empty_every=11)
    d_ref, i_ref, _, K = O.csr_to_ell(as_csr(indptr, cols, data, n))
    scale = 0.21
    h = nat.DeviceELL.from_csr(indptr, cols, data, n, scale=scale, dtype="c64")
    ref = nat.DeviceELL(d_ref, i_ref, n, K, scale=scale, dtype="c64")
    assert np.array_equal(host_indices(h), i_ref)
    assert np.array_equal(host_values(h), host_values(ref))
    hd = nat.DeviceELL.from_csr(torch.from_numpy(indptr).cuda(), torch.from_numpy(cols).cuda(),
                                torch.from_numpy(data).cuda(), n, scale=scale, dtype="c64")
    assert np.array_equal(host_indices(hd), i_ref)
    assert np.array_equal(host_values(hd), host_values(ref))


def test_ingest_explicit_width_and_empty_matrix(torch_gpu):
    n = 50
    indptr, cols, data = ragged_csr(n, 8, max_row=4)
    d_ref, i_ref, _, K = O.csr_to_ell(as_csr(indptr, cols, data, n), len_row=9)
    h = nat.DeviceELL.from_csr(indptr, cols, data, n, len_row=9)
    assert h.k == 9 == K
    assert np.array_equal(host_indices(h), i_ref)
    assert np.array_equal(host_values(h), d_ref)
    # every row empty: one padded slot per row, pointing at the row itself
    z = nat.DeviceELL.from_csr(np.zeros(n + 1, np.int32), np.zeros(0, np.int32),
                               np.zeros(0, np.complex128), n)
    assert z.k == 1
    assert np.array_equal(host_indices(z), np.arange(n, dtype=np.int32))
    assert not np.any(host_values(z))


def test_ingest_rejects_bad_input(torch_gpu):
    n = 20
    indptr, cols, data = ragged_csr(n, 9, max_row=5)
    longest = int(np.diff(indptr).max())
    with pytest.raises(ValueError, match="more than len_row"):
        nat.DeviceELL.from_csr(indptr, cols, data, n, len_row=longest - 1)
    bad = cols.copy()
    bad[3] = n
    with pytest.raises(ValueError, match="outside"):
        nat.DeviceELL.from_csr(indptr, bad, data, n)
    bad = cols.copy()
    bad[0] = -1
    with pytest.raises(ValueError, match="outside"):
        nat.DeviceELL.from_csr(indptr, bad, data, n)
    ip = indptr.copy()
    ip[5], ip[6] = ip[6], ip[5]          # decreasing range
    with pytest.raises(ValueError, match="indptr"):
        nat.DeviceELL.from_csr(ip, cols, data, n)


def test_ingest_moments_equal_row_major_upload(torch_gpu):
    """Honeycomb with bond disorder: the moments of the ingested handle equal
    those of the row-major upload bit for bit, and the oracle's to 1e-10."""
    torch = torch_gpu
    H, _, _ = jcl.honeycomb_lattice(24, t=-2.7, W=1.0, seed=4)
    A = jcl.ell_matrix(H)
    B = float(A.bounds[1])
    N, K = int(A.shape[0]), int(A.len_row)
    csr = A.A
    h = nat.DeviceELL.from_csr(csr.indptr, csr.indices, csr.data, N, scale=1 / B, dtype="c128")
    ref = A.device("c128", 1 / B)
    assert h.k == K
    R, M = 3, 160
    r = np.stack([jcl.random_vector_generator(N, 40 + s) for s in range(R)])
    vr = torch.from_numpy(r.T.copy()).cuda()
    mu = h.moments(vr, M).cpu().numpy()[0]
    mu_ref = ref.moments(vr, M).cpu().numpy()[0]
    assert np.array_equal(mu, mu_ref)
    Ht = A.modifier(jcl.modifier_bounds)
    for k in range(R):
        o = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
        assert np.max(np.abs(mu[:, k] - o)) / abs(o[0]) <= 1e-10


def test_batch_from_csr_equals_row_major(torch_gpu):
    """BatchKPM.from_csr (block-diagonal CSR ingested on the device) gives the
    same moments as BatchKPM over the row-major blocks."""
    rng = np.random.default_rng(12)
    mats, scales, vecs, blocks = [], [], [], []
    R, M = 2, 96
    for g, L in enumerate((5, 9, 7)):
        H, _ = jcl.square_lattice(L, t=-1.0, seed=100 + g, W=float(rng.uniform(0.5, 3)))
        A = jcl.ell_matrix(H)
        s = 1 / float(A.bounds[1])
        mats.append(A.A)
        scales.append(s)
        v = np.stack([jcl.random_vector_generator(L * L, 7 * g + k) for k in range(R)])
        vecs.append(v)
        blocks.append((A.data, A.indices, A.shape[0], A.len_row, s))
    ref = jcl.BatchKPM(blocks, vecs, M, dtype="c128").run().cpu().numpy()
    got = jcl.BatchKPM.from_csr(mats, scales, vecs, M, dtype="c128").run().cpu().numpy()
    assert np.array_equal(got, ref)
