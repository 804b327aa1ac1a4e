"""The reference's namespace on the GPU: its DOS script's call sequence with
only the import changed, the Cython helpers the namespace re-exports, and the
Hermiticity guard of the forms that need A^H = A.

* dos_test_both.py:12-47 (lattice_hexagonal -> H_graphene 'ELL' -> kpm_dos /
  kpm_dos_gpu, normalised by N*B): both DOS tables equal the oracle's DOS of
  the oracle's moments on the same H and vector (<= 1e-10 of the peak), and
  integrate to 1 (the script's own printed check).
* blas_funs helpers against numpy and the oracle's restatements.
* Hermiticity: exactly 0 for the package's lattices (complex128 and
  complex64 handles, ragged padding included); the reference's 1e-6 cross
  term makes the doubling form, the Kubo moments and gradients refuse the
  matrix while the plain recursion still matches the oracle.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.integrate import simpson

import minijclsquant_amd as jcl
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_reference_dos_script_sequence(torch_gpu):
    """The call sequence of dos_test_both.py:12-47 at its size (N = 2^18,
    M = sqrt(N) = 512, honeycomb flake with mass 0.5, periodic, 'ELL')."""
    N = 2 ** (3 + 15)
    S = jcl.lattice_hexagonal(N)
    t = -2.7 + 0j
    m = 0.5 + 0j
    H = jcl.H_graphene(S, t, m, 0 + 0j, True, 'ELL')
    assert H.Omega > 0 and H.shape[0] == N
    random_vector = jcl.random_vector_generator(N)
    dos_cpu = jcl.kpm_dos(H, int(np.sqrt(N)), random_vector)
    dos_cpu[:, 1] = dos_cpu[:, 1] / (N * H.bounds[1])
    dos_gpu = jcl.kpm_dos_gpu(H, int(np.sqrt(N)), random_vector)
    dos_gpu[:, 1] = dos_gpu[:, 1] / (N * H.bounds[1])
    M = int(np.sqrt(N))
    B = H.bounds[1]
    mu = O.rec_A_tab(M, H.data * (1 / B), H.indices, H.len_row, random_vector)
    want = O.dos(mu, dos_cpu[:, 0], H.bounds[0], H.bounds[1], True) / (N * B)
    for got in (dos_cpu, dos_gpu):
        assert np.array_equal(got[:, 0], dos_cpu[:, 0])
        assert np.max(np.abs(got[:, 1] - want)) <= 1e-10 * np.max(np.abs(want))
        assert abs(simpson(got[:, 1], x=got[:, 0]) - 1.0) <= 1e-6
    assert abs(np.sum(dos_gpu[:, 1] - dos_cpu[:, 1])) <= 1e-12


def test_blas_funs_against_numpy_and_oracle(torch_gpu):
    rng = np.random.default_rng(3)
    n = 1000
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    y = rng.normal(size=n) + 1j * rng.normal(size=n)
    w = rng.normal(size=n) + 1j * rng.normal(size=n)
    a, b = 0.3 - 1.1j, -0.7 + 0.25j
    z = x.copy()
    jcl.scal(a, z, n, 8)
    ref = x.copy()
    O.scal(a, ref)
    assert np.max(np.abs(z - ref)) <= 1e-15 * np.max(np.abs(ref))
    y1 = y.copy()
    jcl.axpy(a, x, y1, n, 8)
    assert np.max(np.abs(y1 - (a * x + y))) <= 1e-14
    y2 = y.copy()
    jcl.axpy2(a, b, x, y2, n - 7, 8)  # only the first size elements change
    assert np.max(np.abs(y2[:n - 7] - (a * x + b * y)[:n - 7])) <= 1e-14
    assert np.array_equal(y2[n - 7:], y[n - 7:])
    out = np.zeros(n, complex)
    jcl.vec_mul(a, x, y, out, n, 8)
    assert np.max(np.abs(out - O.vec_mul(a, x, y))) <= 1e-14
    y3 = y.copy()
    jcl.vec_mul2(a, x, y3, n, 8)
    assert np.max(np.abs(y3 - a * x * y)) <= 1e-14
    jcl.vec_mul3(a, x, y, w, out, n, 8)
    assert np.max(np.abs(out - a * x * y * w)) <= 1e-13
    assert abs(jcl.vdot(a, x, y, n, 8) - O.vdot(a, x, y)) <= 1e-12 * n
    assert abs(jcl.dot(a, x, y, n, 8) - O.dot(a, x, y)) <= 1e-12 * n
    assert jcl.vdot(1, x, y, n, 8) == jcl.vdot(1, x, y, n, 8)  # fixed order
    assert jcl.vdot(1, x, y, 0, 8) == 0
    Hs, _ = jcl.square_lattice(9, W=1.0, seed=1)
    d, idx, _, K = O.csr_to_ell(Hs)
    N = Hs.shape[0]
    xx = rng.normal(size=N) + 1j * rng.normal(size=N)
    yy = rng.normal(size=N) + 1j * rng.normal(size=N)
    got = yy.copy()
    jcl.aAxby(a, b, xx, got, d, idx, K, N, 8)
    want = yy.copy()
    O.aAxby(a, b, xx, want, d, idx, K)
    assert np.max(np.abs(got - want)) <= 1e-13
    with pytest.raises(ValueError):
        jcl.scal(a, x.astype(np.complex64), n, 8)
    with pytest.raises(ValueError):
        jcl.axpy(a, x, y1, n + 1, 8)
    with pytest.raises(ValueError):
        jcl.aAxby(a, b, got, got, d, idx, K, N, 8)


def test_light_modifier_cython_modes(torch_gpu):
    """light_modifier_cython's four branches (Tp = 0 / != 0, eta = 0 / != 0)
    against the reference's elementwise factors, complex parameters kept."""
    H, Dx, Dy = jcl.honeycomb_lattice(6, W=1.0, seed=4)
    A = jcl.ell_matrix(H, Dx, Dy)
    n = A.data.size
    a, A0, t, w = 0.9 + 0.1j, 3.1 + 0j, 2.3 + 0j, 0.7 + 0j
    for Tp, eta in ((0, 0), (0, 1), (0, -1), (4.0, 0), (4.0, 1), (4.0, -1)):
        out = np.zeros(n, complex)
        jcl.light_modifier_cython(a, A0, t, w, Tp, eta, A.data, A.dx_vec, A.dy_vec, out, n, 8)
        env = 1.0 if Tp == 0 else np.cosh((t - 2 * Tp) / (0.5673 * Tp))
        if eta == 0:
            ph = A0 * np.sin(w * t) * A.dy_vec / env
        else:
            ph = A0 * (np.cos(w * t) * A.dx_vec + eta * np.sin(w * t) * A.dy_vec) / env
        want = a * A.data * np.exp(1j * ph)
        assert np.max(np.abs(out - want)) <= 1e-13, (Tp, eta)


def _hermitian_ragged(N=60, seed=0):
    rng = np.random.default_rng(seed)
    nnz = rng.integers(1, 5, N)
    rows = np.repeat(np.arange(N), nnz)
    cols = rng.integers(0, N, rows.size)
    X = sp.coo_matrix((rng.normal(size=rows.size) + 1j * rng.normal(size=rows.size),
                       (rows, cols)), shape=(N, N)).tocsr()
    H = (X + X.getH()).tocsr()
    H.sort_indices()
    return H


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_hermiticity_exact_for_hermitian_inputs(torch_gpu, dtype):
    ragged = _hermitian_ragged()
    assert np.any(jcl.ell_matrix(ragged).valid_row_values == 0)  # padded slots present
    for H in (jcl.honeycomb_lattice(8, W=1.0, seed=1)[0], jcl.square_lattice(7, W=2.0, seed=2)[0],
              ragged):
        A = jcl.ell_matrix(H)
        h = A.device(dtype, 1 / A.bounds[1])
        err, mag = h.hermiticity()
        assert err == 0.0 and mag > 0 and h.is_hermitian()


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_hermiticity_reports_nan_values(torch_gpu, dtype):
    """A NaN value makes the asymmetry NaN (not a maximum that skips it), so
    the doubling form refuses the matrix instead of trusting it."""
    H = jcl.square_lattice(7, W=2.0, seed=2)[0]
    A = jcl.ell_matrix(H)
    A.data[5] = np.nan
    h = A.device(dtype, 1 / A.bounds[1])
    err, mag = h.hermiticity()
    assert np.isnan(err) and not h.is_hermitian()
    r = jcl.random_vector_generator(A.shape[0], 1)
    with pytest.raises(ValueError, match="Hermitian"):
        jcl.rec_A_tab(16, A.modifier(jcl.modifier_bounds), r, doubling=True)


def test_hermiticity_guard_on_reference_graphene(torch_gpu):
    """H_graphene 'ELL' (jclsquant/hams.py:160-162) carries H += 1e-6 (Dx + Dy),
    antisymmetric in the bond: |A_ij - conj(A_ji)| = 2e-6 |dx_ij + dy_ij| / B."""
    torch = torch_gpu
    N = 1024
    A = jcl.H_graphene(jcl.lattice_hexagonal(N), -2.7 + 0j, 0.5 + 0j, 0, True, 'ELL')
    B = A.bounds[1]
    h = A.device("c128", 1 / B)
    err, mag = h.hermiticity()
    rows = np.arange(A.data.size) // A.len_row
    off = (rows != A.indices) & (A.valid_row_values == 1)
    # dx_vec holds Dx + 1e-6 H: 3e-5 relative to Dx, inside the tolerance below
    dsum = np.real(A.dx_vec[off] + A.dy_vec[off])
    want = np.max(np.abs(2e-6 * dsum)) / B
    assert abs(err - want) <= 1e-3 * want
    assert not h.is_hermitian()
    r = jcl.random_vector_generator(N, 3)
    # the plain recursion accepts the matrix and matches the oracle
    mu = jcl.rec_A_tab(64, A.modifier(jcl.modifier_bounds), r)
    ref = O.rec_A_tab(64, A.data * (1 / B), A.indices, A.len_row, r)
    assert np.max(np.abs(mu - ref)) <= 1e-10 * N
    with pytest.raises(ValueError, match="Hermitian"):
        jcl.rec_A_tab(64, A.modifier(jcl.modifier_bounds), r, doubling=True)
    with pytest.raises(ValueError, match="Hermitian"):
        jcl.kpm_dos(A, 64, r, doubling=True)
    with pytest.raises(ValueError, match="Hermitian"):
        jcl.kpm_kubo_moments(A, 16, r, "x")
    from minijclsquant_amd.autograd import chebyshev_moments

    vals = torch.from_numpy(A.data.copy()).cuda().requires_grad_(True)
    idx = torch.from_numpy(A.indices.copy()).cuda()
    rd = torch.from_numpy(r.reshape(N, 1).copy()).cuda()
    with pytest.raises(ValueError, match="Hermitian"):
        chebyshev_moments(vals, idx, rd, 16, 1 / B)
    with pytest.raises(ValueError, match="Hermitian"):
        h.grad_values(rd, 16, torch.zeros((16, 1), dtype=torch.complex128, device="cuda"))
    # a Hermitian lattice of the same geometry passes every form
    Hh, Dx, Dy = jcl.honeycomb_lattice(16, t=-2.7, W=1.0, seed=2)
    Ah = jcl.ell_matrix(Hh, Dx, Dy)
    rr = jcl.random_vector_generator(Hh.shape[0], 5)
    m1 = jcl.rec_A_tab(64, Ah.modifier(jcl.modifier_bounds), rr, doubling=True)
    m0 = jcl.rec_A_tab(64, Ah.modifier(jcl.modifier_bounds), rr)
    assert np.max(np.abs(m1 - m0)) <= 1e-10 * Hh.shape[0]


def test_dropin_dos_with_tuned_launch(torch_gpu, tmp_path, monkeypatch):
    """kpm_dos(tune=True) tunes this shape once into the table; tune=None
    (the default) then applies the table's launch shape and vectors per pass
    and gives bitwise the same DOS; tune=False runs the library defaults,
    equal to the complex64 bar."""
    monkeypatch.setenv("JKPM_TUNE_TABLE", str(tmp_path / "table.json"))
    monkeypatch.delenv("JKPM_LAUNCH", raising=False)
    H, Dx, Dy = jcl.honeycomb_lattice(48, t=-2.7, W=1.0, seed=9)
    A = jcl.ell_matrix(H, Dx, Dy)
    N = A.shape[0]
    r = np.stack([jcl.random_vector_generator(N, 600 + s) for s in range(16)])
    M = 96
    tuned = jcl.kpm_dos(A, M, r, dtype="c64", tune=True)
    assert (tmp_path / "table.json").exists()
    cached = jcl.kpm_dos(A, M, r, dtype="c64")
    assert np.array_equal(tuned, cached)
    plain = jcl.kpm_dos(A, M, r, dtype="c64", tune=False)
    assert np.max(np.abs(plain[:, 1] - tuned[:, 1])) <= 1e-5 * np.max(np.abs(plain[:, 1]))
