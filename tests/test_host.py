"""Host logic of the package (CPU only): ELL conversion, bounds, lattice
builders, modifiers, kernels -- checked against the oracle and closed forms."""
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

import minijclsquant_amd as jcl
from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent


def ragged_csr(N=80, seed=0, kmax=7):
    rng = np.random.default_rng(seed)
    nnz = rng.integers(1, kmax, N)
    rows = np.repeat(np.arange(N), nnz)
    cols = rng.integers(0, N, rows.size)
    H = sp.coo_matrix((rng.normal(size=rows.size) + 1j * rng.normal(size=rows.size),
                       (rows, cols)), shape=(N, N)).tocsr()
    H.sum_duplicates()
    return H


@pytest.mark.parametrize("maker", [
    lambda: jcl.square_lattice(9, W=1.0, seed=1)[0],
    lambda: jcl.honeycomb_lattice(5, W=1.0, seed=2)[0],
    lambda: ragged_csr(),
])
def test_ell_conversion_matches_oracle(maker):
    H = maker()
    A = jcl.ell_matrix(H)
    d, i, v, K = O.csr_to_ell(H)
    assert A.len_row == K
    assert np.array_equal(A.data, d)
    assert np.array_equal(A.indices, i)
    assert np.array_equal(A.valid_row_values, v)
    assert A.indices.dtype == np.int32 and A.data.dtype == np.complex128
    back = jcl.ell_to_csr(A)
    assert abs(back - H).max() == 0


def test_ell_with_bond_vectors():
    H, Dx, Dy = jcl.honeycomb_lattice(4, W=0.5, seed=0)
    A = jcl.ell_matrix(H, Dx, Dy)
    # dx_vec/dy_vec sit in the same slots as data; padding is zero
    for M, vec in ((Dx, A.dx_vec), (Dy, A.dy_vec)):
        rows = np.arange(A.data.size) // A.len_row
        dense = M.toarray()
        ok = A.valid_row_values == 1
        assert np.allclose(vec[ok], dense[rows[ok], A.indices[ok]])
        assert np.all(vec[~ok] == 0)


def test_bounds_gershgorin_hand_computed():
    # clean square lattice t=-1: centres 0, radii 4 -> [-4, 4] widened by 2.5% of 8
    H, _ = jcl.square_lattice(8, W=0.0)
    b = jcl.bounds(H)
    assert np.allclose(b, [-4.2, 4.2], rtol=0, atol=1e-14)
    # onsite +/-0.5 extremes: centre e_i, radius 4 -> [min e - 4, max e + 4]
    eps = np.linspace(-0.5, 0.5, 64)
    H, _ = jcl.square_lattice(8, onsite=eps)
    lo, hi = -4.5, 4.5
    ext = (hi - lo) * 0.025
    assert np.allclose(jcl.bounds(H), [-(hi + ext), hi + ext], atol=1e-12)
    # column sums (axis=0), not row sums: an asymmetric matrix tells them apart
    A = sp.csr_matrix(np.array([[0, 3.0], [1.0, 0]], dtype=np.complex128))
    g = jcl.Gershgorin(A)
    # column sums are (1, 3): discs [-1, 1] and [-3, 3]; row sums would give (3, 1)
    # in the other order, so only the per-column radii give [-3, 3] with centre 0
    assert np.allclose(g.real, [-3.0, 3.0])
    B = sp.csr_matrix(np.array([[1.0, 3.0], [0.0, -1.0]], dtype=np.complex128))
    # columns: |3| in column 1 -> disc of the -1 centre is [-4, 2]; rows would
    # put the radius 3 on the +1 centre instead ([-2, 4])
    assert np.allclose(jcl.Gershgorin(B).real, [-4.0, 2.0])


def test_square_lattice_spectrum():
    L = 6
    H, _ = jcl.square_lattice(L, t=-1.0, W=0.0)
    k = 2 * np.pi * np.arange(L) / L
    ek = np.sort((-2 * (np.cos(k)[:, None] + np.cos(k)[None, :])).ravel())
    assert np.allclose(np.linalg.eigvalsh(H.toarray()), ek, atol=1e-12)
    assert H.shape == (36, 36) and np.diff(H.indptr).max() == 5


def test_honeycomb_spectrum_and_velocity():
    L, t = 6, -2.7
    H, Dx, Dy = jcl.honeycomb_lattice(L, t=t, W=0.0)
    k = 2 * np.pi * np.arange(L) / L
    f = 1 + np.exp(-1j * k)[:, None] + np.exp(-1j * k)[None, :]
    e = np.sort(np.concatenate([np.abs(t * f).ravel(), -np.abs(t * f).ravel()]))
    assert np.allclose(np.linalg.eigvalsh(H.toarray()), e, atol=1e-10)
    assert np.diff(H.indptr).max() == 4
    # bond vectors are antisymmetric with |d| = a_cc, and v = -i H.*Dx is Hermitian
    assert abs(Dx + Dx.T).max() == 0 and abs(Dy + Dy.T).max() == 0
    Dd = np.sqrt(np.abs(Dx.toarray()) ** 2 + np.abs(Dy.toarray()) ** 2)
    off = H.toarray() != 0
    np.fill_diagonal(off, False)
    assert np.allclose(Dd[off], jcl.A_CC)
    A = jcl.ell_matrix(H, Dx, Dy)
    v = jcl.modifier_velocity(A, 'x')
    V = sp.csr_matrix((v, (np.arange(v.size) // A.len_row, A.indices)), shape=H.shape)
    assert abs(V - V.getH()).max() <= 1e-15
    assert np.array_equal(v, O.vec_mul(-1j, A.data, A.dx_vec))


def test_open_boundaries():
    H, _ = jcl.square_lattice(5, periodic=False)
    assert np.diff(H.indptr).min() == 3  # corner: diagonal + 2 bonds
    Hh, Dx, _ = jcl.honeycomb_lattice(4, periodic=False)
    assert abs(Hh - Hh.getH()).max() == 0
    assert np.diff(Hh.indptr).min() >= 2


def test_modifiers_match_oracle():
    H, _ = jcl.square_lattice(7, W=1.0, seed=5)
    A = jcl.ell_matrix(H)
    B = A.bounds[1]
    scaled = jcl.modifier_bounds(A)
    ref = A.data.copy()
    O.scal(1 / ((A.bounds[1] - A.bounds[0]) / 2), ref)
    assert np.array_equal(scaled, ref)
    assert np.allclose(scaled, A.data / B, rtol=1e-15)
    diag = np.random.default_rng(1).normal(size=A.shape[0])
    got = jcl.modifier_diagonal(A, None, diag)
    want = O.diagonal_modifier(A.data, A.indices, A.valid_row_values, diag, A.len_row)
    assert np.array_equal(got, want)
    got = jcl.modifier_diagonal_anderson(A, 0.5, seed=3)
    eps = 0.5 * (2 * np.random.default_rng(3).random(A.shape[0]) - 1)
    want = O.diagonal_modifier(A.data, A.indices, A.valid_row_values, eps, A.len_row)
    assert np.array_equal(got, want)
    # asymmetric bounds: diagonal shift by -(b1+b0)/2, then scale
    A2 = A.deep_copy()
    A2.bounds = np.array([-3.0, 5.0])
    got = jcl.modifier_bounds(A2)
    shift = O.diagonal_modifier(A2.data, A2.indices, A2.valid_row_values,
                                np.full(A2.shape[0], -1.0), A2.len_row)
    assert np.allclose(got, shift / 4.0, rtol=1e-15, atol=0)


def _reference_light_factor(mode, params, t, dx, dy):
    """The elementwise factor each light mode of jclsquant/modifiers.py:133-181
    hands to vec_mul, written out per mode as the reference spells it."""
    if mode == "linear":
        A0, w = params
        return np.exp(1j * A0 * np.sin(w * t) * dy)
    if mode == "linear_packed":
        A0, w, Tp = params
        return np.exp(1j * A0 * np.sin(w * t) * dy / np.cosh((t - 2 * Tp) / (0.5673 * Tp)))
    if mode == "circle":
        A0, w, pol = params
        s = 1.0 if pol == "r" else -1.0
        return np.exp(1j * A0 * (np.cos(w * t) * dx + s * np.sin(w * t) * dy))
    A0, w, pol, Tp = params
    s = 1.0 if pol == "r" else -1.0
    return np.exp(1j * A0 * (np.cos(w * t) * dx + s * np.sin(w * t) * dy)
                  / np.cosh((t - 2 * Tp) / (0.5673 * Tp)))


def _check_light_phases(phases):
    """Hand-computed (ax, ay) of each mode at special times, where cos/sin
    and the envelope 1/cosh((t - 2 Tp)/(0.5673 Tp)) take known values."""
    A0, w, Tp = 0.8, 1.7, 3.0
    quarter = np.pi / (2 * w)  # cos(w t) = 0, sin(w t) = 1
    sech1 = 1 / np.cosh(1.0)   # 0.6480542736638853...
    cases = [
        (("linear", (A0, w), 0.0), (0.0, 0.0)),
        (("linear", (A0, w), quarter), (0.0, A0)),
        (("circle", (A0, w, "r"), 0.0), (A0, 0.0)),
        (("circle", (A0, w, "r"), quarter), (0.0, A0)),
        (("circle", (A0, w, "l"), quarter), (0.0, -A0)),
        (("circle_packed", (A0, w, "r", Tp), 2 * Tp), (A0 * np.cos(2 * w * Tp),
                                                       A0 * np.sin(2 * w * Tp))),
        (("circle_packed", (A0, w, "l", Tp), 2 * Tp + 0.5673 * Tp),
         (A0 * np.cos(w * 2.5673 * Tp) * sech1, -A0 * np.sin(w * 2.5673 * Tp) * sech1)),
        (("linear_packed", (A0, w, Tp), 2 * Tp - 0.5673 * Tp),
         (0.0, A0 * np.sin(w * 1.4327 * Tp) * sech1)),
    ]
    for (mode, params, t), (ax, ay) in cases:
        got = phases(mode, params, t)
        if not (abs(got[0] - ax) <= 1e-14 and abs(got[1] - ay) <= 1e-14):
            return False
    return True


def test_light_phases_pinned():
    """light_phases against hand-computed values at special times and against
    the reference's elementwise factor for every mode and polarisation; a
    swapped 'l' sign, a changed envelope constant or cos/sin swapped fail."""
    assert _check_light_phases(jcl.light_phases)
    rng = np.random.default_rng(3)
    dx, dy = rng.normal(size=64) * 0.14, rng.normal(size=64) * 0.14
    for mode, params in (("linear", (2.1, 0.7)), ("linear_packed", (2.1, 0.7, 4.0)),
                         ("circle", (2.1, 0.7, "r")), ("circle", (2.1, 0.7, "l")),
                         ("circle_packed", (2.1, 0.7, "r", 4.0)),
                         ("circle_packed", (2.1, 0.7, "l", 4.0))):
        for t in (0.0, 0.9, 3.3, 8.0, 11.7):
            ax, ay = jcl.light_phases(mode, params, t)
            got = np.exp(1j * (ax * dx + ay * dy))
            want = _reference_light_factor(mode, params, t, dx, dy)
            assert np.max(np.abs(got - want)) <= 1e-14, (mode, params, t)
    assert jcl.light_phases("circle", (1.0, 1.0, "x"), 0.0) is None
    assert jcl.light_phases("square", (1.0, 1.0), 0.0) is None

    def l_sign_swapped(mode, params, t):
        ax, ay = jcl.light_phases(mode, params, t)
        return (ax, -ay) if mode.startswith("circle") and params[2] == "l" else (ax, ay)

    def envelope_constant(mode, params, t):
        ax, ay = jcl.light_phases(mode, params, t)
        if mode.endswith("packed"):
            Tp = params[-1]
            f = np.cosh((t - 2 * Tp) / (0.5673 * Tp)) / np.cosh((t - 2 * Tp) / (0.567 * Tp))
            return ax * f, ay * f
        return ax, ay

    def cos_sin_swapped(mode, params, t):
        ax, ay = jcl.light_phases(mode, params, t)
        return (ay, ax) if mode.startswith("circle") else (ax, ay)

    for bad in (l_sign_swapped, envelope_constant, cos_sin_swapped):
        assert not _check_light_phases(bad), bad.__name__


def test_kernel_and_moments_match_oracle():
    for M in (1, 3, 64, 1000):
        assert np.max(np.abs(jcl.JacksonKernel(M) - O.jackson(M))) <= 1e-14
    for e in (-0.9, 0.0, 0.37):
        assert np.max(np.abs(jcl.moments_delta(e, 50) - O.moments_delta(e, 50))) <= 1e-13
    with pytest.raises(ValueError):
        jcl.moments_delta(1.5, 10)


def test_random_vector_generator():
    r = jcl.random_vector_generator(1000, 42)
    assert np.allclose(np.abs(r), 1.0)
    assert np.array_equal(r, jcl.random_vector_generator(1000, 42))
    ph = np.exp(1j * np.random.default_rng(42).uniform(0, 2 * np.pi, 1000))
    assert np.array_equal(r, ph)
    assert jcl.random_vector_generator((3, 10), 1).shape == (3, 10)


def test_modifier_returns_new_object_and_drops_device_cache():
    H, _ = jcl.square_lattice(4)
    A = jcl.ell_matrix(H)
    A._dev[("sentinel",)] = object()
    B = A.modifier(jcl.modifier_bounds)
    assert B is not A and B._dev == {} and ("sentinel",) in A._dev
    A.data = A.data.copy()
    assert A._dev == {}
    assert np.array_equal(B.indices, A.indices)


def test_square_lattice_ell_matches_csr_path():
    for L in (3, 5, 16):
        d, i = jcl.square_lattice_ell(L, t=-1.0, W=2.0, seed=L)
        A = jcl.ell_matrix(jcl.square_lattice(L, t=-1.0, W=2.0, seed=L)[0])
        assert np.array_equal(d, A.data) and np.array_equal(i, A.indices)


def test_block_diagonal_layout():
    mats = [jcl.ell_matrix(jcl.square_lattice(L, W=1.0, seed=L)[0]) for L in (3, 4)]
    Hh, _, _ = jcl.honeycomb_lattice(2, W=1.0, seed=1)
    mats.append(jcl.ell_matrix(Hh))  # K = 4 < 5: padded
    blocks = [(A.data, A.indices, A.shape[0], A.len_row, 0.5) for A in mats]
    data, idx, off, K = jcl.block_diagonal(blocks)
    assert K == 5 and list(off) == [0, 9, 25, 33]
    big = sp.csr_matrix((data, (np.arange(data.size) // K, idx)), shape=(33, 33))
    want = sp.block_diag([0.5 * A.A for A in mats]).tocsr()
    assert abs(big - want).max() == 0
    # padded slots of the narrower block repeat their row's last column
    last = idx.reshape(33, K)[25:, 3]
    assert np.array_equal(idx.reshape(33, K)[25:, 4], last)
    assert np.all(data.reshape(33, K)[25:, 4] == 0)


def test_recursion_api_argument_behaviour(capsys):
    H, _ = jcl.square_lattice(4)
    A = jcl.ell_matrix(H)
    assert jcl.rec_A_tab(10, A) is None
    assert "must introduce" in capsys.readouterr().out
    assert jcl.rec_A_tab2v(10, A, None, None) is None


def test_coefficient_tables_closed_forms():
    """Each table against the function it expands, summed with numpy's own
    Chebyshev series (Clenshaw), or against an independent integral."""
    from numpy.polynomial import chebyshev as C
    import scipy.integrate as I

    x = np.linspace(-0.95, 0.95, 41)
    # Jacobi-Anger: sum_n c_n T_n(x) = exp(-i x t)
    for t in (0.5, 7.3):
        c = jcl.moments_U(t, 80)
        assert np.max(np.abs(C.chebval(x, c) - np.exp(-1j * x * t))) < 1e-12
    # 1/(z - x) and its Lorentzian imaginary part
    z = 0.3 + 0.2j
    c = jcl.moments_Gmas_2(0.3, 0.2, 600)
    assert np.max(np.abs(C.chebval(x, c) - 1 / (z - x))) < 1e-10
    lor = 0.2 / np.pi / ((0.3 - x) ** 2 + 0.2 ** 2)
    assert np.max(np.abs(C.chebval(x, jcl.moments_delta_2(0.3, 0.2, 600)) - lor)) < 1e-10
    # the i0 limits: Gmas = lim Gmas_2, Gmin = conj(Gmas), Gsum = Gmas + Gmin
    g = jcl.moments_Gmas(0.3, 50)
    assert np.max(np.abs(g - jcl.moments_Gmas_2(0.3, 1e-12, 50))) < 1e-8
    assert np.allclose(jcl.moments_Gmin(0.3, 50), np.conj(g), rtol=0, atol=1e-14)
    assert np.allclose(jcl.moments_Gsum(0.3, 50), g + np.conj(g), rtol=0, atol=1e-13)
    # step theta(Ef - x): c_n = (2 - [n==0]) / pi * integral over theta in [acos Ef, pi]
    Ef = -0.35
    c = jcl.moments_FD_T0(Ef, 30)
    th0 = np.arccos(Ef)
    ex = [(1 if n == 0 else 2) / np.pi * I.quad(lambda th: np.cos(n * th), th0, np.pi)[0]
          for n in range(30)]
    assert np.max(np.abs(c - ex)) < 1e-12
    # Fermi-Dirac tables: the two entry points sample the same occupation
    N, T = 32, 0.05
    E = np.linspace(-1 + 1e-3, 1 - 1e-3, 2 * N)
    fd = jcl.moments_FD_T(0.1, T, N)
    assert np.array_equal(fd, jcl.moments_FD(N, 1 / (1 + np.exp((E - 0.1) / T))))
    # ... and at low temperature approach the step (Simpson on 2N points is coarse)
    # (the grid stops 1e-3 short of the 1/sqrt(1-x^2) end points: ~1% off)
    assert abs(jcl.moments_FD_T(0.1, 1e-3, 400)[0] - jcl.moments_FD_T0(0.1, 1)[0]) < 2e-2
    # sum_i c_i matrix[i]
    rng = np.random.default_rng(0)
    mat = rng.normal(size=(7, 5)) + 1j * rng.normal(size=(7, 5))
    cc = rng.normal(size=4)
    assert np.allclose(jcl.moments_by_matrix(cc, mat), cc @ mat[:4], rtol=0, atol=1e-13)


def test_coefficient_tables_reject_planted_mistakes():
    """The closed-form checks above catch plausible slips in a table."""
    from numpy.polynomial import chebyshev as C

    x = np.linspace(-0.95, 0.95, 41)
    c = jcl.moments_U(3.0, 60)
    bad = c.copy()
    bad[1:] /= 2  # forgotten factor 2 on n >= 1
    assert np.max(np.abs(C.chebval(x, bad) - np.exp(-1j * x * 3.0))) > 1e-2
    bad = np.conj(c)  # (+i)^n instead of (-i)^n
    assert np.max(np.abs(C.chebval(x, bad) - np.exp(-1j * x * 3.0))) > 1e-2


@pytest.mark.parametrize("cap", [12, 64, 512])
def test_tile_plan_host_check(cap):
    """The patch renumbering of the two-steps-per-launch path, built and
    fully checked on the host (bijection, halos outside their patch, every
    patch-local column naming the right row) for lattices and a random
    graph; halos stay a modest share of the rows on lattices."""
    from minijclsquant_amd import _native as nat
    from minijclsquant_amd.batch import square_lattice_ell

    H, _, _ = jcl.honeycomb_lattice(24, t=-2.7, W=1.0, seed=1)
    A = jcl.ell_matrix(H)
    r = nat.tiles_plan_check(A.indices, A.shape[0], A.len_row, cap)
    assert r["max_ext"] <= cap
    if cap == 512:
        assert r["halo_rows"] < 0.25 * A.shape[0]
    d, ix = square_lattice_ell(20, -1.0, 1.0, 0)
    r = nat.tiles_plan_check(ix, 400, 5, cap)
    assert r["max_ext"] <= cap
    # a random graph: large halos, still a consistent plan
    rng = np.random.default_rng(cap)
    n, k = 300, 4
    idx = rng.integers(0, n, size=(n, k)).astype(np.int32)
    r = nat.tiles_plan_check(idx.ravel(), n, k, cap)
    assert r["max_ext"] <= cap and r["npatch"] >= 1


def test_tile_plan_rejects_bad_input():
    from minijclsquant_amd import _native as nat

    with pytest.raises(ValueError):
        nat.tiles_plan_check(np.array([0, 5], np.int32), 2, 1, 64)  # column out of range
    with pytest.raises(ValueError):
        nat.tiles_plan_check(np.zeros(8, np.int32), 2, 4, 6)  # cap below 2K+2


def test_wave_plan_host_check():
    """The wavefront plan (two steps per launch with readiness flags), built
    and independently checked on the host.  Honeycomb 1024 x 1024 cells in
    1024 chunks of one lattice row (2048 rows, blocks of 32): every kept row
    reads at most one block ahead; the x = 0 site of each row reads x =
    Lx - 1, 63 blocks ahead, and is a tail row; the rows next to the 8 XCD
    boundaries (and the periodic wrap of the chunk ring), half of each of the
    16 boundary chunks, are fix-up rows."""
    from minijclsquant_amd import _native as nat

    H, _, _ = jcl.honeycomb_lattice(1024, t=-2.7, W=1.0, seed=2024)
    A = jcl.ell_matrix(H)
    N = A.shape[0]
    r = nat.wave_plan_check(A.indices, N, A.len_row, 1024, 2048, 8, 1)
    assert r == {"lag": 1, "tail_rows": 1016, "fixup_rows": 16384}
    # blocks of 64 rows (R = 8): the same rows, fewer blocks
    assert nat.wave_plan_check(A.indices, N, A.len_row, 1024, 2048, 16, 1) == r
    # chunks of half a lattice row: the neighbour rows sit two chunks away
    r2 = nat.wave_plan_check(A.indices, N, A.len_row, 2048, 1024, 8, 1)
    assert r2["lag"] == 1 and r2["fixup_rows"] == 16384


def test_wave_plan_rejects_unfit_columns():
    from minijclsquant_amd import _native as nat

    rng = np.random.default_rng(7)
    n, k = 65536, 4
    idx = rng.integers(0, n, size=(n, k)).astype(np.int32)
    with pytest.raises(ValueError, match="wavefront plan"):
        nat.wave_plan_check(idx.ravel(), n, k, 256, 256, 8, 1)  # random columns
    H, _, _ = jcl.honeycomb_lattice(256, t=-2.7, W=1.0, seed=1)
    A = jcl.ell_matrix(H)
    with pytest.raises(ValueError, match="chunk"):
        nat.wave_plan_check(A.indices, A.shape[0], A.len_row, 1000, 128, 8, 1)  # not a multiple of 8
    with pytest.raises(ValueError, match="every chunk"):
        nat.wave_plan_check(A.indices, A.shape[0], A.len_row, 2048, 128, 8, 1)  # empty chunks


def test_lattice_stride():
    """The dominant long-range column distance of lattice matrices."""
    from minijclsquant_amd.batch import lattice_stride, square_lattice_ell, _ell_stride, _csr_stride

    for L in (65, 100, 257):
        d, i = square_lattice_ell(L, -1.0, 1.0, seed=0)
        assert _ell_stride(i, L * L, 5) == L
    H, _ = jcl.square_lattice(90, seed=1)
    assert _csr_stride(H) == 90
    # short-range only, or no dominant distance: 0
    d, i = square_lattice_ell(40, -1.0, 0.0, seed=0)
    assert _ell_stride(i, 1600, 5) == 0
    rng = np.random.default_rng(0)
    n = 5000
    rows = np.repeat(np.arange(n), 4)
    cols = rng.integers(0, n, size=4 * n)
    assert lattice_stride(rows, cols, n) == 0


def test_tuned_times_short_steps_longer(tmp_path, monkeypatch):
    """autotune.tuned: a first pass whose timed span is under 2 ms (8 steps of
    25 us) is repeated with enough steps for about 2 ms per candidate (81),
    at most 256; a long step is tuned once; a cached shape is not re-tuned.
    A stand-in handle records the step counts it is asked to time."""
    from minijclsquant_amd import autotune

    monkeypatch.setenv("JKPM_TUNE_TABLE", str(tmp_path / "table.json"))
    monkeypatch.delenv("JKPM_LAUNCH", raising=False)
    monkeypatch.setattr(autotune, "_key", lambda h, R, d: f"k|R={R}|{d}")
    # the finalist round times real launches; the stand-in keeps the sweep's pick
    monkeypatch.setattr(autotune, "_finalists", lambda h, v_r, v, m, us, d: (v, us))

    class Handle:
        def __init__(self, us):
            self.us, self.calls, self.launch = us, [], (0, 8)

        def autotune(self, v_r, steps, doubling=False):
            self.calls.append(steps)
            self.launch = (5, 4)
            return self.us

        def get_launch(self, doubling=False):
            return self.launch

        def set_launch(self, v, m, doubling=False):
            self.launch = (v, m)

        def set_pass(self, g, doubling=False):
            pass

    class Vecs:
        shape = (100, 4)

    h = Handle(25.0)
    assert autotune.tuned(h, Vecs()) == (5, 4)
    assert h.calls == [8, 81]
    h = Handle(300.0)
    autotune.tuned(h, Vecs(), doubling=True)
    assert h.calls == [8]
    monkeypatch.setattr(autotune, "_key", lambda h, R, d: "tiny")
    h = Handle(0.5)
    autotune.tuned(h, Vecs())
    assert h.calls == [8, 256]
    h = Handle(25.0)
    assert autotune.tuned(h, Vecs()) == (5, 4) and h.calls == []


def test_tuned_picks_vectors_per_pass(tmp_path, monkeypatch):
    """autotune.tuned tries all R vectors at once and passes of R/2, R/4, ...
    down to 4 vectors, tunes the launch shape for each, keeps the fastest
    whole step, stores the pass size in the table and applies it again from
    the table."""
    from minijclsquant_amd import autotune

    assert autotune.pass_candidates(16) == [0, 8, 4]
    assert autotune.pass_candidates(8) == [0, 4]
    assert autotune.pass_candidates(4) == [0]
    assert autotune.pass_candidates(1) == [0]
    assert autotune.pass_candidates(40) == [0, 20, 10, 5]
    monkeypatch.setenv("JKPM_TUNE_TABLE", str(tmp_path / "table.json"))
    monkeypatch.delenv("JKPM_LAUNCH", raising=False)
    monkeypatch.setattr(autotune, "_key", lambda h, R, d: f"k|R={R}|{d}")
    monkeypatch.setattr(autotune, "_finalists", lambda h, v_r, v, m, us, d: (v, us))
    # the passes are compared by their per-step slope, not the sweep's time
    monkeypatch.setattr(autotune, "_step_slope", lambda h, v_r, d, us: slope[h.g])
    step_us = {0: 180.0, 8: 190.0, 4: 185.0}
    slope = {0: 180.0, 8: 190.0, 4: 170.0}

    class Handle:
        def __init__(self):
            self.g, self.launch, self.passes = 0, (0, 8), []

        def set_pass(self, g, doubling=False):
            self.g = g
            self.passes.append(g)

        def autotune(self, v_r, steps, doubling=False):
            self.launch = (10 + self.g, 4)
            return step_us[self.g]

        def get_launch(self, doubling=False):
            return self.launch

        def set_launch(self, v, m, doubling=False):
            self.launch = (v, m)

    class Vecs:
        shape = (100, 16)

    h = Handle()
    assert autotune.tuned(h, Vecs()) == (14, 4)
    assert h.passes == [0, 8, 4, 4] and h.g == 4
    h2 = Handle()
    assert autotune.tuned(h2, Vecs()) == (14, 4)
    assert h2.passes == [4]
    # less than 2 % faster per step: all vectors at once stay
    slope[4] = 177.0
    h3 = Handle()
    assert autotune.tuned(h3, Vecs(), retune=True) == (10, 4)
    assert h3.g == 0


def test_finalist_neighbours_flip_one_choice():
    """The tuner's finalist round times the sweep's winner against the
    variants that differ from it in exactly one of: non-temporal matrix
    loads (+-24), non-temporal stores (+-12 within a 24-block), non-temporal
    stream loads (bit 1); the unroll bit and lane shape stay."""
    from minijclsquant_amd import autotune

    def fields(v):
        return {"ntm": v >= 24, "ns": v % 24 >= 12, "nt": bool(v & 2), "unroll": v & 1,
                "shape": (v % 12) >> 2}

    assert autotune.N_VARIANTS == 48
    for v in range(48):
        nb = autotune._neighbours(v)
        assert len(nb) == 3 and len(set(nb)) == 3
        for u, flipped in zip(nb, ("ntm", "ns", "nt")):
            assert 0 <= u < 48
            a, b = fields(v), fields(u)
            assert a[flipped] != b[flipped]
            assert all(a[k] == b[k] for k in a if k != flipped), (v, u)
    # XOR with 12 or 24 is not a one-choice flip everywhere
    assert autotune._neighbours(20)[1] == 8 and (20 ^ 12) == 24
    assert autotune._neighbours(38)[0] == 14 and (38 ^ 24) == 62


def _min_image_bonds(S, ncx, ncy, thr):
    """Brute-force nearest neighbours of the 4-atom-cell flake on its torus:
    every pair, separation wrapped to the nearest periodic image (periods
    ncx sqrt(3) a_cc and 3 ncy a_cc)."""
    a_cc = jcl.A_CC
    P = np.array([ncx * np.sqrt(3) * a_cc, 3 * ncy * a_cc])
    d = S[None, :, :] - S[:, None, :]  # d[i, j] = pos_j - pos_i
    d = d - P * np.round(d / P)
    dist = np.sqrt((d ** 2).sum(-1))
    i, j = np.nonzero((dist < thr) & (dist > 0))
    return i, j, d[i, j]


@pytest.mark.parametrize("N", [16, 64, 256, 1024])
def test_H_graphene_periodic_bonds_vs_minimum_image(N):
    """The border pairing of H_graphene (the shift rule of
    jclsquant/hams.py:108-133) builds exactly the minimum-image torus: every
    atom has three neighbours at a_cc, the hoppings are t and Dx/Dy hold the
    wrapped x_j - x_i, y_j - y_i."""
    S = jcl.lattice_hexagonal(N)
    t = -2.7 + 0j
    H, Dx, Dy, Omega, cj = jcl.H_graphene(S, t, 0, 0, True, 'CSR')
    n = int(np.sqrt(N // 4))
    i, j, d = _min_image_bonds(S, n, n, 1.1 * jcl.A_CC)
    want = sp.csr_matrix((np.full(i.size, t), (i, j)), shape=(N, N))
    assert abs(H - want).max() == 0 and H.nnz == want.nnz == 3 * N
    wdx = sp.csr_matrix((d[:, 0], (i, j)), shape=(N, N))
    wdy = sp.csr_matrix((d[:, 1], (i, j)), shape=(N, N))
    assert abs(Dx - wdx).max() <= 1e-12 and abs(Dy - wdy).max() <= 1e-12
    assert np.allclose(np.sqrt(Dx.data ** 2 + Dy.data ** 2), jcl.A_CC, rtol=1e-12)
    Lx = (n - 0.5) * np.sqrt(3) * jcl.A_CC
    Ly = (3 * n - 0.5) * jcl.A_CC
    assert abs(Omega - Lx * Ly) <= 1e-12 * Omega
    rows = np.repeat(np.arange(N), np.diff(cj.indptr))
    assert np.array_equal(cj.indices, H.indices) and np.array_equal(cj.indptr, H.indptr)
    assert np.array_equal(cj.data, (rows > cj.indices).astype(cj.data.dtype))


def test_H_graphene_clean_spectrum():
    """Clean graphene torus of 2 n^2 two-atom cells: eigenvalues
    +-|t| |sum_d exp(i k.d)| over the NN vectors d at the k allowed by the
    rectangular supercell (2 pi m / (n sqrt(3) a_cc), 2 pi q / (3 n a_cc)),
    m < n, q < 2n."""
    N = 64
    n = 4
    S = jcl.lattice_hexagonal(N)
    H, _, _, _, _ = jcl.H_graphene(S, -2.7, 0, 0, True, 'CSR')
    a = jcl.A_CC
    deltas = np.array([[0.0, a], [np.sqrt(3) * a / 2, -a / 2], [-np.sqrt(3) * a / 2, -a / 2]])
    ev = []
    for mm in range(n):
        for q in range(2 * n):
            k = np.array([2 * np.pi * mm / (n * np.sqrt(3) * a), 2 * np.pi * q / (3 * n * a)])
            f = abs(np.exp(1j * deltas @ k).sum())
            ev += [2.7 * f, -2.7 * f]
    got = np.linalg.eigvalsh(H.toarray())
    assert np.max(np.abs(np.sort(ev) - got)) <= 1e-12


def test_H_graphene_mass_disorder_and_ell_form():
    """m: +m on even, -m on odd sites; W: onsite W (2u - 1) from numpy's
    global generator (or default_rng(seed)); 'ELL': the reference's 1e-6
    cross terms (H += 1e-6 (Dx + Dy), then Dx, Dy += 1e-6 H), Omega kept,
    and the result is not Hermitian."""
    N = 64
    S = jcl.lattice_hexagonal(N)
    H, Dx, Dy, Om, cj = jcl.H_graphene(S, -2.7, 0.3, 0, True, 'CSR')
    assert np.array_equal(H.diagonal(), np.where(np.arange(N) % 2 == 0, 0.3, -0.3))
    assert np.all(cj.diagonal() == 2)
    np.random.seed(5)
    Hw = jcl.H_graphene(S, -2.7, 0, 0.8, True, 'CSR')[0]
    np.random.seed(5)
    assert np.array_equal(Hw.diagonal(), 0.8 * (2 * np.random.rand(N) - 1))
    Hs = jcl.H_graphene(S, -2.7, 0, 0.8, True, 'CSR', seed=9)[0]
    assert np.array_equal(Hs.diagonal(), 0.8 * (2 * np.random.default_rng(9).random(N) - 1))
    A = jcl.H_graphene(S, -2.7 + 0j, 0.3, 0, True, 'ELL')
    assert abs(A.Omega - Om) == 0
    He = H + 1e-6 * (Dx + Dy)
    want = jcl.ell_matrix(He, Dx + 1e-6 * He, Dy + 1e-6 * He)
    assert np.array_equal(A.data, want.data) and np.array_equal(A.indices, want.indices)
    assert np.array_equal(A.dx_vec, want.dx_vec) and np.array_equal(A.dy_vec, want.dy_vec)
    Hc = jcl.ell_to_csr(A)
    assert abs(Hc - Hc.conj().T).max() > 1e-7
    assert A.conjugates.size == A.data.size


def test_extra_funs_match_oracle():
    """The Extra_funs entries of the namespace: data_indices_c(_0) fill the
    caller's arrays like the oracle's CSR -> ELL, diagonal_modifier writes
    data_new in place like the oracle's, table_G_delta equals the oracle's
    table (pinned by eval_chebyu) including the 1e-16 guards at E = +-1."""
    H, Dx, Dy = jcl.honeycomb_lattice(5, W=1.0, seed=2)
    N = H.shape[0]
    d, i, v, K = O.csr_to_ell(H)
    out_i, out_d, out_v = np.zeros(N * K, np.int32), np.zeros(N * K, complex), np.zeros(N * K, np.int32)
    rd, ri, rv = jcl.data_indices_c_0(K, N, H.indptr, H.data, H.indices, out_i, out_d, out_v, 8)
    assert np.array_equal(rd, d) and np.array_equal(ri, i) and np.array_equal(rv, v)
    assert np.array_equal(out_d, d)  # filled in place
    arrs = [np.zeros(N * K, t) for t in (np.int32, complex, complex, complex, np.int32, np.int32)]
    cj = (np.repeat(np.arange(N), np.diff(H.indptr)) > H.indices).astype(np.int32)
    res = jcl.data_indices_c(K, N, H.indptr, H.data, H.indices, Dx.data.astype(complex),
                             Dy.data.astype(complex), cj, *arrs, 8)
    A = jcl.ell_matrix(H, Dx, Dy)
    assert np.array_equal(res[0], A.data) and np.array_equal(res[1], A.indices)
    assert np.array_equal(res[2], A.dx_vec) and np.array_equal(res[3], A.dy_vec)
    assert np.array_equal(res[5], A.valid_row_values)
    rows = np.arange(N * K) // K
    assert np.array_equal(res[4][res[5] == 1], (rows > A.indices)[res[5] == 1])
    diag = np.random.default_rng(4).normal(size=N) + 0j
    new = np.zeros_like(A.data)
    jcl.diagonal_modifier(A.data, A.indices, A.valid_row_values, new, diag, K, A.data.size, 8)
    assert np.array_equal(new, O.diagonal_modifier(A.data, A.indices, A.valid_row_values, diag, K))
    for M, alpha, B in ((9, 0.97, 2.5), (12, 1.0, 3.1)):
        bm, bn = jcl.table_G_delta(M, alpha, B, 8)
        E = np.linspace(-alpha, alpha, 2 * M)
        om, on = O.table_G_delta(M, E, B)
        assert bm.shape == (M, M, 2 * M)
        assert np.allclose(bm, om, rtol=1e-12, atol=1e-300)
        assert np.allclose(bn, on, rtol=1e-12, atol=1e-300)


def test_bench_reads_traffic_from_committed_profiles():
    """bench.py fills every roofline's `traffic` from the committed counter
    summaries (FETCH_SIZE x 2 + WRITE_SIZE per dispatch), and only for the
    workload shape each summary names."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench_mod", ROOT / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    P = bench.PROFILES
    got = {
        "headline": bench.traffic_from_profile(2097152, 4, 16),
        "headline_dbl": bench.traffic_from_profile(2097152, 4, 16,
                                                   "doubling_traffic_bytes_per_launch"),
        "headline_pass4": bench.traffic_from_profile(2097152, 4, 16, g=4),
        "headline_dbl_pass4": bench.traffic_from_profile(2097152, 4, 16,
                                                         "doubling_traffic_bytes_per_launch", g=4),
        "batch": bench.profile_traffic(P / "cheb_step_c64_r8_batch64_square_trace.md",
                                       "traffic_bytes_per_launch", N=24054938, K=5, R=8),
        "config1": bench.profile_traffic(P / "cheb_step_c128_r4_square256_pmc.md",
                                         "traffic_bytes_per_launch", N=65536, K=5, R=4),
        "grad": bench.profile_traffic(P / "grad_adjoint_square512_pmc.md",
                                      "traffic_bytes_per_launch", N=262144, K=5, R=8),
        "kubo_gemm": bench.profile_traffic(P / "kubo_step_and_gemm_honeycomb512_pmc.md",
                                           "gemm_traffic_bytes_per_launch", N=524288, M=512, R=4),
        "kubo_step": bench.profile_traffic(P / "kubo_step_and_gemm_honeycomb512_pmc.md",
                                           "step_traffic_bytes_per_launch", N=524288, M=512, R=4),
    }
    # algorithmic bytes per launch of each kernel (the step kernels of these
    # real-valued lattices read 4 or 8 bytes per slot value, plus a 4-byte
    # column; the complex64 plain forms read their random-phase left vectors
    # as 4-byte phases); counters sit 0-7 % above
    alg = {"headline": 2097152 * (4 * 8 + 16 * 28), "headline_dbl": 2097152 * (4 * 8 + 16 * 24),
           "headline_pass4": 2097152 * (4 * 8 + 4 * 28),
           "headline_dbl_pass4": 2097152 * (4 * 8 + 4 * 24),
           "batch": 24054938 * (5 * 8 + 8 * 28), "config1": 65536 * (5 * 12 + 4 * 64),
           "grad": 262144 * (5 * 12 + 8 * 64), "kubo_gemm": 2 * 512 * 2097152 * 8,
           "kubo_step": 524288 * 4 * 12 + 4 * 524288 * 8 * 7}
    for k, v in got.items():
        assert v is not None, k
        assert 1.0 <= v / alg[k] <= 1.07, (k, v / alg[k])
    # another shape gets no figure
    assert bench.traffic_from_profile(2097152, 4, 8) is None
    assert bench.profile_traffic(P / "cheb_step_c128_r4_square256_pmc.md",
                                 "traffic_bytes_per_launch", N=65536, K=5, R=8) is None


def test_bench_pass_roofline_fields():
    """bench.pass_note for R vectors in passes of g: a step moves the matrix
    once per pass plus every vector stream once, and the single-pass figure
    counts the matrix once over the same step time (R/g launches)."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("bench_mod2", ROOT / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    N, K, R, g, mv, s_v, s_l = 2097152, 4, 16, 4, 4, 8, 4
    launch_ms = 0.04274
    f = bench.pass_note(N, K, R, g, mv, s_v, s_l, launch_ms)
    assert f["passes_per_step"] == 4
    assert f["step_bytes_all_passes"] == 4 * N * K * 8 + R * N * 28
    assert f["single_pass_bytes"] == N * K * 8 + R * N * 28 == 1006632960
    step_s = 4 * launch_ms / 1e3
    assert abs(f["single_pass_equiv_GBs"] - 1006632960 / step_s / 1e9) < 0.1
    assert abs(f["single_pass_equiv_frac"] - 1006632960 / step_s / 8e12) < 1e-4
    # doubling form: no left stream; a last pass shorter than g counts as a pass
    d = bench.pass_note(N, K, 10, 4, mv, s_v, 0, 0.03)
    assert d["passes_per_step"] == 3
    assert d["step_bytes_all_passes"] == 3 * N * K * 8 + 10 * N * 24


def test_apply_cached_uses_the_table_only(tmp_path, monkeypatch):
    """autotune.apply_cached applies a table entry (launch shape and vectors
    per pass) without timing anything, and reports when there is none."""
    import json

    from minijclsquant_amd import autotune

    monkeypatch.setenv("JKPM_TUNE_TABLE", str(tmp_path / "table.json"))
    monkeypatch.delenv("JKPM_LAUNCH", raising=False)
    monkeypatch.setattr(autotune, "_key", lambda h, R, d: f"k|R={R}|{d}")

    class Handle:
        def __init__(self):
            self.launch, self.g = (0, 8), 0

        def set_launch(self, v, m, doubling=False):
            self.launch = (v, m)

        def set_pass(self, g, doubling=False):
            self.g = g

        def autotune(self, *a, **k):
            raise AssertionError("apply_cached must not time anything")

    class Vecs:
        shape = (100, 16)

    h = Handle()
    assert autotune.apply_cached(h, Vecs()) is False and h.launch == (0, 8)
    (tmp_path / "table.json").write_text(json.dumps(
        {"k|R=16|False": {"variant": 12, "chunk_mult": 4, "pass": 4, "step_us": 1.0}}))
    assert autotune.apply_cached(h, Vecs()) is True
    assert h.launch == (12, 4) and h.g == 4
    assert autotune.apply_cached(h, Vecs(), doubling=True) is False
    monkeypatch.setenv("JKPM_LAUNCH", "3,16")
    assert autotune.apply_cached(h, Vecs(), doubling=True) is True and h.launch == (3, 16)
