"""GPU parity of the Chebyshev moment path against the CPU oracle.

Every test runs the HIP library through its C-ABI (via the package API or
the ctypes wrapper) and compares with oracle/ (complex128 restatement of the
reference).  Tolerances: complex128 recursion max_m |dmu_m| / mu_0 <= 1e-10;
complex64 recursion (FP64 dots) max_m |dmu_m| / mu_0 <= 1e-5.
"""
import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _native as nat
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL128 = 1e-10
TOL64 = 1e-5


def square(L, W=1.0, seed=0, B=None):
    H, _ = jcl.square_lattice(L, W=W, seed=seed)
    return jcl.ell_matrix(H, bounds=None if B is None else [-B, B])


def honeycomb(L, W=1.0, seed=0, t=-2.7):
    H, Dx, Dy = jcl.honeycomb_lattice(L, t=t, W=W, seed=seed)
    return jcl.ell_matrix(H, Dx, Dy)


def vectors(N, R, seed0=0):
    return np.stack([jcl.random_vector_generator(N, seed0 + s) for s in range(R)])


def rel_err(mu, ref):
    return float(np.max(np.abs(mu - ref)) / abs(ref[0]))


def test_config1_square_c128_parity(torch_gpu):
    """Square 256x256, W=1, K=5, B=(4+W/2)*1.01, M=512, R=4, complex128."""
    A = square(256, W=1.0, seed=11, B=(4 + 0.5) * 1.01)
    assert A.len_row == 5
    R, M = 4, 512
    r = vectors(A.shape[0], R)
    Ht = A.modifier(jcl.modifier_bounds)
    mu = jcl.rec_A_tab(M, Ht, r)
    assert mu.shape == (R, M)
    for k in range(R):
        ref = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
        assert rel_err(mu[k], ref) <= TOL128
    # the single-vector form returns [M]
    mu0 = jcl.rec_A_tab(M, Ht, r[0])
    assert mu0.shape == (M,)
    assert np.array_equal(mu0, mu[0]) or rel_err(mu0, mu[0]) <= 1e-13


def test_two_vector_variants(torch_gpu):
    A = honeycomb(16, seed=3)
    Ht = A.modifier(jcl.modifier_bounds)
    N, M = A.shape[0], 200
    vl, vr = jcl.random_vector_generator(N, 1), jcl.random_vector_generator(N, 2)
    mu = jcl.rec_A_tab2v(M, Ht, vl, vr)
    ref = O.rec_A_tab2v(M, Ht.data, Ht.indices, Ht.len_row, vl, vr)
    assert np.max(np.abs(mu - ref)) / N <= TOL128
    m1, m2 = jcl.rec_A_tab2v_2(M, Ht, vl, vr)
    r1, r2 = O.rec_A_tab2v_2(M, Ht.data, Ht.indices, Ht.len_row, vl, vr)
    assert np.max(np.abs(m1 - r1)) / N <= TOL128
    assert np.max(np.abs(m2 - r2)) / N <= TOL128


@pytest.mark.parametrize("R", [1, 2, 3, 5, 8, 16, 17, 24, 32, 33, 40])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_batch_sizes(torch_gpu, R, dtype):
    """Every batch size (power-of-two groups, odd leading dimensions)."""
    A = honeycomb(8, seed=R)
    Ht = A.modifier(jcl.modifier_bounds)
    M = 96
    r = vectors(A.shape[0], R, seed0=100 * R)
    mu = jcl.rec_A_tab(M, Ht, r, dtype=dtype)
    tol = TOL128 if dtype == "c128" else TOL64
    for k in range(R):
        ref = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
        assert rel_err(mu[k], ref) <= tol, (R, k)


def test_c64_recursion_accuracy(torch_gpu):
    """complex64 recursion + FP64 dots vs the complex128 oracle."""
    A = honeycomb(64, seed=5)
    Ht = A.modifier(jcl.modifier_bounds)
    M, R = 512, 4
    r = vectors(A.shape[0], R, 7)
    mu = jcl.rec_A_tab(M, Ht, r, dtype="c64")
    errs = []
    for k in range(R):
        ref = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
        errs.append(rel_err(mu[k], ref))
    assert max(errs) <= TOL64, errs


def test_non_hermitian_ragged(torch_gpu):
    """Random complex non-Hermitian ELL with ragged rows (padding slots repeat
    the last column with value 0) -- the reference H_graphene 'ELL' matrices
    are non-Hermitian at the 1e-6 level."""
    import scipy.sparse as sp

    rng = np.random.default_rng(42)
    N = 700
    nnz_row = rng.integers(1, 8, N)
    rows = np.repeat(np.arange(N), nnz_row)
    cols = rng.integers(0, N, rows.size)
    vals = 0.1 * (rng.normal(size=rows.size) + 1j * rng.normal(size=rows.size))
    Hs = sp.coo_matrix((vals, (rows, cols)), shape=(N, N)).tocsr()
    Hs.sum_duplicates()
    A = jcl.ell_matrix(Hs)
    Ht = A.modifier(jcl.modifier_bounds)
    M = 150
    r = vectors(N, 3, 9)
    mu = jcl.rec_A_tab(M, Ht, r)
    for k in range(3):
        ref = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
        # complex eigenvalues: |mu_m| may grow with m, so compare to the largest
        assert np.max(np.abs(mu[k] - ref)) <= TOL128 * np.max(np.abs(ref))


@pytest.mark.parametrize("M", [1, 2, 3, 63, 64, 65, 129])
def test_moment_counts(torch_gpu, M):
    """M at and around the reduction ring boundary (64 steps per fold)."""
    A = square(12, seed=M)
    Ht = A.modifier(jcl.modifier_bounds)
    r = vectors(A.shape[0], 2, 3)
    mu = jcl.rec_A_tab(M, Ht, r)
    for k in range(2):
        ref = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r[k])
        assert mu[k].shape == (M,)
        assert rel_err(mu[k], ref) <= TOL128


def test_tiny_matrices(torch_gpu):
    import scipy.sparse as sp

    for N in (1, 2, 3):
        Hs = sp.csr_matrix(np.diag(np.linspace(-0.5, 0.5, N)).astype(np.complex128))
        A = jcl.ell_matrix(Hs, bounds=[-1.0, 1.0])
        r = vectors(N, 1, N)[0]
        mu = jcl.rec_A_tab(10, A, r)
        ref = O.rec_A_tab(10, A.data, A.indices, A.len_row, r)
        assert np.max(np.abs(mu - ref)) <= 1e-14


def test_bitwise_reproducible(torch_gpu):
    A = honeycomb(32, seed=1)
    Ht = A.modifier(jcl.modifier_bounds)
    r = vectors(A.shape[0], 16)
    a = jcl.rec_A_tab(300, Ht, r, dtype="c64")
    b = jcl.rec_A_tab(300, Ht, r, dtype="c64")
    assert np.array_equal(a, b)


def test_dropin_host_symbols(torch_gpu):
    """obv_gpu-style calls through the extern "C" rec_A_tab / rec_A_tab_2."""
    A = honeycomb(16, seed=2)
    Ht = A.modifier(jcl.modifier_bounds)
    N, M = A.shape[0], 128
    v = jcl.random_vector_generator(N, 5)
    w = jcl.random_vector_generator(N, 6)
    out = np.zeros(M, np.complex128)
    jcl.rec_A_tab_gpu(Ht.data, Ht.indices, Ht.len_row, v, out, M)
    ref = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, v)
    assert rel_err(out, ref) <= TOL128
    out2 = np.zeros(M, np.complex128)
    jcl.rec_A_tab_2_gpu(Ht.data, Ht.indices, Ht.len_row, v, w, out2, M)
    ref2 = O.rec_A_tab2v(M, Ht.data, Ht.indices, Ht.len_row, v, w)
    assert np.max(np.abs(out2 - ref2)) / N <= TOL128


def test_aAxby_dot(torch_gpu):
    A = honeycomb(10, seed=4)
    N = A.shape[0]
    rng = np.random.default_rng(0)
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    y = rng.normal(size=N) + 1j * rng.normal(size=N)
    y_ref = y.copy()
    a, b = 0.7 - 0.2j, -1.3 + 0.4j
    A.dot(a, b, x, y)
    O.aAxby(a, b, x, y_ref, A.data, A.indices, A.len_row)
    assert np.max(np.abs(y - y_ref)) <= 1e-12


def test_kpm_dos_parity(torch_gpu):
    """kpm_dos / kpm_dos_gpu: device Jackson DOS vs the oracle's DOS loop."""
    A = square(64, W=1.0, seed=8)
    N, M = A.shape[0], 256
    r = jcl.random_vector_generator(N, 3)
    for fn in (jcl.kpm_dos, jcl.kpm_dos_gpu):
        d = fn(A, M, r)
        assert d.shape == (2 * M + 1, 2)
        Ht = A.modifier(jcl.modifier_bounds)
        mu = O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, r)
        ref = O.dos(mu, d[:, 0], A.bounds[0], A.bounds[1], True)
        assert np.max(np.abs(d[:, 1] - ref)) / np.max(np.abs(ref)) <= TOL128
    # explicit energy grid and R-averaged moments
    E = np.linspace(-3.9, 3.9, 1024)
    rs = vectors(N, 4)
    d = jcl.kpm_dos(A, M, rs, energies=E)
    Ht = A.modifier(jcl.modifier_bounds)
    mu = np.mean([O.rec_A_tab(M, Ht.data, Ht.indices, Ht.len_row, v) for v in rs], axis=0)
    ref = O.dos(mu, E, A.bounds[0], A.bounds[1], True)
    assert np.max(np.abs(d[:, 1] - ref)) / np.max(np.abs(ref)) <= TOL128


def test_kpm_two_vector_dos(torch_gpu):
    A = honeycomb(12, seed=6)
    N, M = A.shape[0], 128
    vl, vr = jcl.random_vector_generator(N, 1), jcl.random_vector_generator(N, 2)
    P = jcl.ell_matrix(A.A)  # projector: any ELL matrix
    for fn in (jcl.kpm_n_dos_n, jcl.kpm_dos_2_gpu):
        d = fn(A, M, vl, vr, True, P)
        pl = np.zeros(N, np.complex128)
        O.aAxby(1, 0, vl, pl, P.data, P.indices, P.len_row)
        Ht = A.modifier(jcl.modifier_bounds)
        mu = O.rec_A_tab2v(M, Ht.data, Ht.indices, Ht.len_row, pl, vr)
        ref = O.dos(mu, d[:, 0], A.bounds[0], A.bounds[1], True)
        assert np.max(np.abs(d[:, 1] - ref)) / np.max(np.abs(ref)) <= TOL128


def test_argument_errors(torch_gpu):
    torch = torch_gpu
    A = square(4)
    bad = A.indices.copy()
    bad[3] = 10_000
    with pytest.raises(ValueError):
        nat.DeviceELL(A.data, bad, A.shape[0], A.len_row)
    h = nat.DeviceELL(A.data, A.indices, A.shape[0], A.len_row)
    v = torch.zeros((A.shape[0], 2), dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        h.moments(v, 8)  # dtype mismatch (handle is complex128)
    v = torch.zeros((A.shape[0], 2), dtype=torch.complex128, device="cuda")
    small = torch.empty(16, dtype=torch.uint8, device="cuda")
    with pytest.raises(nat.JkpmError):
        nat.check(nat.lib().jkpm_moments(h.handle, v.data_ptr(), v.data_ptr(), None, 2, 8,
                                         v.data_ptr(), small.data_ptr(), 16, None))
    with pytest.raises(ValueError):
        h.moments(v, 0)
    # device-side index validation
    dd = torch.from_numpy(A.data).cuda()
    di = torch.from_numpy(bad).cuda()
    with pytest.raises(ValueError):
        nat.DeviceELL(dd, di, A.shape[0], A.len_row)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_launch_variants_and_autotune(torch_gpu, dtype, tmp_path, monkeypatch):
    """Every launch shape gives the same moments (same per-row arithmetic;
    only the FP64 order of the chunk partials changes), and the autotuner
    keeps one of them and records it in its table."""
    from minijclsquant_amd import _kpm, autotune

    A = square(40, seed=2)  # K = 5 (variants exist)
    Ht = A.modifier(jcl.modifier_bounds)
    r = vectors(A.shape[0], 16, 5)
    h = nat.DeviceELL(Ht.data, Ht.indices, A.shape[0], A.len_row, 1.0, dtype)
    vr = _kpm.as_batch(r, A.shape[0], dtype)
    ref = h.moments(vr, 200)[0].cpu().numpy()
    for v in range(nat.JKPM_STEP_VARIANTS):
        for m in (1, 4, 32):
            h.set_launch(v, m)
            got = h.moments(vr, 200)[0].cpu().numpy()
            assert np.max(np.abs(got - ref)) / abs(ref[0, 0]) <= 1e-12, (v, m)
    # odd batch (narrow lanes and power-of-two groups) under every variant
    r5 = _kpm.as_batch(r[:5], A.shape[0], dtype)
    ref5 = h.moments(r5, 90)[0].cpu().numpy()
    for v in range(nat.JKPM_STEP_VARIANTS):
        h.set_launch(v, 8)
        got = h.moments(r5, 90)[0].cpu().numpy()
        assert np.max(np.abs(got - ref5)) / abs(ref5[0, 0]) <= 1e-12, v
    # the doubling form has its own setting; every shape of it agrees too
    h.set_launch(0, 8)
    refd = nat.moments_ex(h, None, vr, 200, True).cpu().numpy()
    # (the two forms round differently: complex64 vectors leave ~1e-7)
    tol_form = 1e-11 if dtype == "c128" else 1e-6
    assert np.max(np.abs(refd - ref)) / abs(ref[0, 0]) <= tol_form
    for v in range(nat.JKPM_STEP_VARIANTS):
        for m in (1, 4, 32):
            h.set_launch(v, m, doubling=True)
            assert h.get_launch() == (0, 8)
            got = nat.moments_ex(h, None, vr, 200, True).cpu().numpy()
            assert np.max(np.abs(got - refd)) / abs(ref[0, 0]) <= 1e-12, (v, m)
    monkeypatch.setenv("JKPM_TUNE_TABLE", str(tmp_path / "t.json"))
    v, m = autotune.tuned(h, vr, steps=4)
    assert 0 <= v < nat.JKPM_STEP_VARIANTS and m in (2, 4, 8, 16, 32)
    assert (tmp_path / "t.json").exists()
    vd, md = autotune.tuned(h, vr, steps=4, doubling=True)
    assert 0 <= vd < nat.JKPM_STEP_VARIANTS and md in (2, 4, 8, 16, 32)
    assert h.get_launch(doubling=True) == (vd, md) and h.get_launch() == (v, m)
    h.set_launch(0, 8)
    h.set_launch(0, 8, doubling=True)
    assert autotune.tuned(h, vr) == (v, m)  # from the table
    assert autotune.tuned(h, vr, doubling=True) == (vd, md)
    with pytest.raises(ValueError):
        h.set_launch(nat.JKPM_STEP_VARIANTS, 8)
    assert nat.lib().jkpm_ell_set_launch(h.handle, 0, 8, 2) == nat.JKPM_ERR_ARG  # unknown flag


def test_precision_form_full_size(torch_gpu):
    """Honeycomb 1024x1024 cells (N = 2,097,152), K = 4, M = 2048, R = 16,
    complex64 recursion + FP64 dots: size-independent properties, and the
    DOS error against the complex128 GPU path (pinned to the oracle above)."""
    torch = torch_gpu
    A = honeycomb(1024, W=1.0, seed=2024)
    N, M, R = A.shape[0], 2048, 16
    assert N == 2_097_152 and A.len_row == 4
    B = (3 * 2.7 + 0.5) * 1.01
    A.bounds = np.array([-B, B])
    r = vectors(N, R, 1000)
    mu64 = jcl.kpm_moments(A, M, r, dtype="c64")  # device [R, M]
    mu128 = jcl.kpm_moments(A, M, r, dtype="c128")
    m64 = mu64.cpu().numpy()
    m128 = mu128.cpu().numpy()
    # mu_0 = sum |r_i|^2 = N; Hermitian H~ -> real moments; |mu_m| <= mu_0
    assert np.max(np.abs(m128[:, 0] - N)) / N <= 1e-12
    assert np.max(np.abs(m64[:, 0] - N)) / N <= 1e-6
    assert np.max(np.abs(m128.imag)) / N <= 1e-10
    assert np.max(np.abs(m64.imag)) / N <= 1e-5
    assert np.max(np.abs(m128)) <= N * (1 + 1e-10)
    # one vector's complex128 moments against the oracle at full size (the
    # C restatement runs this in a few seconds on the host), and every
    # complex64 moment against complex128 (the error of every mu_n)
    want = O.rec_A_tab(M, A.data * (1 / B), A.indices, A.len_row, r[3])
    assert np.max(np.abs(m128[3] - want)) / N <= 1e-10
    assert np.max(np.abs(m64 - m128)) / N <= 1e-6
    # DOS error of the complex64 recursion on 4096 energies
    E = np.linspace(-B + 1e-3, B - 1e-3, 4096)
    d64 = jcl.kpm_dos(A, M, r, energies=E, dtype="c64")[:, 1]
    d128 = jcl.kpm_dos(A, M, r, energies=E, dtype="c128")[:, 1]
    err = np.max(np.abs(d64 - d128)) / np.max(np.abs(d128))
    assert err <= 1e-5, err
    del torch


def test_size_limits_and_degenerate_shapes(torch_gpu):
    """Inputs past the 2^31 - 1 slot limit are refused before any
    allocation; a one-row matrix and M = 1, 2 still give the oracle's
    moments."""
    torch = torch_gpu
    lib = nat.lib()
    data = np.zeros(8, np.complex128)
    idx = np.zeros(8, np.int32)
    h = nat._vp()
    rc = lib.jkpm_ell_create(data.ctypes.data, idx.ctypes.data, 1 << 20, 1 << 12, 1.0,
                             nat.JKPM_C128, 0, None, nat._c.byref(h))
    assert rc == nat.JKPM_ERR_ARG and "2^31" in lib.jkpm_last_error().decode()
    # CSR whose requested width makes n_rows * len_row too large
    ip = np.array([0, 1, 2, 3, 4], np.int32)
    ci = np.array([0, 1, 2, 3], np.int32)
    dv = np.ones(4, np.complex128)
    with pytest.raises(ValueError, match="2\\^31"):
        nat.DeviceELL.from_csr(ip, ci, dv, 4, len_row=1 << 30)
    # one row, one slot: T_n = a^n r
    a = 0.3 - 0.4j
    h1 = nat.DeviceELL(np.array([a]), np.array([0], np.int32), 1, 1, 1.0, "c128")
    v = torch.tensor([[1.0 + 0j]], dtype=torch.complex128, device="cuda")
    for M in (1, 2, 7):
        mu = h1.moments(v, M)[0].cpu().numpy()[:, 0]
        ref = O.rec_A_tab(M, np.array([a]), np.array([0], np.int32), 1, np.array([1.0 + 0j]))
        assert np.allclose(mu, ref, rtol=0, atol=1e-14), M


def test_profiler_times_runs_of_consecutive_steps(torch_gpu):
    """jkpm_profile_enable(16): runs of up to 8 consecutive step launches,
    one started every 16 steps, never at the first step and closed before a
    reduction.  M = 130 is 129 steps with reductions after steps 63, 127 and
    129: runs of 8 start at steps 17, 33, 49, 65, 81, 97 and 113, and a run
    of 1 at step 129, closed by the last reduction: 57 launches.  The spans
    must fit inside the whole call."""
    torch = torch_gpu
    A = honeycomb(32, seed=4)
    h = A.device("c64", 1.0 / A.bounds[1])
    v = torch.ones((A.shape[0], 16), dtype=torch.complex64, device="cuda")
    h.moments(v, 130)
    torch.cuda.synchronize()
    nat.profile_read(reset=True)
    nat.profile_enable(16)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    h.moments(v, 130)
    t1.record()
    t1.synchronize()
    nat.profile_enable(0)
    ms, launches = nat.profile_read(reset=True)
    assert launches == 57
    assert 0.0 < ms < t0.elapsed_time(t1)
    nat.profile_enable(0)
    assert nat.profile_read(reset=True) == (0.0, 0)
