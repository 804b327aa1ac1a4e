"""Failure detection on the GPU path: non-finite moments are reported (NaN
inputs, a spectral bound below the largest |eigenvalue|), the device check
finds planted NaN/Inf, and the allocation-failure hook (JKPM_FAULT=alloc)
reaches the caller as JKPM_ERR_ALLOC without touching device memory.  Every
case is an ordinary input: no kernel faults or hangs."""
import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _native as nat

pytestmark = pytest.mark.gpu


def lattice(L=16, W=1.0, seed=1):
    H, _ = jcl.square_lattice(L, W=W, seed=seed)
    return jcl.ell_matrix(H)


def test_check_finite_finds_planted_values(torch_gpu):
    torch = torch_gpu
    x = torch.zeros((4, 50, 3), dtype=torch.complex128, device="cuda")
    nat.check_finite(x)  # all finite: no error
    x[2, 17, 1] = complex(float("nan"), 0.0)
    x[3, 0, 0] = complex(0.0, float("inf"))
    with pytest.raises(FloatingPointError, match=r"2 non-finite values.*\(2, 17, 1\)"):
        nat.check_finite(x, "probe")
    y = torch.zeros(1000, dtype=torch.float64, device="cuda")
    y[999] = float("-inf")
    with pytest.raises(FloatingPointError, match=r"\(999,\)"):
        nat.check_finite(y)


def test_nan_input_vector_is_reported(torch_gpu):
    A = lattice()
    r = jcl.random_vector_generator(A.shape[0], 3)
    r[5] = np.nan
    with pytest.raises(FloatingPointError, match="non-finite"):
        jcl.rec_A_tab(32, A.modifier(jcl.modifier_bounds), r)
    with pytest.raises(FloatingPointError, match="non-finite"):
        jcl.kpm_dos(A, 32, r, dtype="c64")


def test_bound_too_small_is_reported(torch_gpu):
    """Bounds at half the spectrum: T_n(H/B) grows like cosh(n acosh 2) and
    overflows complex64 within a few hundred moments."""
    A = lattice(W=0.0)
    A.bounds = A.bounds / 2
    r = jcl.random_vector_generator(A.shape[0], 4)
    with pytest.raises(FloatingPointError, match="spectral bound"):
        jcl.kpm_moments(A, 1024, r, dtype="c64")
    A.bounds = A.bounds * 2  # the right bounds work again
    mu = jcl.kpm_moments(A, 1024, r, dtype="c64").cpu().numpy()
    assert np.all(np.isfinite(mu))


def test_allocation_failure_hook(torch_gpu, monkeypatch):
    A = lattice()
    monkeypatch.setenv("JKPM_FAULT", "alloc")
    with pytest.raises(nat.JkpmError) as ei:
        nat.DeviceELL(A.data, A.indices, A.shape[0], A.len_row, 1.0, "c128")
    assert ei.value.code == nat.JKPM_ERR_ALLOC
    assert "JKPM_FAULT" in str(ei.value)
    monkeypatch.delenv("JKPM_FAULT")
    h = nat.DeviceELL(A.data, A.indices, A.shape[0], A.len_row, 1.0, "c128")
    assert h.n == A.shape[0] and h.handle
