"""jkpm_moments_noalloc (library-owned scratch, no allocation) against
jkpm_moments (caller workspace), called through the C-ABI: the same kernels
on the same layout give bitwise-equal moments.  Two calls in a row, the
second on a smaller input, so the second reuses scratch the first left
behind; a third with a second left set; and a size beyond the scratch is
refused with JKPM_ERR_WORKSPACE."""
import ctypes

import numpy as np
import pytest

import minijclsquant_amd as jcl
from minijclsquant_amd import _native as nat

pytestmark = pytest.mark.gpu


def _call_noalloc(h, vr, vl0, vl1, M):
    torch = nat.require_device()
    R = vr.shape[1]
    mu = torch.empty((2 if vl1 is not None else 1, M, R), dtype=torch.complex128, device="cuda")
    rc = nat.lib().jkpm_moments_noalloc(h.handle, vr.data_ptr(), vl0.data_ptr(),
                                        vl1.data_ptr() if vl1 is not None else None, R, M,
                                        mu.data_ptr(), nat.stream_ptr(None))
    return rc, mu


def _case(torch, L, R, M, dtype, seed):
    H, _ = jcl.square_lattice(L, W=1.0, seed=seed)
    A = jcl.ell_matrix(H)
    h = A.device(dtype, 1 / A.bounds[1])
    r = np.stack([jcl.random_vector_generator(A.shape[0], seed + s) for s in range(R)])
    tdt = torch.complex128 if dtype == "c128" else torch.complex64
    vr = torch.from_numpy(r.T.copy()).to(device="cuda", dtype=tdt)
    return h, vr


def test_noalloc_matches_workspace_entry(torch_gpu):
    torch = torch_gpu
    assert nat.lib().jkpm_moments_noalloc_setup() == nat.JKPM_OK
    cases = [(64, 4, 300, "c128", 1), (23, 3, 77, "c128", 7), (40, 16, 129, "c64", 3)]
    for L, R, M, dtype, seed in cases:
        h, vr = _case(torch, L, R, M, dtype, seed)
        rc, got = _call_noalloc(h, vr, vr, None, M)
        assert rc == nat.JKPM_OK, nat.lib().jkpm_last_error()
        want = h.moments(vr, M)
        torch.cuda.synchronize()
        assert torch.equal(got, want), (L, R, M, dtype)
    # a second left set
    h, vr = _case(torch, 30, 2, 50, "c128", 11)
    vl1 = torch.flip(vr, dims=[0]).contiguous()
    rc, got = _call_noalloc(h, vr, vr, vl1, 50)
    assert rc == nat.JKPM_OK, nat.lib().jkpm_last_error()
    want = h.moments(vr, 50, v_l0=vr, v_l1=vl1)
    assert torch.equal(got, want)


def test_noalloc_refuses_sizes_beyond_scratch(torch_gpu):
    torch = torch_gpu
    need = ctypes.c_size_t()
    h, vr = _case(torch, 1024, 8, 4, "c128", 5)  # 2 x 128 MiB of vectors alone
    assert nat.lib().jkpm_moments_workspace(h.handle, 8, 1, ctypes.byref(need)) == 0
    assert need.value > nat.JKPM_NOALLOC_SCRATCH_BYTES
    rc, _ = _call_noalloc(h, vr, vr, None, 4)
    assert rc == nat.JKPM_ERR_WORKSPACE
    assert b"use jkpm_moments with a workspace" in nat.lib().jkpm_last_error()
