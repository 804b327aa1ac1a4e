"""ctypes binding of libjclkpm_hip.so (the C-ABI declared in include/jclkpm.h).

This is the only route from Python to the compute path.  There is no CPU
fallback: when the library is missing or no HIP device is visible, every call
raises.  Device buffers are PyTorch-ROCm tensors; the library receives their
raw pointers and the current HIP stream.

The reference binds its CUDA library the same way, through a thin extension
over ``extern "C"`` symbols (jclsquant/cuda_cython/obv.pyx:9-49).
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libjclkpm_hip.so"

JKPM_OK = 0
JKPM_ERR_ARG = 1
JKPM_ERR_HIP = 2
JKPM_ERR_ALLOC = 3
JKPM_ERR_WORKSPACE = 4

JKPM_C128 = 0
JKPM_C64 = 1

JKPM_HERMITIAN_RTOL = HERMITIAN_RTOL = 1e-10

JKPM_NOALLOC_SCRATCH_BYTES = 128 << 20

JKPM_VOP_AXPBY = 0
JKPM_VOP_MUL = 1
JKPM_VOP_MUL2 = 2
JKPM_VOP_MUL3 = 3

JKPM_KERNEL_NONE = 0
JKPM_KERNEL_JACKSON = 1

_c = ctypes
_vp = _c.c_void_p
_int = _c.c_int
_dbl = _c.c_double
_size = _c.c_size_t
_dp = _c.POINTER(_c.c_double)
_ip = _c.POINTER(_c.c_int)

# name -> (restype, argtypes); every symbol include/jclkpm.h declares.
SIGNATURES = {
    "jkpm_version": (_int, []),
    "jkpm_last_error": (_c.c_char_p, []),
    "rec_A_tab": (_int, [_vp, _vp, _int, _int, _vp, _vp, _int]),
    "rec_A_tab_2": (_int, [_vp, _vp, _int, _int, _vp, _vp, _vp, _int]),
    "jkpm_ell_create": (_int, [_vp, _vp, _int, _int, _dbl, _int, _int, _vp, _c.POINTER(_vp)]),
    "jkpm_ell_destroy": (_int, [_vp]),
    "jkpm_ell_info": (_int, [_vp, _ip, _ip, _ip]),
    "jkpm_ell_set_values": (_int, [_vp, _vp, _dbl, _int, _vp]),
    "jkpm_ell_get_values": (_int, [_vp, _vp, _int, _vp]),
    "jkpm_ell_hermiticity": (_int, [_vp, _dp, _dp, _vp]),
    "jkpm_ell_is_real": (_int, [_vp, _ip]),
    "jkpm_ell_is_imag": (_int, [_vp, _ip]),
    "jkpm_ell_use_real_values": (_int, [_vp, _int]),
    "jkpm_ell_use_left_phases": (_int, [_vp, _int]),
    "jkpm_phase_pack": (_int, [_vp, _c.c_longlong, _vp, _vp, _vp]),
    "jkpm_check_finite": (_int, [_vp, _c.c_longlong, _c.POINTER(_c.c_longlong),
                                 _c.POINTER(_c.c_longlong), _vp]),
    "jkpm_zvec": (_int, [_int, _c.c_longlong, _dp, _dp, _vp, _vp, _vp, _vp, _vp]),
    "jkpm_moments_noalloc_setup": (_int, []),
    "jkpm_profile_read_kind": (_int, [_int, _dp, _c.POINTER(_c.c_longlong), _int]),
    "jkpm_moments_noalloc": (_int, [_vp, _vp, _vp, _vp, _int, _int, _vp, _vp]),
    "jkpm_zdot_workspace": (_int, [_c.POINTER(_size)]),
    "jkpm_zdot": (_int, [_c.c_longlong, _dp, _vp, _vp, _int, _vp, _vp, _size, _vp]),
    "jkpm_ell_create_csr": (_int, [_vp, _vp, _vp, _int, _int, _dbl, _int, _int, _vp,
                                   _c.POINTER(_vp)]),
    "jkpm_ell_get_indices": (_int, [_vp, _vp, _int, _vp]),
    "jkpm_tiles_create": (_int, [_vp, _int, _int, _c.POINTER(_vp)]),
    "jkpm_tiles_destroy": (_int, [_vp]),
    "jkpm_tiles_plan_check": (_int, [_int, _int, _vp, _int, _ip, _c.POINTER(_c.c_longlong), _ip]),
    "jkpm_tiles_info": (_int, [_vp, _ip, _c.POINTER(_c.c_longlong), _ip, _ip]),
    "jkpm_tiles_refresh": (_int, [_vp, _vp]),
    "jkpm_tiles_workspace": (_int, [_vp, _int, _int, _int, _c.POINTER(_size)]),
    "jkpm_tiles_moments": (_int, [_vp, _vp, _int, _int, _int, _vp, _vp, _size, _vp]),
    "jkpm_wave_create": (_int, [_vp, _int, _int, _c.POINTER(_vp)]),
    "jkpm_wave_destroy": (_int, [_vp]),
    "jkpm_wave_info": (_int, [_vp, _ip, _ip, _c.POINTER(_c.c_longlong),
                              _c.POINTER(_c.c_longlong)]),
    "jkpm_wave_plan_check": (_int, [_int, _int, _vp, _int, _int, _int, _int, _ip,
                                    _c.POINTER(_c.c_longlong), _c.POINTER(_c.c_longlong)]),
    "jkpm_wave_workspace": (_int, [_vp, _int, _int, _c.POINTER(_size)]),
    "jkpm_wave_moments": (_int, [_vp, _vp, _int, _int, _vp, _vp, _size, _vp]),
    "jkpm_ell_peierls": (_int, [_vp, _vp, _vp, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _vp]),
    "jkpm_aAxby": (_int, [_vp, _dp, _dp, _vp, _vp, _int, _vp]),
    "jkpm_moments_workspace": (_int, [_vp, _int, _int, _c.POINTER(_size)]),
    "jkpm_moments": (_int, [_vp, _vp, _vp, _vp, _int, _int, _vp, _vp, _size, _vp]),
    "jkpm_dos": (_int, [_vp, _int, _int, _vp, _int, _int, _vp, _vp, _size, _vp]),
    "jkpm_profile_enable": (_int, [_int]),
    "jkpm_profile_read": (_int, [_dp, _c.POINTER(_c.c_longlong), _int]),
    "jkpm_copy": (_int, [_vp, _vp, _size, _vp]),
    "jkpm_ell_create_like": (_int, [_vp, _vp, _dbl, _int, _int, _vp, _c.POINTER(_vp)]),
    "jkpm_moments_keep": (_int, [_vp, _vp, _int, _int, _vp, _vp, _size, _vp, _size, _vp]),
    "jkpm_grad_kept_workspace": (_int, [_vp, _int, _int, _c.POINTER(_size)]),
    "jkpm_moments_grad_kept": (_int, [_vp, _vp, _int, _int, _vp, _vp, _vp, _vp, _size, _vp]),
    "jkpm_kubo_workspace": (_int, [_vp, _int, _int, _c.POINTER(_size)]),
    "jkpm_kubo_mu_nm": (_int, [_vp, _vp, _vp, _int, _int, _vp, _vp, _size, _vp]),
    "jkpm_kubo_sigma": (_int, [_vp, _int, _vp, _int, _int, _dbl, _vp, _vp]),
    "jkpm_grad_workspace": (_int, [_vp, _int, _int, _int, _c.POINTER(_size)]),
    "jkpm_moments_grad": (_int, [_vp, _vp, _int, _int, _vp, _int, _vp, _vp, _size, _vp]),
    "jkpm_dos_grad": (_int, [_vp, _int, _int, _vp, _int, _int, _vp, _vp]),
    "jkpm_batch_create": (_int, [_vp, _int, _vp, _c.POINTER(_vp)]),
    "jkpm_batch_create_strided": (_int, [_vp, _int, _vp, _vp, _c.POINTER(_vp)]),
    "jkpm_batch_destroy": (_int, [_vp]),
    "jkpm_batch_workspace": (_int, [_vp, _int, _int, _c.POINTER(_size)]),
    "jkpm_batch_moments": (_int, [_vp, _vp, _vp, _vp, _int, _int, _vp, _vp, _size, _vp]),
    "jkpm_graph_begin": (_int, [_vp]),
    "jkpm_graph_end": (_int, [_vp, _c.POINTER(_vp)]),
    "jkpm_graph_launch": (_int, [_vp, _vp]),
    "jkpm_graph_destroy": (_int, [_vp]),
    "jkpm_ell_set_launch": (_int, [_vp, _int, _int, _int]),
    "jkpm_ell_get_launch": (_int, [_vp, _ip, _ip, _int]),
    "jkpm_ell_set_pass": (_int, [_vp, _int, _int]),
    "jkpm_ell_get_pass": (_int, [_vp, _ip, _int]),
    "jkpm_autotune_workspace": (_int, [_vp, _int, _int, _int, _c.POINTER(_size)]),
    "jkpm_autotune": (_int, [_vp, _vp, _int, _int, _int, _vp, _size, _vp, _dp]),
    "jkpm_moments_ex_workspace": (_int, [_vp, _vp, _int, _int, _int, _c.POINTER(_size)]),
    "jkpm_moments_ex": (_int, [_vp, _vp, _vp, _int, _int, _int, _vp, _vp, _size, _vp]),
    "jkpm_series_workspace": (_int, [_vp, _int, _int, _c.POINTER(_size)]),
    "jkpm_series": (_int, [_vp, _vp, _vp, _int, _int, _vp, _vp, _vp, _vp, _size, _vp]),
}


def series(h, v_op, v, M: int, c, want_sum: bool = True, want_table: bool = False,
           stream=None):
    """sum_m c_m X_m ([N, R]) and/or the table c_m X_m ([M, N, R]) with
    X = T_m(H) v (v_op None) or the commutator recursion (v_op on H's slots)."""
    torch = _torch()
    R = h._check_vecs(v)
    M = int(M)
    cc = np.ascontiguousarray(np.asarray(c, dtype=np.complex128).reshape(-1))
    if cc.size < M or M <= 0:
        raise ValueError("need M > 0 coefficients")
    ws_n = _size()
    check(lib().jkpm_series_workspace(h.handle, R, 1 if v_op is not None else 0,
                                      _c.byref(ws_n)))
    ws = torch.empty(int(ws_n.value), dtype=torch.uint8, device=v.device)
    out = torch.empty_like(v) if want_sum else None
    tab = torch.empty((M,) + tuple(v.shape), dtype=v.dtype, device=v.device) if want_table else None
    check(lib().jkpm_series(h.handle, v_op.handle if v_op is not None else None, v.data_ptr(), R,
                            M, cc.ctypes.data, out.data_ptr() if out is not None else None,
                            tab.data_ptr() if tab is not None else None, ws.data_ptr(),
                            ws.numel(), stream_ptr(stream)))
    return out, tab

JKPM_MOMENTS_DOUBLING = 1
# step-kernel launch variants (jkpm_ell_set_launch)
JKPM_STEP_VARIANTS = 48


def moments_ex_workspace_bytes(h, b, R: int, M: int, doubling: bool) -> int:
    need = _size()
    check(lib().jkpm_moments_ex_workspace(h.handle if h is not None else None,
                                          b._b if b is not None else None, int(R), int(M),
                                          JKPM_MOMENTS_DOUBLING if doubling else 0,
                                          _c.byref(need)))
    return int(need.value)


def moments_ex(h, b, v_r, M: int, doubling: bool, workspace=None, out=None, stream=None):
    """Single-left-set moments through jkpm_moments_ex: handle h ([M, R]) or
    batch b ([G, M, R]); doubling selects the two-moments-per-step form."""
    torch = _torch()
    H = b.H if b is not None else h
    R = H._check_vecs(v_r)
    M = int(M)
    if M <= 0:
        raise ValueError("M must be positive")
    flags = JKPM_MOMENTS_DOUBLING if doubling else 0
    hp = h.handle if h is not None else None
    bp = b._b if b is not None else None
    need = _size()
    check(lib().jkpm_moments_ex_workspace(hp, bp, R, M, flags, _c.byref(need)))
    if workspace is None or workspace.numel() < need.value:
        workspace = torch.empty(int(need.value), dtype=torch.uint8, device=v_r.device)
    shape = (b.G, M, R) if b is not None else (M, R)
    if out is None:
        out = torch.empty(shape, dtype=torch.complex128, device=v_r.device)
    elif tuple(out.shape) != shape or out.dtype != torch.complex128 or not out.is_contiguous():
        raise ValueError(f"out must be contiguous complex128 {list(shape)}")
    check(lib().jkpm_moments_ex(hp, bp, v_r.data_ptr(), R, M, flags, out.data_ptr(),
                                workspace.data_ptr(), workspace.numel(), stream_ptr(stream)))
    return out


class JkpmError(RuntimeError):
    """A jclkpm call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"jclkpm error {code}: {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()


def load_library(path: Path | None = None) -> ctypes.CDLL:
    """Loads the HIP library (no GPU needed to load it) and declares symbols."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise RuntimeError(
                f"{p} is not built: run `python -m minijclsquant_amd._build` "
                "(or __graft_entry__.build()) first"
            )
        lib = ctypes.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def lib() -> ctypes.CDLL:
    return load_library()


def check(rc: int) -> None:
    if rc != JKPM_OK:
        msg = lib().jkpm_last_error()
        msg = msg.decode(errors="replace") if msg else ""
        if rc == JKPM_ERR_ARG:
            raise ValueError(f"jclkpm: {msg}")
        raise JkpmError(rc, msg)


# ---------------------------------------------------------------- torch glue

def _torch():
    import torch

    return torch


def require_device():
    """Returns torch after checking that a HIP device is visible."""
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError(
            "no HIP device visible: the jclkpm compute path runs on an MI355X GPU only"
        )
    lib()
    return torch


def stream_ptr(stream=None) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def vtype_of(dtype) -> int:
    torch = _torch()
    if dtype in (torch.complex128, np.complex128, "c128", "complex128"):
        return JKPM_C128
    if dtype in (torch.complex64, np.complex64, "c64", "complex64"):
        return JKPM_C64
    raise ValueError(f"unsupported value type {dtype!r}; use complex128 or complex64")


def torch_cdtype(vtype: int):
    torch = _torch()
    return torch.complex128 if vtype == JKPM_C128 else torch.complex64


class DeviceELL:
    """Device-resident, slot-major ELL matrix behind a jkpm_ell handle.

    ``data`` is the reference's row-major complex128 ELL value array
    (``ell_matrix.data``), ``indices`` its int32 column array; both may be
    numpy arrays (uploaded) or contiguous torch device tensors.  The values are
    multiplied by ``scale`` in complex128 and stored in ``dtype``.
    """

    def __init__(self, data, indices, n_rows: int, len_row: int, scale: float = 1.0,
                 dtype="c128", stream=None):
        torch = require_device()
        self.n = int(n_rows)
        self.k = int(len_row)
        self.vtype = vtype_of(dtype)
        self.scale = float(scale)
        self._h = _vp()
        on_dev = isinstance(data, torch.Tensor)
        if on_dev:
            if not (data.is_cuda and indices.is_cuda):
                raise ValueError("torch inputs must be device tensors")
            if data.dtype != torch.complex128 or indices.dtype != torch.int32:
                raise ValueError("device ELL input must be complex128 values and int32 indices")
            data = data.contiguous()
            indices = indices.contiguous()
            if data.numel() != self.n * self.k or indices.numel() != self.n * self.k:
                raise ValueError("data/indices must hold n_rows*len_row entries")
            dp, ip = data.data_ptr(), indices.data_ptr()
            keep = (data, indices)
        else:
            data = np.ascontiguousarray(data, dtype=np.complex128)
            indices = np.ascontiguousarray(indices, dtype=np.int32)
            if data.size != self.n * self.k or indices.size != self.n * self.k:
                raise ValueError("data/indices must hold n_rows*len_row entries")
            dp, ip = data.ctypes.data, indices.ctypes.data
            keep = (data, indices)
        check(lib().jkpm_ell_create(dp, ip, self.n, self.k, self.scale, self.vtype,
                                    1 if on_dev else 0, stream_ptr(stream), _c.byref(self._h)))
        del keep

    @classmethod
    def like(cls, pattern: "DeviceELL", data, scale: float = 1.0, dtype="c64", stream=None):
        """A handle with new row-major complex128 values on `pattern`'s slot
        pattern (shares its column indices), e.g. a velocity operator."""
        torch = require_device()
        self = cls.__new__(cls)
        self.n, self.k = pattern.n, pattern.k
        self.vtype = vtype_of(dtype)
        self.scale = float(scale)
        self._h = _vp()
        if isinstance(data, torch.Tensor):
            d = data.contiguous()
            if not d.is_cuda or d.dtype != torch.complex128 or d.numel() != self.n * self.k:
                raise ValueError("values must be a complex128 device tensor of n_rows*len_row")
            ptr, on_dev = d.data_ptr(), 1
        else:
            d = np.ascontiguousarray(data, dtype=np.complex128)
            if d.size != self.n * self.k:
                raise ValueError("values must hold n_rows*len_row entries")
            ptr, on_dev = d.ctypes.data, 0
        check(lib().jkpm_ell_create_like(pattern._h, ptr, self.scale, self.vtype, on_dev,
                                         stream_ptr(stream), _c.byref(self._h)))
        return self

    @classmethod
    def from_csr(cls, indptr, indices, data, n_rows: int, scale: float = 1.0, dtype="c128",
                 len_row: int = 0, stream=None):
        """Device-side CSR -> ELL ingest (jkpm_ell_create_csr): the row-major
        host ELL of ell_matrix.csr_to_ell (jclsquant/ell_matrix.py:187-200) is
        never built.  indptr int32 [n_rows+1], indices int32 [nnz], data
        complex128 [nnz]: numpy arrays (uploaded) or contiguous torch device
        tensors.  len_row 0 takes the longest row."""
        torch = require_device()
        self = cls.__new__(cls)
        self.n = int(n_rows)
        self.vtype = vtype_of(dtype)
        self.scale = float(scale)
        self._h = _vp()
        on_dev = isinstance(indptr, torch.Tensor)
        if on_dev:
            for t, dt in ((indptr, torch.int32), (indices, torch.int32), (data, torch.complex128)):
                if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != dt:
                    raise ValueError("device CSR input must be int32 indptr/indices and "
                                     "complex128 data, all device tensors")
            keep = (indptr.contiguous(), indices.contiguous(), data.contiguous())
            if keep[0].numel() != self.n + 1:
                raise ValueError("indptr must hold n_rows+1 entries")
            nnz = int(keep[0][-1].item())
            if keep[1].numel() < nnz or keep[2].numel() < nnz:
                raise ValueError("indices/data must hold indptr[n_rows] entries")
            ptrs = [t.data_ptr() for t in keep]
        else:
            keep = (np.ascontiguousarray(indptr, dtype=np.int32),
                    np.ascontiguousarray(indices, dtype=np.int32),
                    np.ascontiguousarray(data, dtype=np.complex128))
            if keep[0].size != self.n + 1:
                raise ValueError("indptr must hold n_rows+1 entries")
            nnz = int(keep[0][-1])
            if keep[1].size < nnz or keep[2].size < nnz:
                raise ValueError("indices/data must hold indptr[n_rows] entries")
            ptrs = [t.ctypes.data for t in keep]
        check(lib().jkpm_ell_create_csr(ptrs[0], ptrs[1], ptrs[2], self.n, int(len_row),
                                        self.scale, self.vtype, 1 if on_dev else 0,
                                        stream_ptr(stream), _c.byref(self._h)))
        del keep
        n, k, vt = _c.c_int(), _c.c_int(), _c.c_int()
        check(lib().jkpm_ell_info(self._h, _c.byref(n), _c.byref(k), _c.byref(vt)))
        self.k = int(k.value)
        return self

    @property
    def handle(self):
        return self._h

    @property
    def dtype(self):
        return torch_cdtype(self.vtype)

    def set_values(self, data, scale: float | None = None, stream=None) -> None:
        torch = _torch()
        sc = self.scale if scale is None else float(scale)
        if isinstance(data, torch.Tensor):
            d = data.contiguous()
            if d.dtype != torch.complex128 or d.numel() != self.n * self.k:
                raise ValueError("values must be complex128 with n_rows*len_row entries")
            check(lib().jkpm_ell_set_values(self._h, d.data_ptr(), sc, 1, stream_ptr(stream)))
        else:
            d = np.ascontiguousarray(data, dtype=np.complex128)
            if d.size != self.n * self.k:
                raise ValueError("values must hold n_rows*len_row entries")
            check(lib().jkpm_ell_set_values(self._h, d.ctypes.data, sc, 0, stream_ptr(stream)))
        self.scale = sc

    def close(self) -> None:
        if self._h:
            lib().jkpm_ell_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ operations

    def _check_vecs(self, *vs):
        torch = _torch()
        R = None
        for v in vs:
            if v is None:
                continue
            if not isinstance(v, torch.Tensor) or not v.is_cuda:
                raise ValueError("vectors must be device tensors")
            if v.dtype != self.dtype:
                raise ValueError(f"vectors must be {self.dtype}, got {v.dtype}")
            if v.dim() != 2 or v.shape[0] != self.n or not v.is_contiguous():
                raise ValueError(f"vectors must be contiguous [N={self.n}, R]")
            if R is None:
                R = v.shape[1]
            elif v.shape[1] != R:
                raise ValueError("all vector batches must have the same R")
        return R

    def workspace_bytes(self, R: int, n_left: int = 1) -> int:
        out = _size()
        check(lib().jkpm_moments_workspace(self._h, int(R), int(n_left), _c.byref(out)))
        return int(out.value)

    def moments(self, v_r, M: int, v_l0=None, v_l1=None, workspace=None, out=None,
                stream=None):
        """mu[l, m, r] = <v_l{l}[:, r] | T_m v_r[:, r]>, complex128 [n_left, M, R].

        v_r, v_l0, v_l1: device tensors [N, R] of the handle's dtype
        (v_l0 defaults to v_r)."""
        torch = _torch()
        if v_l0 is None:
            v_l0 = v_r
        R = self._check_vecs(v_r, v_l0, v_l1)
        M = int(M)
        if M <= 0:
            raise ValueError("M must be positive")
        n_left = 2 if v_l1 is not None else 1
        need = self.workspace_bytes(R, n_left)
        if workspace is None or workspace.numel() < need:
            workspace = torch.empty(need, dtype=torch.uint8, device=v_r.device)
        if out is None:
            out = torch.empty((n_left, M, R), dtype=torch.complex128, device=v_r.device)
        elif out.shape != (n_left, M, R) or out.dtype != torch.complex128 or not out.is_contiguous():
            raise ValueError("out must be contiguous complex128 [n_left, M, R]")
        check(lib().jkpm_moments(self._h, v_r.data_ptr(), v_l0.data_ptr(),
                                 v_l1.data_ptr() if v_l1 is not None else None, R, M,
                                 out.data_ptr(), workspace.data_ptr(), workspace.numel(),
                                 stream_ptr(stream)))
        return out

    def hermiticity(self, stream=None):
        """(max |a_ij - conj(a_ji)|, max |a_ij|) of the stored values, computed
        once per value change on the device (jkpm_ell_hermiticity)."""
        err, mag = _dbl(), _dbl()
        check(lib().jkpm_ell_hermiticity(self._h, _c.byref(err), _c.byref(mag),
                                         stream_ptr(stream)))
        return float(err.value), float(mag.value)

    def is_real(self) -> bool:
        """True when every stored value has a zero imaginary part; the step
        kernels then read a packed real copy (jkpm_ell_is_real)."""
        v = _c.c_int()
        check(lib().jkpm_ell_is_real(self._h, _c.byref(v)))
        return bool(v.value)

    def is_imag(self) -> bool:
        """True when every stored real part is zero and some imaginary part is
        not; the Kubo step then reads the packed imaginary parts (jkpm_ell_is_imag)."""
        v = _c.c_int()
        check(lib().jkpm_ell_is_imag(self._h, _c.byref(v)))
        return bool(v.value)

    def use_real_values(self, on: bool = True) -> None:
        """on=False: the step kernels read the complex values even when they
        are all real (jkpm_ell_use_real_values)."""
        check(lib().jkpm_ell_use_real_values(self._h, 1 if on else 0))
        self._use_real = bool(on)

    def use_left_phases(self, on: bool = True) -> None:
        """on=False: complex64 plain recursions read the left vectors as
        complex numbers even when they are unit complex numbers, whose 32-bit
        phases they read otherwise (jkpm_ell_use_left_phases).  Workspace
        sizes change with it."""
        check(lib().jkpm_ell_use_left_phases(self._h, 1 if on else 0))
        self._left_phases = bool(on)

    def is_hermitian(self, stream=None) -> bool:
        """True when max |a_ij - conj(a_ji)| <= HERMITIAN_RTOL * max |a_ij|, the
        condition the doubling form, the Kubo moments and the adjoint need."""
        err, mag = self.hermiticity(stream)
        return err <= HERMITIAN_RTOL * mag

    def values(self, stream=None):
        """Row-major complex128 values [n*k] (device tensor) in the units they
        were uploaded in."""
        torch = _torch()
        out = torch.empty(self.n * self.k, dtype=torch.complex128, device="cuda")
        check(lib().jkpm_ell_get_values(self._h, out.data_ptr(), 1, stream_ptr(stream)))
        return out

    def indices(self, stream=None):
        """Row-major int32 column indices [n*k] (device tensor)."""
        torch = _torch()
        out = torch.empty(self.n * self.k, dtype=torch.int32, device="cuda")
        check(lib().jkpm_ell_get_indices(self._h, out.data_ptr(), 1, stream_ptr(stream)))
        return out

    def peierls(self, src: "DeviceELL", bonds: "DeviceELL", a=1.0, ax=0.0, ay=0.0,
                stream=None) -> "DeviceELL":
        """self <- a * src * exp(i (ax dx + ay dy)) slot by slot, (dx, dy) the
        real and imaginary parts of bonds' values; self, src and bonds share
        one slot pattern.  Asynchronous on stream."""
        a, ax, ay = complex(a), complex(ax), complex(ay)
        check(lib().jkpm_ell_peierls(self._h, src.handle, bonds.handle, a.real, a.imag, ax.real,
                                     ax.imag, ay.real, ay.imag, stream_ptr(stream)))
        self.scale = src.scale
        return self

    def set_launch(self, variant: int, chunk_mult: int, doubling: bool = False) -> None:
        """Step-kernel launch shape of the plain recursion (or of the
        doubling form with doubling=True)."""
        check(lib().jkpm_ell_set_launch(self._h, int(variant), int(chunk_mult),
                                        JKPM_MOMENTS_DOUBLING if doubling else 0))

    def set_pass(self, vectors: int, doubling: bool = False) -> None:
        """Vectors per recursion pass of the plain recursion (or of the
        doubling form); 0 runs all vectors at once (jkpm_ell_set_pass)."""
        check(lib().jkpm_ell_set_pass(self._h, int(vectors),
                                      JKPM_MOMENTS_DOUBLING if doubling else 0))

    def get_pass(self, doubling: bool = False) -> int:
        v = _c.c_int()
        check(lib().jkpm_ell_get_pass(self._h, _c.byref(v),
                                      JKPM_MOMENTS_DOUBLING if doubling else 0))
        return int(v.value)

    def get_launch(self, doubling: bool = False):
        v, m = _c.c_int(), _c.c_int()
        check(lib().jkpm_ell_get_launch(self._h, _c.byref(v), _c.byref(m),
                                        JKPM_MOMENTS_DOUBLING if doubling else 0))
        return int(v.value), int(m.value)

    def autotune(self, v_r, steps: int = 8, stream=None, doubling: bool = False) -> float:
        """Picks the fastest step-kernel launch shape for these vectors (for
        the doubling form with doubling=True); returns its mean step time (us)."""
        torch = _torch()
        R = self._check_vecs(v_r)
        flags = JKPM_MOMENTS_DOUBLING if doubling else 0
        need = _size()
        check(lib().jkpm_autotune_workspace(self._h, R, int(steps), flags, _c.byref(need)))
        ws = torch.empty(int(need.value), dtype=torch.uint8, device=v_r.device)
        best = _c.c_double()
        check(lib().jkpm_autotune(self._h, v_r.data_ptr(), R, int(steps), flags, ws.data_ptr(),
                                  ws.numel(), stream_ptr(stream), _c.byref(best)))
        return float(best.value)

    def moments_keep(self, v_r, M: int, stack=None, stream=None):
        """(mu [M, R] complex128, stack): the moments <v_r|T_m|v_r> with every
        T_0 .. T_{M-1} left in stack [max(M,2), N, R] (jkpm_moments_keep)."""
        torch = _torch()
        R = self._check_vecs(v_r)
        M = int(M)
        shape = (max(M, 2), self.n, R)
        if stack is None:
            stack = torch.empty(shape, dtype=v_r.dtype, device=v_r.device)
        elif tuple(stack.shape) != shape or stack.dtype != v_r.dtype or not stack.is_contiguous():
            raise ValueError(f"stack must be a contiguous {v_r.dtype} tensor {shape}")
        ws = torch.empty(self.workspace_bytes(R, 1), dtype=torch.uint8, device=v_r.device)
        mu = torch.empty((M, R), dtype=torch.complex128, device=v_r.device)
        check(lib().jkpm_moments_keep(self._h, v_r.data_ptr(), R, M, mu.data_ptr(),
                                      stack.data_ptr(), stack.numel() * stack.element_size(),
                                      ws.data_ptr(), ws.numel(), stream_ptr(stream)))
        return mu, stack

    def grad_values_kept(self, r, M: int, grad_mu, stack, workspace=None, stream=None):
        """grad_values from the vectors moments_keep left in stack: the
        adjoint without a forward pass (jkpm_moments_grad_kept).  workspace:
        optional uint8 device tensor of at least the required size."""
        torch = _torch()
        R = self._check_vecs(r)
        M = int(M)
        g = grad_mu.to(dtype=torch.complex128).contiguous()
        if g.shape != (M, R) or not g.is_cuda:
            raise ValueError(f"grad_mu must be a complex128 device tensor [M={M}, R={R}]")
        if tuple(stack.shape) != (max(M, 2), self.n, R) or not stack.is_contiguous():
            raise ValueError("stack must be the one moments_keep filled for these r and M")
        need = _size()
        check(lib().jkpm_grad_kept_workspace(self._h, R, M, _c.byref(need)))
        ws = workspace
        if ws is None or ws.numel() < int(need.value):
            ws = torch.empty(int(need.value), dtype=torch.uint8, device=r.device)
        out = torch.empty(self.n * self.k, dtype=torch.complex128, device=r.device)
        check(lib().jkpm_moments_grad_kept(self._h, r.data_ptr(), R, M, g.data_ptr(),
                                           stack.data_ptr(), out.data_ptr(), ws.data_ptr(),
                                           ws.numel(), stream_ptr(stream)))
        return out

    def grad_workspace_bytes(self, R: int, M: int, seg: int = 0) -> int:
        out = _size()
        check(lib().jkpm_grad_workspace(self._h, int(R), int(M), int(seg), _c.byref(out)))
        return int(out.value)

    def grad_values(self, r, M: int, grad_mu, seg: int = 0, workspace=None, stream=None):
        """dL/dRe(values) + i dL/dIm(values) for the row-major complex128
        values this handle was made from, given grad_mu [M, R] (complex128)
        of mu = moments(r, M)[0].  seg: checkpoint spacing (0: sqrt(M-1));
        < 0 keeps every T_n (no recomputation, M-1 vector batches).
        workspace: optional uint8 device tensor of at least the required size."""
        torch = _torch()
        R = self._check_vecs(r)
        M = int(M)
        g = grad_mu.to(dtype=torch.complex128).contiguous()
        if g.shape != (M, R) or not g.is_cuda:
            raise ValueError(f"grad_mu must be a complex128 device tensor [M={M}, R={R}]")
        ws = workspace
        if ws is None or ws.numel() < self.grad_workspace_bytes(R, M, seg):
            ws = torch.empty(self.grad_workspace_bytes(R, M, seg), dtype=torch.uint8,
                             device=r.device)
        out = torch.empty(self.n * self.k, dtype=torch.complex128, device=r.device)
        check(lib().jkpm_moments_grad(self._h, r.data_ptr(), R, M, g.data_ptr(), int(seg),
                                      out.data_ptr(), ws.data_ptr(), ws.numel(),
                                      stream_ptr(stream)))
        return out

    def aAxby(self, a: complex, b: complex, x, y, stream=None) -> None:
        """y <- a*A*x + b*y on device tensors [N, R] (in place on y)."""
        R = self._check_vecs(x, y)
        if x.data_ptr() == y.data_ptr():
            raise ValueError("x must not alias y")
        av = (_c.c_double * 2)(complex(a).real, complex(a).imag)
        bv = (_c.c_double * 2)(complex(b).real, complex(b).imag)
        check(lib().jkpm_aAxby(self._h, av, bv, x.data_ptr(), y.data_ptr(), R,
                               stream_ptr(stream)))


def dos(mu, energies_scaled, kernel: int = JKPM_KERNEL_JACKSON, stream=None, out=None,
        workspace=None, check_range: bool = True):
    """DOS on device from complex128 moments mu [M, R] (averaged over R).
    out (float64 [n_e]) and workspace (float64 [M]) may be given, and the
    range check of the energies (a device-to-host read) skipped, so that the
    call can be captured into a HIP graph."""
    torch = _torch()
    if mu.dtype != torch.complex128 or mu.dim() != 2 or not mu.is_contiguous():
        raise ValueError("mu must be contiguous complex128 [M, R]")
    e = energies_scaled
    if e.device != mu.device or e.dtype != torch.float64 or not e.is_contiguous():
        e = e.to(device=mu.device, dtype=torch.float64).contiguous()
    if check_range and e.numel() and float(e.abs().max()) >= 1.0:
        raise ValueError("rescaled energies must lie strictly inside (-1, 1)")
    M, R = mu.shape
    if out is None:
        out = torch.empty(e.numel(), dtype=torch.float64, device=mu.device)
    elif out.numel() != e.numel() or out.dtype != torch.float64 or not out.is_contiguous():
        raise ValueError("out must be contiguous float64 with one entry per energy")
    ws = workspace
    if ws is None:
        ws = torch.empty(M, dtype=torch.float64, device=mu.device)
    elif ws.numel() < M or ws.dtype != torch.float64:
        raise ValueError("workspace must be float64 with at least M entries")
    check(lib().jkpm_dos(mu.data_ptr(), M, R, e.data_ptr(), e.numel(), int(kernel),
                         out.data_ptr(), ws.data_ptr(), ws.numel() * 8, stream_ptr(stream)))
    return out


def tiles_plan_check(indices, n_rows: int, len_row: int, cap: int):
    """Host-only build + full check of a tile plan for a row-major ELL column
    array; returns {npatch, halo_rows, max_ext}.  Raises ValueError when the
    plan is inconsistent."""
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.int32).reshape(n_rows, len_row).T)
    npatch, mext = _c.c_int(), _c.c_int()
    halo = _c.c_longlong()
    check(lib().jkpm_tiles_plan_check(int(n_rows), int(len_row), idx.ctypes.data, int(cap),
                                      _c.byref(npatch), _c.byref(halo), _c.byref(mext)))
    return {"npatch": npatch.value, "halo_rows": halo.value, "max_ext": mext.value}


def wave_plan_check(indices, n_rows: int, len_row: int, nchunks: int, rpc: int,
                    rows_per_wave: int, max_lag: int = 1):
    """Host-only build + independent check of a wavefront plan for a row-major
    ELL column array; returns {lag, tail_rows, fixup_rows}.  Raises ValueError when the
    columns do not fit the wavefront or the plan is inconsistent."""
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.int32).reshape(n_rows, len_row).T)
    lag = _c.c_int()
    tail, fix = _c.c_longlong(), _c.c_longlong()
    check(lib().jkpm_wave_plan_check(int(n_rows), int(len_row), idx.ctypes.data, int(nchunks),
                                     int(rpc), int(rows_per_wave), int(max_lag),
                                     _c.byref(lag), _c.byref(tail), _c.byref(fix)))
    return {"lag": lag.value, "tail_rows": tail.value, "fixup_rows": fix.value}


class WavePlan:
    """Two Chebyshev steps per launch by a wavefront over the row chunks of a
    complex64 DeviceELL (jkpm_wave_*): R = 8 or 16 vectors, ELL width 4 or 5,
    one form (plain or doubling) per plan.  The handle must outlive the plan.
    moments() synchronises its stream (it reads the kernel's error flag) and
    cannot be captured into a graph; one call at a time per plan (the plan's
    progress flags are shared by its calls)."""

    def __init__(self, H: DeviceELL, R: int, doubling: bool = False):
        require_device()
        self.H = H
        self.R = int(R)
        self.doubling = bool(doubling)
        self._w = _vp()
        check(lib().jkpm_wave_create(H.handle, self.R, JKPM_MOMENTS_DOUBLING if doubling else 0,
                                     _c.byref(self._w)))

    def info(self):
        nch, lag = _c.c_int(), _c.c_int()
        tail, fix = _c.c_longlong(), _c.c_longlong()
        check(lib().jkpm_wave_info(self._w, _c.byref(nch), _c.byref(lag), _c.byref(tail),
                                   _c.byref(fix)))
        return {"nchunks": nch.value, "lag": lag.value, "tail_rows": tail.value,
                "fixup_rows": fix.value}

    def workspace_bytes(self, M: int) -> int:
        out = _size()
        check(lib().jkpm_wave_workspace(self._w, self.R, int(M), _c.byref(out)))
        return int(out.value)

    def moments(self, v_r, M: int, workspace=None, out=None, stream=None):
        """mu [M, R] complex128 of <r|T_m|r> (device)."""
        torch = _torch()
        R = self.H._check_vecs(v_r)
        if R != self.R:
            raise ValueError(f"the plan is for R = {self.R} vectors, got {R}")
        M = int(M)
        if M <= 0:
            raise ValueError("M must be positive")
        need = self.workspace_bytes(M)
        if workspace is None or workspace.numel() < need:
            workspace = torch.empty(need, dtype=torch.uint8, device=v_r.device)
        if out is None:
            out = torch.empty((M, R), dtype=torch.complex128, device=v_r.device)
        elif tuple(out.shape) != (M, R) or out.dtype != torch.complex128 or not out.is_contiguous():
            raise ValueError("out must be contiguous complex128 [M, R]")
        check(lib().jkpm_wave_moments(self._w, v_r.data_ptr(), R, M, out.data_ptr(),
                                      workspace.data_ptr(), workspace.numel(),
                                      stream_ptr(stream)))
        return out

    def __del__(self):
        w = getattr(self, "_w", None)
        if w is not None and w.value and _lib is not None:
            _lib.jkpm_wave_destroy(w)
            self._w = _vp()


class TilePlan:
    """Rows of a DeviceELL renumbered into LDS patches for the two-steps-per-
    launch kernel (jkpm_tiles_*).  Built for up to R vectors per call; the
    handle must outlive the plan.  After the handle's values change, call
    refresh()."""

    def __init__(self, H: DeviceELL, R: int, max_rows: int = 0):
        require_device()
        self.H = H
        self.R = int(R)
        self._t = _vp()
        check(lib().jkpm_tiles_create(H.handle, self.R, int(max_rows), _c.byref(self._t)))

    def info(self):
        npatch, mext, cap = _c.c_int(), _c.c_int(), _c.c_int()
        halo = _c.c_longlong()
        check(lib().jkpm_tiles_info(self._t, _c.byref(npatch), _c.byref(halo), _c.byref(mext),
                                    _c.byref(cap)))
        return {"npatch": npatch.value, "halo_rows": halo.value, "max_ext": mext.value,
                "cap": cap.value}

    def refresh(self, stream=None):
        check(lib().jkpm_tiles_refresh(self._t, stream_ptr(stream)))

    def workspace_bytes(self, R: int, M: int, doubling: bool = False) -> int:
        out = _size()
        check(lib().jkpm_tiles_workspace(self._t, int(R), int(M),
                                         JKPM_MOMENTS_DOUBLING if doubling else 0,
                                         _c.byref(out)))
        return int(out.value)

    def moments(self, v_r, M: int, doubling: bool = False, workspace=None, out=None,
                stream=None):
        """mu [M, R] complex128 of <r|T_m|r> (device)."""
        torch = _torch()
        R = self.H._check_vecs(v_r)
        M = int(M)
        if M <= 0:
            raise ValueError("M must be positive")
        need = self.workspace_bytes(R, M, doubling)
        if workspace is None or workspace.numel() < need:
            workspace = torch.empty(need, dtype=torch.uint8, device=v_r.device)
        if out is None:
            out = torch.empty((M, R), dtype=torch.complex128, device=v_r.device)
        elif tuple(out.shape) != (M, R) or out.dtype != torch.complex128 or not out.is_contiguous():
            raise ValueError("out must be contiguous complex128 [M, R]")
        check(lib().jkpm_tiles_moments(self._t, v_r.data_ptr(), R, M,
                                       JKPM_MOMENTS_DOUBLING if doubling else 0, out.data_ptr(),
                                       workspace.data_ptr(), workspace.numel(),
                                       stream_ptr(stream)))
        return out

    def __del__(self):
        t = getattr(self, "_t", None)
        if t is not None and t.value and _lib is not None:
            _lib.jkpm_tiles_destroy(t)
            self._t = _vp()


class Batch:
    """A DeviceELL seen as a block-diagonal batch of G realizations with rows
    [offsets[g], offsets[g+1]) (jkpm_batch_*)."""

    def __init__(self, H: DeviceELL, offsets, chunk_rows=None):
        """chunk_rows (optional, per realization): chunks of exactly that many
        rows (0: the proportional cut), jkpm_batch_create_strided."""
        require_device()
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        self.H = H
        self.G = off.size - 1
        self.offsets = off
        self._b = _vp()
        if chunk_rows is None:
            check(lib().jkpm_batch_create(H.handle, self.G, off.ctypes.data, _c.byref(self._b)))
        else:
            cr = np.ascontiguousarray(chunk_rows, dtype=np.int32)
            if cr.size != self.G:
                raise ValueError("chunk_rows needs one entry per realization")
            check(lib().jkpm_batch_create_strided(H.handle, self.G, off.ctypes.data,
                                                  cr.ctypes.data, _c.byref(self._b)))

    def workspace_bytes(self, R: int, n_left: int = 1) -> int:
        out = _size()
        check(lib().jkpm_batch_workspace(self._b, int(R), int(n_left), _c.byref(out)))
        return int(out.value)

    def moments(self, v_r, M: int, v_l0=None, workspace=None, out=None, stream=None):
        """mu [G, M, R] complex128 of every realization (v_l0 defaults to v_r)."""
        torch = _torch()
        if v_l0 is None:
            v_l0 = v_r
        R = self.H._check_vecs(v_r, v_l0)
        M = int(M)
        need = self.workspace_bytes(R, 1)
        if workspace is None or workspace.numel() < need:
            workspace = torch.empty(need, dtype=torch.uint8, device=v_r.device)
        if out is None:
            out = torch.empty((self.G, M, R), dtype=torch.complex128, device=v_r.device)
        elif out.shape != (self.G, M, R) or out.dtype != torch.complex128 or not out.is_contiguous():
            raise ValueError("out must be contiguous complex128 [G, M, R]")
        check(lib().jkpm_batch_moments(self._b, v_r.data_ptr(), v_l0.data_ptr(), None, R, M,
                                       out.data_ptr(), workspace.data_ptr(), workspace.numel(),
                                       stream_ptr(stream)))
        return out

    def close(self):
        if self._b:
            lib().jkpm_batch_destroy(self._b)
            self._b = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """HIP graph of the library launches issued on `stream` inside a
    ``with Graph.capture(stream) as g:`` block; ``g.replay()`` relaunches
    them (same pointers and sizes)."""

    def __init__(self, stream):
        self.stream = stream
        self._g = _vp()

    @classmethod
    def capture(cls, stream):
        return _Capture(cls(stream))

    def replay(self, stream=None):
        check(lib().jkpm_graph_launch(self._g, stream_ptr(stream or self.stream)))

    def close(self):
        if self._g:
            lib().jkpm_graph_destroy(self._g)
            self._g = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Capture:
    def __init__(self, g: Graph):
        self.g = g

    def __enter__(self):
        torch = require_device()
        if self.g.stream is None or self.g.stream == torch.cuda.default_stream():
            raise ValueError("graph capture needs a non-default stream")
        check(lib().jkpm_graph_begin(stream_ptr(self.g.stream)))
        return self.g

    def __exit__(self, et, ev, tb):
        rc = lib().jkpm_graph_end(stream_ptr(self.g.stream), _c.byref(self.g._g))
        if et is None:
            check(rc)
        return False


def dos_grad(w, M: int, R: int, energies_scaled, kernel: int = JKPM_KERNEL_JACKSON, stream=None):
    """Transpose of dos(): complex128 [M, R] gradient of mu from w = dL/d dos."""
    torch = require_device()
    w = w.to(dtype=torch.float64).contiguous()
    e = energies_scaled.to(device=w.device, dtype=torch.float64).contiguous()
    if w.numel() != e.numel():
        raise ValueError("w and energies must have the same length")
    out = torch.empty((int(M), int(R)), dtype=torch.complex128, device=w.device)
    check(lib().jkpm_dos_grad(w.data_ptr(), int(M), int(R), e.data_ptr(), e.numel(), int(kernel),
                              out.data_ptr(), stream_ptr(stream)))
    return out


def kubo_workspace_bytes(H: DeviceELL, R: int, M: int) -> int:
    out = _size()
    check(lib().jkpm_kubo_workspace(H.handle, int(R), int(M), _c.byref(out)))
    return int(out.value)


def kubo_mu_nm(H: DeviceELL, V: DeviceELL, r, M: int, workspace=None, out=None, stream=None):
    """mu[m, n] = (1/R) sum_r <r| v T_m(H) v T_n(H) |r>, complex128 [M, M].
    H, V: complex64 handles (V made with DeviceELL.like(H, ...)); r: device
    complex64 [N, R]."""
    torch = require_device()
    if not (isinstance(r, torch.Tensor) and r.is_cuda and r.dtype == torch.complex64
            and r.dim() == 2 and r.shape[0] == H.n and r.is_contiguous()):
        raise ValueError(f"r must be a contiguous complex64 device tensor [N={H.n}, R]")
    R, M = r.shape[1], int(M)
    need = kubo_workspace_bytes(H, R, M)
    if workspace is None or workspace.numel() < need:
        workspace = torch.empty(need, dtype=torch.uint8, device=r.device)
    if out is None:
        out = torch.empty((M, M), dtype=torch.complex128, device=r.device)
    check(lib().jkpm_kubo_mu_nm(H.handle, V.handle, r.data_ptr(), R, M, out.data_ptr(),
                                workspace.data_ptr(), workspace.numel(), stream_ptr(stream)))
    return out


def kubo_sigma(mu, e_scaled, kind: int, bound: float, stream=None):
    """sigma[i] = sum_mn (a_m b_n + c_m d_n) mu[m, n] at rescaled energies."""
    torch = require_device()
    if mu.dtype != torch.complex128 or mu.dim() != 2 or mu.shape[0] != mu.shape[1]:
        raise ValueError("mu must be complex128 [M, M]")
    mu = mu.contiguous()
    e = torch.as_tensor(e_scaled, dtype=torch.float64, device=mu.device).contiguous()
    if e.numel() and float(e.abs().max()) > 1.0:
        raise ValueError("rescaled energies must lie in [-1, 1]")
    out = torch.empty(e.numel(), dtype=torch.complex128, device=mu.device)
    check(lib().jkpm_kubo_sigma(mu.data_ptr(), mu.shape[0], e.data_ptr(), e.numel(), int(kind),
                                float(bound), out.data_ptr(), stream_ptr(stream)))
    return out


def profile_enable(every: int) -> None:
    """Record HIP events around every `every`-th step launch (0 = off)."""
    check(lib().jkpm_profile_enable(int(every)))


def check_finite(t, what="result", hint="", stream=None) -> None:
    """Raises FloatingPointError when the complex128/float64 device tensor t
    holds NaN or Inf (jkpm_check_finite); JKPM_CHECK_FINITE=0 skips the check."""
    if os.environ.get("JKPM_CHECK_FINITE", "1") == "0":
        return
    torch = _torch()
    if t.dtype not in (torch.complex128, torch.float64) or not t.is_cuda:
        raise ValueError("check_finite takes complex128 or float64 device tensors")
    t = t.contiguous()
    n = t.numel() * (2 if t.dtype == torch.complex128 else 1)
    first, count = _c.c_longlong(), _c.c_longlong()
    check(lib().jkpm_check_finite(t.data_ptr(), n, _c.byref(first), _c.byref(count),
                                  stream_ptr(stream)))
    if count.value:
        idx = int(first.value) // (2 if t.dtype == torch.complex128 else 1)
        pos = tuple(int(i) for i in np.unravel_index(idx, tuple(t.shape)))
        raise FloatingPointError(f"{count.value} non-finite values in the {what}, the first at "
                                 f"index {pos}{hint}")


JKPM_PROF_STEP = 0
JKPM_PROF_KUBO_STEP = 1
JKPM_PROF_KUBO_GEMM = 2
JKPM_PROF_WAVE = 3


def profile_read(reset: bool = True, kind: int = JKPM_PROF_STEP):
    """(summed ms, launches) recorded for one kind of kernel since the last
    reset of that kind."""
    ms = _c.c_double()
    n = _c.c_longlong()
    check(lib().jkpm_profile_read_kind(int(kind), _c.byref(ms), _c.byref(n), 1 if reset else 0))
    return float(ms.value), int(n.value)


def copy(dst, src, nbytes: int, stream=None) -> None:
    check(lib().jkpm_copy(dst.data_ptr(), src.data_ptr(), int(nbytes), stream_ptr(stream)))


def phase_pack(v, stream=None):
    """(phases, ok) of a complex64 device tensor: int32 phases (2 pi = 2^32,
    as the step kernels read them; the bits are those of the unsigned phase)
    and whether every element is reproduced within 2^-21 per component
    (jkpm_phase_pack)."""
    torch = _torch()
    if v.dtype != torch.complex64 or not v.is_cuda:
        raise ValueError("phase_pack takes a complex64 device tensor")
    v = v.contiguous()
    ph = torch.empty(v.shape, dtype=torch.int32, device=v.device)
    bad = torch.empty(1, dtype=torch.int32, device=v.device)
    check(lib().jkpm_phase_pack(v.data_ptr(), v.numel(), ph.data_ptr(), bad.data_ptr(),
                                stream_ptr(stream)))
    return ph, int(bad.item()) == 0


def rec_A_tab_host(data, indices, len_row: int, v, M: int, v_r=None):
    """Calls the reference-compatible host-pointer symbols rec_A_tab /
    rec_A_tab_2 and returns complex128 [M]."""
    require_device()
    data = np.ascontiguousarray(data, dtype=np.complex128)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.complex128)
    N = v.shape[0]
    if data.size != N * len_row or indices.size != N * len_row:
        raise ValueError("data/indices must hold N*len_row entries")
    out = np.zeros(M, dtype=np.complex128)
    if v_r is None:
        check(lib().rec_A_tab(data.ctypes.data, indices.ctypes.data, int(len_row), N,
                              out.ctypes.data, v.ctypes.data, int(M)))
    else:
        v_r = np.ascontiguousarray(v_r, dtype=np.complex128)
        if v_r.shape != v.shape:
            raise ValueError("v_l and v_r must have the same length")
        check(lib().rec_A_tab_2(data.ctypes.data, indices.ctypes.data, int(len_row), N,
                                out.ctypes.data, v.ctypes.data, v_r.ctypes.data, int(M)))
    return out
