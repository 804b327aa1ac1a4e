"""Loading ELL matrices saved with ``ell_matrix.save``.

The reference saves a matrix with ``joblib.dump(self, path)``
(jclsquant/ell_matrix.py:202-207) and loads it with ``joblib.load(path)``
(``load_ell``, jclsquant/hams.py:14-19).  A joblib file is a pickle stream in
which every numpy array is a ``NumpyArrayWrapper`` record followed inline by
the array's raw bytes; optionally the whole file is zlib / gzip / bz2 / xz /
lzma compressed.

``load_ell`` here reads that format without executing anything from the
file: it runs Python's pure-Python unpickler with ``find_class`` restricted to
a fixed list of constructors (numpy arrays, dtypes and scalars, scipy CSR
matrices, a few builtins and the ell_matrix class itself) and reads the inline
array blocks itself, refusing object arrays.  Anything else in the stream, such
as a reference to a function, raises ``pickle.UnpicklingError``.  Both the
reference's class path (``jclsquant.ell_matrix.ell_matrix``) and this
package's load as this package's ``ell_matrix``.
"""
from __future__ import annotations

import bz2
import gzip
import io
import lzma
import os
import pickle
import zlib

import numpy as np
import scipy.sparse as sp

from .ell_matrix import ell_matrix

__all__ = ["load_ell"]


def _restore_ell(state):
    """ell_matrix from its saved state (what ell_matrix.__reduce__ names)."""
    obj = ell_matrix.__new__(ell_matrix)
    obj.__setstate__(state)
    return obj


class _ArrayRecord:
    """Stands for joblib's NumpyArrayWrapper: the pickle stream only sets its
    attributes (subclass, shape, order, dtype, allow_mmap and, from joblib
    1.4 on, numpy_array_alignment_bytes); the array bytes follow inline."""


def _numpy_reconstruct():
    from numpy._core.multiarray import _reconstruct

    return _reconstruct


def _numpy_scalar():
    from numpy._core.multiarray import scalar

    return scalar


def _allowed():
    out = {}
    for mod in ("numpy.core.multiarray", "numpy._core.multiarray"):
        out[(mod, "_reconstruct")] = _numpy_reconstruct()
        out[(mod, "scalar")] = _numpy_scalar()
    out[("numpy", "ndarray")] = np.ndarray
    out[("numpy", "dtype")] = np.dtype
    out[("joblib.numpy_pickle", "NumpyArrayWrapper")] = _ArrayRecord
    for mod in ("scipy.sparse._csr", "scipy.sparse.csr"):
        out[(mod, "csr_matrix")] = sp.csr_matrix
    out[("scipy.sparse._csr", "csr_array")] = sp.csr_array
    for mod in ("jclsquant.ell_matrix", "minijclsquant_amd.ell_matrix"):
        out[(mod, "ell_matrix")] = ell_matrix
    out[("minijclsquant_amd.hams", "_restore_ell")] = _restore_ell
    import copyreg

    out[("copyreg", "_reconstructor")] = copyreg._reconstructor
    for name, obj in (("object", object), ("complex", complex), ("slice", slice),
                      ("frozenset", frozenset), ("set", set), ("bytearray", bytearray)):
        out[("builtins", name)] = obj
        out[("__builtin__", name)] = obj
    return out


class _Unpickler(pickle._Unpickler):
    """Pure-Python unpickler with a fixed constructor list."""

    def __init__(self, fh):
        super().__init__(fh)
        self._ok = _allowed()

    def find_class(self, module, name):
        obj = self._ok.get((module, name))
        if obj is None:
            raise pickle.UnpicklingError(
                f"refusing to load {module}.{name}: not a numpy, scipy CSR or "
                "ell_matrix constructor")
        return obj

    def persistent_load(self, pid):
        raise pickle.UnpicklingError("persistent references are not supported")

    def _read_array(self, rec: _ArrayRecord) -> np.ndarray:
        dtype = np.dtype(rec.dtype)
        if dtype.hasobject:
            raise pickle.UnpicklingError("object arrays are not supported")
        subclass = getattr(rec, "subclass", np.ndarray)
        if subclass not in (np.ndarray, np.memmap):
            raise pickle.UnpicklingError("array subclasses are not supported")
        shape = tuple(int(x) for x in rec.shape)
        count = int(np.prod(shape, dtype=np.int64)) if shape else 1
        if getattr(rec, "numpy_array_alignment_bytes", None) is not None:
            pad = self.read(1)
            if len(pad) != 1:
                raise pickle.UnpicklingError("truncated array padding")
            if pad[0]:
                self.read(pad[0])
        nbytes = count * dtype.itemsize
        buf = self.read(nbytes)
        if len(buf) != nbytes:
            raise pickle.UnpicklingError("truncated array data")
        arr = np.frombuffer(buf, dtype=dtype, count=count).copy()
        if getattr(rec, "order", "C") == "F":
            arr = arr.reshape(shape[::-1]).transpose()
        else:
            arr = arr.reshape(shape)
        if not arr.dtype.isnative:
            arr = arr.astype(arr.dtype.newbyteorder("="))
        return arr

    def load_build(self):
        super().load_build()
        if isinstance(self.stack[-1], _ArrayRecord):
            rec = self.stack.pop()
            self.stack.append(self._read_array(rec))

    dispatch = dict(pickle._Unpickler.dispatch)
    dispatch[pickle.BUILD[0]] = load_build


def _open_stream(path) -> io.BufferedIOBase:
    """The (decompressed) byte stream of a joblib file."""
    with open(path, "rb") as f:
        head = f.read(8)
    if head.startswith(b"\x1f\x8b"):
        return gzip.open(path, "rb")
    if head.startswith(b"BZh"):
        return bz2.open(path, "rb")
    if head.startswith(b"\xfd7zXZ") or head.startswith(b"]\x00\x00"):
        return lzma.open(path, "rb")
    if head[:1] == b"x" and len(head) > 1 and (head[0] * 256 + head[1]) % 31 == 0:
        with open(path, "rb") as f:
            return io.BytesIO(zlib.decompress(f.read()))
    return open(path, "rb")


def load_ell(path=None):
    """Loads an ``ell_matrix`` saved with ``ell_matrix.save`` (here or by the
    reference, jclsquant/hams.py:14-19).  ``path=None`` prints a message and
    returns None, as the reference does."""
    if path is None:
        print('No path inserted')
        return None
    path = os.fspath(path)
    with _open_stream(path) as fh:
        obj = _Unpickler(fh).load()
    if not isinstance(obj, ell_matrix):
        raise pickle.UnpicklingError(f"{path} does not hold an ell_matrix")
    return obj
