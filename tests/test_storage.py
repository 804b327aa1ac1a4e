"""On-disk ELL format: ell_matrix.save / load_ell (CPU only).

The reference saves ``joblib.dump(self, path)`` (jclsquant/ell_matrix.py:202-207)
and loads with ``joblib.load`` (jclsquant/hams.py:14-19).  A file in the
reference's layout is made here by dumping an object whose class path is
``jclsquant.ell_matrix.ell_matrix`` and whose attributes are the ones the
reference's constructor and csr_to_ell set (jclsquant/ell_matrix.py:134-200),
through a throw-away module registered only for the dump.  ``load_ell`` must
read it back as this package's ell_matrix, and must refuse any stream that
names a constructor outside its list, without running it.
"""
import os
import pickle
import sys
import types

import joblib
import numpy as np
import pytest

import minijclsquant_amd as jcl


def reference_state(A):
    """The attribute set of the reference's ell_matrix after construction
    from CSR, with the reference's scalar types (len_row and len_col are
    numpy int32, jclsquant/ell_matrix.py:188-192)."""
    return {
        "A": A.A, "type": "ELLPACK(ELL) sparse matrix format",
        "len_row": np.int32(A.len_row), "len_col": np.int32(A.shape[0]),
        "ell_data": None, "ell_indices": None,
        "shape": np.array(A.shape, dtype=np.int32), "n_threads": 8,
        "DX": A.DX, "DY": A.DY, "CJ": A.CJ,
        "dx_vec": A.dx_vec, "dy_vec": A.dy_vec, "dz_vec": None, "Omega": None,
        "conjugates": A.conjugates, "valid_row_values": A.valid_row_values,
        "data": A.data, "indices": A.indices, "bounds": A.bounds,
        "space": "r", "velocities": "False",
    }


def dump_reference_layout(A, path, compress=0):
    pkg = types.ModuleType("jclsquant")
    mod = types.ModuleType("jclsquant.ell_matrix")

    class ell_matrix:  # noqa: N801 - the reference's class name
        pass

    ell_matrix.__module__ = "jclsquant.ell_matrix"
    ell_matrix.__qualname__ = "ell_matrix"
    mod.ell_matrix = ell_matrix
    pkg.ell_matrix = mod
    saved = {k: sys.modules.get(k) for k in ("jclsquant", "jclsquant.ell_matrix")}
    sys.modules["jclsquant"] = pkg
    sys.modules["jclsquant.ell_matrix"] = mod
    try:
        obj = ell_matrix()
        obj.__dict__.update(reference_state(A))
        joblib.dump(obj, path, compress=compress)
    finally:
        for k, v in saved.items():
            if v is None:
                sys.modules.pop(k, None)
            else:
                sys.modules[k] = v


def assert_same(A, B):
    assert type(B) is jcl.ell_matrix
    for name in ("data", "indices", "valid_row_values", "bounds", "shape"):
        assert np.array_equal(getattr(A, name), getattr(B, name)), name
    assert int(B.len_row) == int(A.len_row)
    assert int(B.len_col) == int(A.shape[0])
    assert (A.A != B.A).nnz == 0
    assert B._dev == {}


@pytest.mark.parametrize("compress", [0, 3, ("gzip", 3), ("bz2", 3), ("xz", 3), ("lzma", 3)])
def test_reference_layout_loads(tmp_path, compress):
    H, _ = jcl.square_lattice(7, t=-1.0, W=1.0, seed=3)
    A = jcl.ell_matrix(H)
    p = tmp_path / "H.json"
    dump_reference_layout(A, p, compress)
    assert "jclsquant" not in sys.modules
    B = jcl.load_ell(str(p))
    assert_same(A, B)
    # the loaded object behaves like a constructed one
    Ht = B.modifier(jcl.modifier_bounds)
    assert np.allclose(Ht.data, A.data * (1 / A.bounds[1]), rtol=0, atol=1e-15)


def test_reference_layout_with_bond_vectors(tmp_path):
    H, Dx, Dy = jcl.honeycomb_lattice(4, t=-2.7, W=0.5, seed=2)
    A = jcl.ell_matrix(H, Dx, Dy)
    p = tmp_path / "hc.joblib"
    dump_reference_layout(A, p)
    B = jcl.load_ell(p)
    assert_same(A, B)
    assert np.array_equal(A.dx_vec, B.dx_vec) and np.array_equal(A.dy_vec, B.dy_vec)
    assert np.array_equal(A.conjugates, B.conjugates)
    assert (A.DX != B.DX).nnz == 0


def test_own_save_round_trip(tmp_path):
    H, _ = jcl.square_lattice(5, t=-1.0, W=2.0, seed=9)
    A = jcl.ell_matrix(H)
    A.save(str(tmp_path / "a.joblib"))
    B = jcl.load_ell(str(tmp_path / "a.joblib"))
    assert_same(A, B)
    # a modified object keeps its modified values
    C = A.modifier(jcl.modifier_bounds)
    C.save(str(tmp_path / "c.joblib"))
    assert np.array_equal(jcl.load_ell(str(tmp_path / "c.joblib")).data, C.data)


def test_load_none_path(capsys):
    assert jcl.load_ell() is None
    assert "No path" in capsys.readouterr().out


class _Calls:
    """Pickles as a call of os.getcwd (harmless if it ran; it must not)."""

    def __reduce__(self):
        return (os.getcwd, ())


def test_refuses_foreign_constructors(tmp_path):
    p = tmp_path / "bad.pkl"
    with open(p, "wb") as f:
        pickle.dump(_Calls(), f)
    with pytest.raises(pickle.UnpicklingError, match="refusing"):
        jcl.load_ell(str(p))
    # nested inside an otherwise valid ell_matrix
    H, _ = jcl.square_lattice(3, seed=0)
    A = jcl.ell_matrix(H)
    A.space = _Calls()
    joblib.dump(A, str(p))
    with pytest.raises(pickle.UnpicklingError, match="refusing"):
        jcl.load_ell(str(p))


def test_refuses_object_arrays_and_other_objects(tmp_path):
    H, _ = jcl.square_lattice(3, seed=0)
    A = jcl.ell_matrix(H)
    A.bounds = np.array([None, 1.0], dtype=object)
    p = tmp_path / "obj.joblib"
    joblib.dump(A, str(p))
    with pytest.raises(pickle.UnpicklingError, match="object arrays"):
        jcl.load_ell(str(p))
    joblib.dump({"data": np.zeros(3)}, str(p))
    with pytest.raises(pickle.UnpicklingError, match="does not hold an ell_matrix"):
        jcl.load_ell(str(p))
