"""The C-ABI library loads without a GPU and exports exactly what
include/jclkpm.h declares; the compute path refuses to run without a device."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from minijclsquant_amd import _native as nat

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "jclkpm.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def test_header_parses():
    names = declared_functions()
    assert "jkpm_moments" in names and "rec_A_tab" in names and "rec_A_tab_2" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = nat.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), f"{name} declared in jclkpm.h but not exported"
        assert name in nat.SIGNATURES, f"{name} has no ctypes signature"
    assert set(nat.SIGNATURES) == set(declared_functions())


def test_header_constants_match_the_wrapper():
    """Every #define of the header that the ctypes wrapper restates has the
    same value there (status codes, value types, flags, variant count)."""
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    defines = dict(re.findall(r"^#define\s+(JKPM_[A-Z0-9_]+)\s+(.+?)\s*$", text, flags=re.M))
    checked = 0
    for name, expr in defines.items():
        if not hasattr(nat, name):
            continue
        value = eval(expr.replace("u <<", " <<").replace("u)", ")"), {})
        assert getattr(nat, name) == value, name
        checked += 1
    assert checked >= 15
    assert nat.JKPM_STEP_VARIANTS == 48


def test_symbols_are_c_linkage():
    # extern "C": the plain name resolves (no C++ mangling) via dlsym
    lib = ctypes.CDLL(str(nat.LIB_PATH))
    for name in declared_functions():
        getattr(lib, name)


def test_calls_that_need_no_device():
    lib = nat.load_library()
    assert lib.jkpm_version() >= 100
    # a call that succeeds clears the error an earlier call left
    assert lib.jkpm_profile_enable(0) == nat.JKPM_OK
    assert lib.jkpm_last_error() == b""
    # argument errors are reported before any device work
    rc = lib.jkpm_moments(None, None, None, None, 1, 1, None, None, 0, None)
    assert rc == nat.JKPM_ERR_ARG
    assert b"NULL" in lib.jkpm_last_error()
    out = ctypes.c_void_p()
    rc = lib.jkpm_ell_create(None, None, 4, 2, 1.0, 0, 0, None, ctypes.byref(out))
    assert rc == nat.JKPM_ERR_ARG
    rc = lib.jkpm_profile_enable(-1)
    assert rc == nat.JKPM_ERR_ARG


def test_compute_path_fails_loudly_without_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is visible; this checks the no-device failure")
    import minijclsquant_amd as jcl

    H, _ = jcl.square_lattice(4)
    A = jcl.ell_matrix(H)
    with pytest.raises(RuntimeError, match="no HIP device"):
        jcl.rec_A_tab(8, A, np.ones(16, complex))
    with pytest.raises(RuntimeError, match="no HIP device"):
        jcl.kpm_dos(A, 8, np.ones(16, complex))


def test_missing_library_is_an_error(tmp_path):
    with pytest.raises(RuntimeError, match="not built"):
        nat.load_library(tmp_path / "libjclkpm_hip.so")
