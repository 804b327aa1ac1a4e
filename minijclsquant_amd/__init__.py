"""MI355X-native Chebyshev/KPM hot path of jclsquant.

The package namespace mirrors the reference's star-import surface
(jclsquant/__init__.py:1-10) for the KPM path: ell_matrix, recurrence
functions, kernels and moments, modifiers, DOS observables (CPU- and
GPU-named), lattice and Hamiltonian builders (H_graphene), and the Cython
helpers those modules re-export (blas_funs: scal, axpy, axpy2, aAxby, vdot,
dot, vec_mul*; Extra_funs: data_indices_c(_0), diagonal_modifier,
light_modifier_cython, table_G_delta).  The compute path is the HIP library
lib/libjclkpm_hip.so (C-ABI in include/jclkpm.h), called through ctypes.
"""
from .blas_funs import *  # noqa: F401,F403
from .ell_matrix import *  # noqa: F401,F403
from .extra_funs import *  # noqa: F401,F403
from .hams import *  # noqa: F401,F403
from .kernel_and_moments import *  # noqa: F401,F403
from .kubo import *  # noqa: F401,F403
from .lattice import *  # noqa: F401,F403
from .modifiers import *  # noqa: F401,F403
from .observables import *  # noqa: F401,F403
from .observables_gpu import *  # noqa: F401,F403
from .recurrence_functions import *  # noqa: F401,F403

__version__ = "0.1.0"
from .batch import *  # noqa: F401,F403,E402
