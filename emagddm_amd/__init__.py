"""emagddm_amd -- MI355X-native FETI-DP solver for lowest-order Nedelec curl-curl.

The product is ``libemagddm_fetidp.so`` (C ABI, include/emagddm_fetidp.h):
host C++ topology/plans + hand-written gfx950 HIP kernels.  This package holds
the ctypes binding (``_lib``), the solver interface mirror (``fetidp``) and the
seeded synthetic workloads (``workloads``).
"""
from .fetidp import FETIDPSolver, SolverResult, SolverStatus, integrate_elements  # noqa: F401
from . import workloads  # noqa: F401
