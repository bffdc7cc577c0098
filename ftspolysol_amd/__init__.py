"""MI355X-native complex-Langevin field-theoretic simulation engine.

The engine is C++ host code plus hand-written HIP kernels for gfx950, built
in-tree as libfts_hip.so and driven through the C ABI in include/fts.h.  This
package holds the kernels' sources (csrc/), the ctypes mirror of the ABI
(api.py) and the workload descriptions (system.py).
"""
from .api import (Engine, FtsError, hip_engine, load_hip_library, PREC_FP32, PREC_FP64,  # noqa: F401
                  SCHEME_EM, SCHEME_PC_REFERENCE, SCHEME_SEMI_IMPLICIT, SCHEME_SI_LEARNED, SCHEME_SIPC, SCHEME_SIPC_LEARNED,
                  STRESS_ALL, STRESS_PISO)
from .system import System, config, initial_fields, make_system  # noqa: F401
