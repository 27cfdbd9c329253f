"""MI355X-native evaluation of the gtinv machine-learning potential.

The compute path is the HIP library ``_build/libmlip_hip.so`` behind the C
ABI in ``include/mlip_hip.h``; this package only loads it and mirrors the
reference pair-style interface (``pair_style mlip_gtinv`` /
``pair_coeff * * file elements``) for Python callers.
"""
from ._lib import LIB_PATH, MlipError, library, build_library  # noqa: F401
from .pair import (BatchSystem, PairMLIPGtinv, RequestPool, ResidentSystem, inspect_model,  # noqa: F401
                   periodic_full_list)
