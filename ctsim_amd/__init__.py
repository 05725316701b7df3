"""ctsim_amd -- MI355X (gfx950) implementation of CTSim's projection,
filtering and backprojection hot path.

The compute lives in ``libctsim_amd.so`` (hand-written HIP kernels behind the
C ABI declared in include/ctsim_hip.h).  This package only loads that library
and offers thin Python handles over it for tests and benchmarks; there is no
Python or CPU fallback for any kernel.  Importing fails loudly when the
library has not been built (``python -m ctsim_amd.build``).
"""
from __future__ import annotations

import os

from ._lib import LIB_PATH, lib  # noqa: F401  (raises if the library is missing)
from .api import (  # noqa: F401
    PREC_F64,
    PREC_MIXED,
    Backprojector,
    Context,
    CtsimError,
)

__all__ = ["lib", "LIB_PATH", "Context", "Backprojector", "CtsimError", "PREC_F64", "PREC_MIXED"]
__version__ = "0.1"
HERE = os.path.dirname(os.path.abspath(__file__))
