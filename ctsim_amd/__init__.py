"""ctsim_amd -- MI355X (gfx950) implementation of CTSim's projection,
filtering and backprojection hot path.

The compute lives in ``libctsim_amd.so`` (hand-written HIP kernels behind the
C ABI declared in include/ctsim_hip.h).  This package only loads that library
and offers thin Python handles over it for tests and benchmarks; there is no
Python or CPU fallback for any kernel.  Touching any handle below loads the
library and fails loudly when it has not been built
(``python -m ctsim_amd.build``); the ``build`` and ``workloads`` submodules
work without it.
"""
from __future__ import annotations

import importlib

__version__ = "0.1"

_API = {"Context", "Backprojector", "ProcessSignal", "Phantom", "Scanner", "CtsimError",
        "PREC_F64", "PREC_MIXED", "image_stats", "helical180li",
        "halfscan_feather", "Batch", "SliceDesc", "Rebin", "check", "crmath_eval"}


def __getattr__(name):
    if name in ("lib", "LIB_PATH"):
        return getattr(importlib.import_module("._lib", __name__), name)
    if name in _API:
        return getattr(importlib.import_module(".api", __name__), name)
    raise AttributeError(name)
