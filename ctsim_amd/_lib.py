"""ctypes binding of libctsim_amd.so (the C ABI of include/ctsim_hip.h)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libctsim_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -m ctsim_amd.build` (there is no CPU fallback)")

lib = C.CDLL(LIB_PATH)

_i, _d, _vp, _cp = C.c_int, C.c_double, C.c_void_p, C.c_char_p
_pp = C.POINTER(C.c_void_p)

# name: (restype, argtypes)
SIGNATURES = {
    "ctsim_ctx_create": (_i, [_pp, _i, _vp]),
    "ctsim_ctx_destroy": (None, [_vp]),
    "ctsim_ctx_set_stream": (_i, [_vp, _vp]),
    "ctsim_ctx_stream": (_vp, [_vp]),
    "ctsim_ctx_synchronize": (_i, [_vp]),
    "ctsim_last_error": (_cp, []),
    "ctsim_version": (_cp, []),
    "ctsim_backprojector_create": (_i, [_vp, _pp, _i, _i, _d, _i, _d, _d, _d, _cp, _cp, _i, _i,
                                        _i, _vp]),
    "ctsim_backprojector_destroy": (None, [_vp]),
    "ctsim_backproject_views": (_i, [_vp, _vp, _i, _vp, _vp, _i, _vp]),
    "ctsim_backproject_postprocess": (_i, [_vp, _vp, _vp]),
    "ctsim_fbp_slice": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _i, _vp]),
    "ctsim_fbp_slice_static": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "ctsim_backprojector_info": (_i, [_vp, _vp]),
    "ctsim_backprojector_mirror_members": (_i, [_vp, _vp, _i]),
    "ctsim_procsignal_create": (_i, [_vp, _pp, _cp, _cp, _d, _d, _i, _d, _cp, _cp, _i, _i, _i,
                                     _d, _d]),
    "ctsim_procsignal_destroy": (None, [_vp]),
    "ctsim_procsignal_info": (_i, [_vp, _vp, _vp]),
    "ctsim_procsignal_get_filter": (_i, [_vp, _vp, _i]),
    "ctsim_filter_views": (_i, [_vp, _vp, _i, _vp, _i, _vp]),
    "ctsim_phantom_create": (_i, [_pp, _cp]),
    "ctsim_phantom_destroy": (None, [_vp]),
    "ctsim_phantom_add_element": (_i, [_vp, _cp, _d, _d, _d, _d, _d, _d]),
    "ctsim_phantom_extent": (_i, [_vp, _vp]),
    "ctsim_phantom_element": (_i, [_vp, _i, _vp]),
    "ctsim_scanner_create": (_i, [_pp, _vp, _cp, _i, _i, _i, _i, _d, _d, _d, _d, _d]),
    "ctsim_scanner_destroy": (None, [_vp]),
    "ctsim_scanner_info": (_i, [_vp, _vp]),
    "ctsim_collect_projections": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _vp, _vp]),
    "ctsim_image_stats": (_i, [_vp, _vp, _vp, _i, _i, _vp]),
    "ctsim_phantom_rasterize": (_i, [_vp, _vp, _i, _i, _d, _i, _vp, _vp]),
    "ctsim_rebin_create": (_i, [_vp, _vp, _i, _i, _i, _d, _d, _d, _d, _d, _d, _vp]),
    "ctsim_rebin_destroy": (None, [_vp]),
    "ctsim_rebin_info": (_i, [_vp, _vp, _vp]),
    "ctsim_rebin_run": (_i, [_vp, _vp, _vp, _vp]),
    "ctsim_batch_create": (_i, [_vp, _vp, _vp, _i]),
    "ctsim_batch_destroy": (None, [_vp]),
    "ctsim_batch_run": (_i, [_vp, _vp, _vp, _vp]),
    "ctsim_batch_status": (_i, [_vp, _vp]),
    "ctsim_batch_set_shape": (_i, [_vp, _vp, _i, _i, _i, _i, _i]),
    "ctsim_batch_autotune": (_i, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "ctsim_batch_info": (_i, [_vp, _vp]),
    "ctsim_helical180li": (_i, [_vp, _i, _i, _i, _d, _d, _d, _d, _i, _vp, _vp, _i, _vp, _vp]),
    "ctsim_halfscan_feather": (_i, [_vp, _i, _i, _i, _d, _d, _d, _d, _vp, _vp]),
    "ctsim_ctx_profile": (_i, [_vp, _i]),
    "ctsim_ctx_profile_read": (_i, [_vp, _vp, _vp]),
    "ctsim_backprojector_autotune": (_i, [_vp, _vp, _vp, _vp, _i, _vp, _vp]),
    "ctsim_backprojector_set_shape": (_i, [_vp, _vp]),
    "ctsim_crmath_eval": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args
