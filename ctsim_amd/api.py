"""Python handles over the C ABI.  Device arrays are torch tensors on a HIP
device (torch is used for device memory and streams only); every computation
runs in libctsim_amd.so."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib

PREC_F64 = 0
PREC_MIXED = 1

GEOMETRY_IDS = {"parallel": 0, "equiangular": 1, "equilinear": 2}


class CtsimError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib.ctsim_last_error().decode(errors="replace")
        raise CtsimError(f"{what}: status {rc}: {msg}")


def _dptr(t) -> int:
    """Device pointer of a torch tensor (must be contiguous)."""
    if not t.is_contiguous():
        raise ValueError("device arrays must be contiguous")
    return t.data_ptr()


def _hptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """A ctsim_ctx bound to one device and (by default) torch's current stream."""

    def __init__(self, device: int = 0, stream=None):
        import torch
        self.device = device
        if stream is None:
            stream = torch.cuda.current_stream(device)
        self._torch_stream = stream
        h = C.c_void_p()
        check(lib.ctsim_ctx_create(C.byref(h), device, C.c_void_p(stream.cuda_stream)),
              "ctsim_ctx_create")
        self.handle = h

    @property
    def stream(self):
        return self._torch_stream

    def synchronize(self):
        check(lib.ctsim_ctx_synchronize(self.handle), "ctsim_ctx_synchronize")

    def close(self):
        if getattr(self, "handle", None):
            lib.ctsim_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class Backprojector:
    """Backprojector plan (libctsim/backprojectors.cpp:100-319)."""

    def __init__(self, ctx: Context, geometry, nDet: int, detInc: float, nView: int,
                 viewDiameter: float, focalLength: float, sourceDetectorLength: float,
                 backproj: str = "idiff", interp: str = "linear", nx: int = 256, ny: int = 256,
                 roi=None, interpFactor: int = 1):
        self.ctx = ctx
        g = GEOMETRY_IDS[geometry] if isinstance(geometry, str) else int(geometry)
        h = C.c_void_p()
        roi_a = None if roi is None else np.ascontiguousarray(roi, dtype=np.float64)
        check(lib.ctsim_backprojector_create(
            ctx.handle, C.byref(h), g, nDet, detInc, nView, viewDiameter, focalLength,
            sourceDetectorLength, backproj.encode(), interp.encode(), interpFactor, nx, ny,
            None if roi_a is None else _hptr(roi_a)), "ctsim_backprojector_create")
        self.handle = h
        self.nx, self.ny, self.nDet = nx, ny, nDet
        self._roi = roi_a

    def info(self):
        o = np.zeros(8, np.float64)
        check(lib.ctsim_backprojector_info(self.handle, _hptr(o)), "ctsim_backprojector_info")
        return o

    def backproject(self, filtered, angles: np.ndarray, image, precision: int = PREC_F64):
        """Add all rows of ``filtered`` (device [nView, nDet]; float64 for F64,
        float32 for MIXED) into ``image`` (device float32 [nx, ny])."""
        import torch
        want = torch.float64 if precision == PREC_F64 else torch.float32
        if filtered.dtype != want:
            raise TypeError(f"filtered must be {want} for this precision")
        if image.dtype != torch.float32 or tuple(image.shape) != (self.nx, self.ny):
            raise TypeError("image must be float32 [nx, ny]")
        ang = np.ascontiguousarray(angles, dtype=np.float64)
        nView = filtered.shape[0]
        if ang.shape[0] < nView:
            raise ValueError("not enough view angles")
        check(lib.ctsim_backproject_views(self.ctx.handle, self.handle, precision,
                                          C.c_void_p(_dptr(filtered)), _hptr(ang), nView,
                                          C.c_void_p(_dptr(image))), "ctsim_backproject_views")

    def postprocess(self, image):
        check(lib.ctsim_backproject_postprocess(self.ctx.handle, self.handle,
                                                C.c_void_p(_dptr(image))),
              "ctsim_backproject_postprocess")

    def close(self):
        if getattr(self, "handle", None):
            lib.ctsim_backprojector_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()
