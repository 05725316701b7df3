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

    def profile(self, enable: bool = True):
        """Bracket every main backprojection kernel with HIP events."""
        check(lib.ctsim_ctx_profile(self.handle, int(enable)), "ctsim_ctx_profile")

    def profile_read(self) -> tuple[float, int]:
        """(summed kernel ms, launches) since the last read."""
        ms = np.zeros(1, np.float64)
        n = np.zeros(1, np.int32)
        check(lib.ctsim_ctx_profile_read(self.handle, _hptr(ms), _hptr(n)), "ctsim_ctx_profile_read")
        return float(ms[0]), int(n[0])

    def close(self):
        if getattr(self, "handle", None):
            lib.ctsim_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def image_stats(ctx: Context, ref, cmp) -> tuple[float, float, float]:
    """Herman's d, r, e of two device images [nx][ny] float32 (ref first),
    ImageFile::comparativeStatistics (libctsim/imagefile.cpp:185-240)."""
    if ref.shape != cmp.shape or ref.dim() != 2:
        raise ValueError("image_stats: images must be 2-D and of one shape")
    out = np.zeros(3, np.float64)
    nx, ny = ref.shape
    check(lib.ctsim_image_stats(ctx.handle, C.c_void_p(_dptr(ref)), C.c_void_p(_dptr(cmp)), nx,
                                ny, _hptr(out)), "ctsim_image_stats")
    return float(out[0]), float(out[1]), float(out[2])


CRMATH_FN = {"cos": 0, "sin": 1, "atan": 2, "atan2": 3}


def crmath_eval(ctx: Context, fn: str, x, y=None):
    """The FP64 kernels' correctly rounded cos/sin/atan/atan2 evaluated on the
    device over float64 tensors (atan2 takes y, x)."""
    import torch
    out = torch.empty_like(x)
    check(lib.ctsim_crmath_eval(ctx.handle, CRMATH_FN[fn], x.numel(), C.c_void_p(_dptr(x)),
                                C.c_void_p(_dptr(y)) if y is not None else None,
                                C.c_void_p(_dptr(out))), "ctsim_crmath_eval")
    return out


def helical180li(ctx: Context, geometry, sino, rotInc: float, detInc: float, fanBeamAngle: float,
                 sourceDetectorLength: float, interpView: int = -1):
    """Projections::Helical180LI (libctsim/projections.cpp:200-358) of a device
    sinogram float32 [nView, nDet]; returns (device [L+1, nDet], angles)."""
    import torch
    g = GEOMETRY_IDS[geometry] if isinstance(geometry, str) else int(geometry)
    if sino.dtype != torch.float32 or sino.dim() != 2:
        raise TypeError("sino must be float32 [nView, nDet]")
    nView, nDet = sino.shape
    cap = int((np.pi + fanBeamAngle) / rotInc) + 2
    out = torch.zeros((cap, nDet), dtype=torch.float32, device=sino.device)
    ang = np.zeros(cap, np.float64)
    n = C.c_int()
    check(lib.ctsim_helical180li(ctx.handle, g, nDet, nView, rotInc, detInc, fanBeamAngle,
                                 sourceDetectorLength, interpView, C.c_void_p(_dptr(sino)),
                                 C.c_void_p(_dptr(out)), cap, C.byref(n), _hptr(ang)),
          "ctsim_helical180li")
    return out[: n.value], ang[: n.value].copy()


def halfscan_feather(ctx: Context, geometry, views, rotInc: float, detInc: float,
                     fanBeamAngle: float, sourceDetectorLength: float):
    """Projections::HalfScanFeather (projections.cpp:369-439); returns a new
    device array."""
    import torch
    g = GEOMETRY_IDS[geometry] if isinstance(geometry, str) else int(geometry)
    src = views.contiguous()
    out = torch.empty_like(src)
    nView, nDet = src.shape
    check(lib.ctsim_halfscan_feather(ctx.handle, g, nDet, nView, rotInc, detInc, fanBeamAngle,
                                     sourceDetectorLength, C.c_void_p(_dptr(src)),
                                     C.c_void_p(_dptr(out))), "ctsim_halfscan_feather")
    return out


class Rebin:
    """Fan-beam to parallel rebinning plan (Projections::interpolateToParallel,
    libctsim/projections.cpp:1250-1328)."""

    def __init__(self, ctx: Context, geometry, nDet: int, nView: int, detStart: float,
                 detInc: float, focalLength: float, sourceDetectorLength: float,
                 viewDiameter: float, rotStart: float, angles: np.ndarray):
        self.ctx = ctx
        g = GEOMETRY_IDS[geometry] if isinstance(geometry, str) else int(geometry)
        ang = np.ascontiguousarray(angles, dtype=np.float64)
        if ang.shape[0] < nView:
            raise ValueError("not enough view angles")
        h = C.c_void_p()
        check(lib.ctsim_rebin_create(ctx.handle, C.byref(h), g, nDet, nView, detStart, detInc,
                                     focalLength, sourceDetectorLength, viewDiameter, rotStart,
                                     _hptr(ang)), "ctsim_rebin_create")
        self.handle = h
        self.nDet, self.nView = nDet, nView
        o = np.zeros(4, np.float64)
        self.angles = np.zeros(nView, np.float64)
        check(lib.ctsim_rebin_info(h, _hptr(o), _hptr(self.angles)), "ctsim_rebin_info")
        self.rotStart, self.rotInc, self.detStart, self.detInc = (float(x) for x in o)

    def run(self, sino):
        """Parallel projections (new device array) of a float32 [nView, nDet] sinogram."""
        import torch
        if sino.dtype != torch.float32 or tuple(sino.shape) != (self.nView, self.nDet):
            raise TypeError("sino must be float32 [nView, nDet]")
        src = sino.contiguous()
        out = torch.empty_like(src)
        check(lib.ctsim_rebin_run(self.ctx.handle, self.handle, C.c_void_p(_dptr(src)),
                                  C.c_void_p(_dptr(out))), "ctsim_rebin_run")
        return out

    def close(self):
        if getattr(self, "handle", None):
            lib.ctsim_rebin_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class SliceDesc(C.Structure):
    """ctsim_slice_desc of include/ctsim_hip.h."""
    _fields_ = [("geometry", C.c_int), ("nDet", C.c_int), ("nView", C.c_int),
                ("detInc", C.c_double), ("viewDiameter", C.c_double),
                ("focalLength", C.c_double), ("sourceDetectorLength", C.c_double),
                ("filter", C.c_char_p), ("method", C.c_char_p), ("filterParam", C.c_double),
                ("generation", C.c_char_p), ("zeropad", C.c_int), ("backproj", C.c_char_p),
                ("interp", C.c_char_p), ("nx", C.c_int), ("ny", C.c_int),
                ("angles", C.c_void_p)]


class Batch:
    """Batched mixed-precision FBP of many slices (one persistent launch in a
    HIP graph).  ``slices``: dicts with the SliceDesc fields (``angles`` a
    numpy array)."""

    def __init__(self, ctx: Context, slices):
        self.ctx = ctx
        self._keep = []
        arr = (SliceDesc * len(slices))()
        self.shapes = []
        for k, d in enumerate(slices):
            ang = np.ascontiguousarray(d["angles"], dtype=np.float64)
            self._keep.append(ang)
            g = GEOMETRY_IDS[d["geometry"]] if isinstance(d["geometry"], str) else int(d["geometry"])
            arr[k] = SliceDesc(g, d["nDet"], d["nView"], d["detInc"], d["viewDiameter"],
                               d["focalLength"], d["sourceDetectorLength"], d["filter"].encode(),
                               d["method"].encode(), d.get("filterParam", -1.0),
                               d.get("generation", "direct").encode(), d.get("zeropad", 1),
                               d.get("backproj", "idiff").encode(), d["interp"].encode(),
                               d["nx"], d["ny"], ang.ctypes.data)
            self.shapes.append((d["nView"], d["nDet"], d["nx"], d["ny"]))
        h = C.c_void_p()
        check(lib.ctsim_batch_create(ctx.handle, C.byref(h), arr, len(slices)), "ctsim_batch_create")
        self.handle = h

    def _ptrs(self, sinos, images):
        import torch
        if len(sinos) != len(self.shapes) or len(images) != len(self.shapes):
            raise ValueError("one sinogram and one image per slice")
        for (nv, nd, nx, ny), s, im in zip(self.shapes, sinos, images):
            if s.dtype != torch.float32 or tuple(s.shape) != (nv, nd):
                raise TypeError("sinogram must be float32 [nView, nDet]")
            if im.dtype != torch.float32 or tuple(im.shape) != (nx, ny):
                raise TypeError("image must be float32 [nx, ny]")
        sp = (C.c_void_p * len(sinos))(*[_dptr(s) for s in sinos])
        ip = (C.c_void_p * len(images))(*[_dptr(i) for i in images])
        return sp, ip

    def run(self, sinos, images):
        sp, ip = self._ptrs(sinos, images)
        check(lib.ctsim_batch_run(self.ctx.handle, self.handle, sp, ip), "ctsim_batch_run")

    def status(self):
        check(lib.ctsim_batch_status(self.ctx.handle, self.handle), "ctsim_batch_status")

    def set_shape(self, P: int, lanesY: int, chunkBytes: int, fused: bool, sym: int = 1):
        """sym: 0 plain body; 1, 2, 4 symmetric body with that many view groups;
        + 8 the mirror body for the other slices (two-phase mode)."""
        check(lib.ctsim_batch_set_shape(self.ctx.handle, self.handle, P, lanesY, chunkBytes,
                                        int(fused), int(sym)), "ctsim_batch_set_shape")

    def autotune(self, sinos, images):
        """Times the launch shapes; returns (P, lanesY, chunkBytes, fused, sym, ms)."""
        sp, ip = self._ptrs(sinos, images)
        shape = np.zeros(5, np.int32)
        ms = np.zeros(1, np.float64)
        check(lib.ctsim_batch_autotune(self.ctx.handle, self.handle, sp, ip, _hptr(shape),
                                       _hptr(ms)), "ctsim_batch_autotune")
        return (*map(int, shape), float(ms[0]))

    def info(self):
        o = np.zeros(4, np.float64)
        check(lib.ctsim_batch_info(self.handle, _hptr(o)), "ctsim_batch_info")
        return {"slices": int(o[0]), "pvu": o[1], "fused": bool(o[2]), "grid": int(o[3])}

    def close(self):
        if getattr(self, "handle", None):
            lib.ctsim_batch_destroy(self.handle)
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
        self.nx, self.ny, self.nDet, self.interpFactor = nx, ny, nDet, interpFactor
        self._roi = roi_a

    def info(self):
        o = np.zeros(8, np.float64)
        check(lib.ctsim_backprojector_info(self.handle, _hptr(o)), "ctsim_backprojector_info")
        return o

    def mirror_members(self, angles: np.ndarray) -> int:
        """Mirror members (4, 2 or 1) of the mirror kernel for these angles."""
        a = np.ascontiguousarray(angles, np.float64)
        n = lib.ctsim_backprojector_mirror_members(self.handle, _hptr(a), len(a))
        if n < 1:
            check(n, "ctsim_backprojector_mirror_members")
        return n

    def backproject(self, filtered, angles: np.ndarray, image, precision: int = PREC_F64):
        """Add all rows of ``filtered`` (device [nView, nDet]; float64 for F64,
        float32 for MIXED) into ``image`` (device float32 [nx, ny])."""
        import torch
        want = torch.float64 if precision == PREC_F64 else torch.float32
        if filtered.dtype != want:
            raise TypeError(f"filtered must be {want} for this precision")
        if image.dtype != torch.float32 or tuple(image.shape) != (self.nx, self.ny):
            raise TypeError("image must be float32 [nx, ny]")
        if filtered.dim() != 2 or filtered.shape[1] != self.nDet * self.interpFactor:
            raise ValueError("filtered must be [nView, nDet * interpFactor]")
        ang = np.ascontiguousarray(angles, dtype=np.float64)
        nView = filtered.shape[0]
        if ang.shape[0] < nView:
            raise ValueError("not enough view angles")
        check(lib.ctsim_backproject_views(self.ctx.handle, self.handle, precision,
                                          C.c_void_p(_dptr(filtered)), _hptr(ang), nView,
                                          C.c_void_p(_dptr(image))), "ctsim_backproject_views")

    def autotune(self, filtered, angles: np.ndarray):
        """Times the MIXED kernel's launch shapes on these float32 rows and
        keeps the fastest; returns (lanesY, P, split, chunk, window, sym, stagingKB, ms)."""
        import torch
        if filtered.dtype != torch.float32 or filtered.dim() != 2 or \
                filtered.shape[1] != self.nDet * self.interpFactor:
            raise TypeError("filtered must be float32 [nView, nDet * interpFactor]")
        ang = np.ascontiguousarray(angles, dtype=np.float64)
        if ang.shape[0] < filtered.shape[0]:
            raise ValueError("not enough view angles")
        shape = np.zeros(7, np.int32)
        ms = np.zeros(1, np.float64)
        check(lib.ctsim_backprojector_autotune(self.ctx.handle, self.handle,
                                               C.c_void_p(_dptr(filtered)), _hptr(ang),
                                               filtered.shape[0], _hptr(shape), _hptr(ms)),
              "ctsim_backprojector_autotune")
        return (*map(int, shape), float(ms[0]))

    def set_shape(self, lanesY: int, P: int, split: int, chunk: int, sym: int = 0,
                  stagingKB: int = 0):
        """Pins the MIXED launch shape; sym = 1 runs the rotationally symmetric
        kernel when the image and the views allow it (P <= 2), staging its
        windows in stagingKB KB of LDS (0: 62); 2 its pipelined form; 3 the
        mirror kernel when the views have mirror partners."""
        s = np.array([lanesY, P, split, chunk, sym, stagingKB], np.int32)
        check(lib.ctsim_backprojector_set_shape(self.handle, _hptr(s)), "ctsim_backprojector_set_shape")

    def postprocess(self, image):
        check(lib.ctsim_backproject_postprocess(self.ctx.handle, self.handle,
                                                C.c_void_p(_dptr(image))),
              "ctsim_backproject_postprocess")

    def fbp(self, ps: "ProcessSignal", sino, angles: np.ndarray, image,
            precision: int = PREC_MIXED):
        """One whole slice (ctsim_fbp_slice): ``image`` (device float32
        [nx, ny]) is overwritten with the post-scaled reconstruction of every
        row of ``sino`` (device float32 [nView, nDet]) filtered by ``ps``."""
        import torch
        if sino.dtype != torch.float32 or sino.dim() != 2 or sino.shape[1] != self.nDet:
            raise TypeError("sino must be float32 [nView, nDet]")
        if image.dtype != torch.float32 or tuple(image.shape) != (self.nx, self.ny):
            raise TypeError("image must be float32 [nx, ny]")
        ang = np.ascontiguousarray(angles, dtype=np.float64)
        nView = sino.shape[0]
        if ang.shape[0] < nView:
            raise ValueError("not enough view angles")
        check(lib.ctsim_fbp_slice(self.ctx.handle, ps.handle, self.handle, precision,
                                  C.c_void_p(_dptr(sino)), _hptr(ang), nView,
                                  C.c_void_p(_dptr(image))), "ctsim_fbp_slice")

    def close(self):
        if getattr(self, "handle", None):
            lib.ctsim_backprojector_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class ProcessSignal:
    """Filter plan (libctsim/procsignal.cpp:82-442) and filterSignal over
    batches of views (procsignal.cpp:544-629)."""

    def __init__(self, ctx: Context, filter: str, method: str, bw: float, signalInc: float,
                 n: int, param: float = -1.0, domain: str = "spatial",
                 generation: str = "direct", zeropad: int = 1, preinterp: int = 1,
                 geometry=0, focalLength: float = 0.0, sourceDetectorLength: float = 0.0):
        self.ctx = ctx
        g = GEOMETRY_IDS[geometry] if isinstance(geometry, str) else int(geometry)
        h = C.c_void_p()
        check(lib.ctsim_procsignal_create(
            ctx.handle, C.byref(h), filter.encode(), method.encode(), bw, signalInc, n, param,
            domain.encode(), generation.encode(), zeropad, preinterp, g, focalLength,
            sourceDetectorLength), "ctsim_procsignal_create")
        self.handle = h
        self.n = n

    def info(self):
        nfp = C.c_int()
        mmi = np.zeros(3, np.float64)
        check(lib.ctsim_procsignal_info(self.handle, C.byref(nfp), _hptr(mmi)),
              "ctsim_procsignal_info")
        return nfp.value, tuple(mmi)

    def filter_kernel(self):
        nfp, _ = self.info()
        out = np.zeros(nfp, np.float64)
        lib.ctsim_procsignal_get_filter(self.handle, _hptr(out), nfp)
        return out

    def filter(self, sino, out, precision: int = PREC_F64):
        """Filter every row of ``sino`` (device float32 [nView, n]) into ``out``
        (device [nView, n], float64 for F64, float32 for MIXED)."""
        import torch
        want = torch.float64 if precision == PREC_F64 else torch.float32
        if sino.dtype != torch.float32 or sino.shape[-1] != self.n:
            raise TypeError("sino must be float32 [nView, n]")
        if out.dtype != want or tuple(out.shape) != tuple(sino.shape):
            raise TypeError(f"out must be {want} with the shape of sino")
        check(lib.ctsim_filter_views(self.ctx.handle, self.handle, precision,
                                     C.c_void_p(_dptr(sino)), sino.shape[0],
                                     C.c_void_p(_dptr(out))), "ctsim_filter_views")

    def close(self):
        if getattr(self, "handle", None):
            lib.ctsim_procsignal_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class Phantom:
    """Phantom (libctsim/phantom.cpp): a built-in name and/or elements
    (type, cx, cy, u, v, rot_deg, atten) in definition order."""

    def __init__(self, builtin: str | None = None, elems=()):
        h = C.c_void_p()
        check(lib.ctsim_phantom_create(C.byref(h), builtin.encode() if builtin else None),
              "ctsim_phantom_create")
        self.handle = h
        for e in elems:
            self.add(*e)

    def add(self, type: str, cx, cy, u, v, rot_deg, atten):
        check(lib.ctsim_phantom_add_element(self.handle, type.encode(), cx, cy, u, v, rot_deg,
                                            atten), "ctsim_phantom_add_element")

    def rasterize(self, ctx: Context, nx: int, ny: int, viewRatio: float = 1.0, nsample: int = 1):
        """Phantom::convertToImagefile (phantom.cpp:463-538): a new device
        image float32 [nx, ny] and the axis extent (xmin, xmax, ymin, ymax)."""
        import torch
        img = torch.empty((nx, ny), dtype=torch.float32, device=f"cuda:{ctx.device}")
        ext = np.zeros(4, np.float64)
        check(lib.ctsim_phantom_rasterize(ctx.handle, self.handle, nx, ny, viewRatio, nsample,
                                          C.c_void_p(_dptr(img)), _hptr(ext)),
              "ctsim_phantom_rasterize")
        return img, tuple(ext)

    def extent(self):
        e = np.zeros(4, np.float64)
        lib.ctsim_phantom_extent(self.handle, _hptr(e))
        return e

    def close(self):
        if getattr(self, "handle", None):
            lib.ctsim_phantom_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


SCANNER_INFO = ("detInc", "detStart", "rotInc", "viewDiameter", "focalLength",
                "sourceDetectorLength", "fanBeamAngle", "phmLen", "detLen",
                "centerDetectorLength", "scanDiameter", "geometry")


class Scanner:
    """Scanner geometry (libctsim/scanner.cpp:86-224) and
    collectProjections (280-419)."""

    def __init__(self, phantom: Phantom, geometry: str, nDet: int, nView: int,
                 offsetView: int = 0, nSample: int = 1, rotLen: float = np.pi,
                 focalRatio: float = 2.0, centerRatio: float = 2.0, viewRatio: float = 1.0,
                 scanRatio: float = 1.0):
        h = C.c_void_p()
        check(lib.ctsim_scanner_create(C.byref(h), phantom.handle, geometry.encode(), nDet, nView,
                                       offsetView, nSample, rotLen, focalRatio, centerRatio,
                                       viewRatio, scanRatio), "ctsim_scanner_create")
        self.handle = h
        self.phantom = phantom
        self.nDet, self.nView, self.offsetView = nDet, nView, offsetView
        o = np.zeros(12, np.float64)
        lib.ctsim_scanner_info(h, _hptr(o))
        self.info = dict(zip(SCANNER_INFO, o))
        self.geometry = geometry

    def collect(self, ctx: Context, sino, precision: int = PREC_F64, iStartView: int = 0,
                iNumViews: int | None = None, iOffsetView: int | None = None,
                iStorageOffset: int | None = None, phantom: Phantom | None = None):
        """Ray sums into ``sino`` (device float32 [nView, nDet]); returns the
        view angles (host float64).  ``phantom`` (default: the scanner's)
        is the object scanned, as collectProjections' phm argument."""
        import torch
        if sino.dtype != torch.float32 or sino.shape[-1] != self.nDet:
            raise TypeError("sino must be float32 [nView, nDet]")
        n = self.nView if iNumViews is None else iNumViews
        off = self.offsetView if iOffsetView is None else iOffsetView
        store = iStartView if iStorageOffset is None else iStorageOffset
        if store + n > sino.shape[0]:
            raise ValueError("sinogram has too few rows")
        ang = np.zeros(sino.shape[0], np.float64)
        phm = self.phantom if phantom is None else phantom
        check(lib.ctsim_collect_projections(ctx.handle, self.handle, phm.handle,
                                            precision, iStartView, n, off, store,
                                            C.c_void_p(_dptr(sino)), _hptr(ang)),
              "ctsim_collect_projections")
        return ang

    def close(self):
        if getattr(self, "handle", None):
            lib.ctsim_scanner_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()
