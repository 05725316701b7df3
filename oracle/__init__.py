"""ctypes front end to the CPU oracle (TEST INFRASTRUCTURE ONLY).

Two libraries share one calling convention:

* ``liboracle.so``           -- this repository's plain-C restatement
                                (oracle/ctsim_oracle.c)
* ``_ref/libctsimref.so``    -- the CTSim reference itself, compiled from its
                                own sources by oracle/Makefile, driven through
                                oracle/ref_harness.cpp

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this module.  The product package ``ctsim_amd`` never
does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libctsimref.so")
REF_SRC = "/root/reference"

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

# element type ids, include/phantom.h:33-38
ELEM_TYPES = {"rectangle": 1, "triangle": 2, "ellipse": 3, "sector": 4, "segment": 5}
# geometry ids, libctsim/scanner.cpp:29-33
GEOMETRY_IDS = {"parallel": 0, "equiangular": 1, "equilinear": 2}


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so, and the reference library when its sources exist."""
    subprocess.check_call(["make", "-s", "-C", HERE, "oracle"])
    if ref is None:
        ref = os.path.isdir(REF_SRC)
    if ref:
        subprocess.check_call(["make", "-s", "-j8", "-C", HERE, "ref", f"REF={REF_SRC}"])


def _bind(lib: C.CDLL, prefix: str) -> C.CDLL:
    c = C.c_char_p
    i = C.c_int
    d = C.c_double
    vp = C.c_void_p
    f = getattr(lib, f"{prefix}_pipeline")
    f.restype = i
    f.argtypes = [vp, vp, i, c, c, i, i, i, i, d, d, d, d, d, c, d, c, i, c, c, i, c, i, i,
                  vp, vp, vp, vp, vp, vp, vp]
    f = getattr(lib, f"{prefix}_filter")
    f.restype = i
    f.argtypes = [c, c, d, d, i, d, c, c, i, i, i, d, d, vp, i, vp, vp, i, vp, vp]
    f = getattr(lib, f"{prefix}_backproject")
    f.restype = i
    f.argtypes = [vp, vp, i, c, c, i, i, i, i, d, d, d, d, d, c, c, i, i, i, vp, vp, vp, i, vp]
    f = getattr(lib, f"{prefix}_stats")
    f.restype = i
    f.argtypes = [i, i, vp, vp, vp]
    getattr(lib, f"{prefix}_last_error").restype = C.c_char_p
    return lib


_libs: dict[str, C.CDLL] = {}


def lib(kind: str = "oracle") -> C.CDLL:
    """kind = 'oracle' (C restatement) or 'ref' (reference built from source)."""
    if kind not in _libs:
        path = ORACLE_SO if kind == "oracle" else REF_SO
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built (run oracle.build())")
        _libs[kind] = _bind(C.CDLL(path), "orc" if kind == "oracle" else "ref")
    return _libs[kind]


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _b(s):
    return None if s is None else s.encode()


@dataclass
class Phantom:
    """A phantom: optional built-in name plus explicit elements in definition
    order, (type, cx, cy, u, v, rot_deg, atten) as in the .phm file format
    (libctsim/phantom.cpp:184-213)."""
    builtin: str | None = None
    elems: list = field(default_factory=list)

    def arrays(self):
        n = len(self.elems)
        t = np.array([ELEM_TYPES[e[0]] for e in self.elems] or [0], dtype=np.int32)
        p = np.array([e[1:] for e in self.elems] or [[0.0] * 6], dtype=np.float64).reshape(-1, 6)
        return t, np.ascontiguousarray(p), n


@dataclass
class ScanSpec:
    """Scanner arguments, libctsim/scanner.cpp:86-91 (rotLen in radians)."""
    geometry: str = "parallel"
    nDet: int = 65
    nView: int = 60
    offsetView: int = 0
    nSample: int = 1
    rotLen: float = np.pi
    focalRatio: float = 2.0
    centerRatio: float = 2.0
    viewRatio: float = 1.0
    scanRatio: float = 1.0


@dataclass
class ReconSpec:
    """Reconstructor arguments, libctsim/reconstruct.cpp:50-54 (pjrec defaults
    tools/pjrec.cpp:128-137)."""
    filter: str = "abs_bandlimit"
    filterParam: float = -1.0
    method: str = "convolution"
    zeropad: int = 1
    generation: str = "direct"
    interp: str = "linear"
    interpFactor: int = 1
    backproj: str = "idiff"
    nx: int = 64
    ny: int = 64


def pipeline(kind: str, phm: Phantom, scan: ScanSpec, rec: ReconSpec | None = None,
             want_recon: bool = False):
    """Run phantom -> projections -> filter -> backprojection on the CPU.

    Returns a dict with sino [nView, nDet] f32, angles [nView] f64,
    filtered [nView, nDet] f64, image [nx, ny] f32 (column-major layout),
    info (16 geometry numbers) and timings (raysum, filter, backproject s).
    """
    L = lib(kind)
    pre = "orc" if kind == "oracle" else "ref"
    t, p, n = phm.arrays()
    sino = np.zeros((scan.nView, scan.nDet), np.float32)
    ang = np.zeros(scan.nView, np.float64)
    info = np.zeros(16, np.float64)
    tim = np.zeros(3, np.float64)
    out = {"sino": sino, "angles": ang, "info": info, "timings": tim}
    if rec is None:
        rc = getattr(L, f"{pre}_pipeline")(
            _ptr(t), _ptr(p), n, _b(phm.builtin), _b(scan.geometry), scan.nDet, scan.nView,
            scan.offsetView, scan.nSample, scan.rotLen, scan.focalRatio, scan.centerRatio,
            scan.viewRatio, scan.scanRatio, None, 0.0, None, 0, None, None, 1, None, 0, 0,
            _ptr(sino), _ptr(ang), None, None, None, _ptr(info), _ptr(tim))
    else:
        nf = scan.nDet * rec.interpFactor
        filt = np.zeros((scan.nView, nf), np.float64)
        img = np.zeros((rec.nx, rec.ny), np.float32)
        img2 = np.zeros((rec.nx, rec.ny), np.float32) if want_recon else None
        rc = getattr(L, f"{pre}_pipeline")(
            _ptr(t), _ptr(p), n, _b(phm.builtin), _b(scan.geometry), scan.nDet, scan.nView,
            scan.offsetView, scan.nSample, scan.rotLen, scan.focalRatio, scan.centerRatio,
            scan.viewRatio, scan.scanRatio, _b(rec.filter), rec.filterParam, _b(rec.method),
            rec.zeropad, _b(rec.generation), _b(rec.interp), rec.interpFactor, _b(rec.backproj),
            rec.nx, rec.ny, _ptr(sino), _ptr(ang), _ptr(filt), _ptr(img), _ptr(img2),
            _ptr(info), _ptr(tim))
        out.update(filtered=filt, image=img, image_recon=img2)
    if rc:
        raise RuntimeError(f"{kind} pipeline failed ({rc}): "
                           f"{getattr(L, pre + '_last_error')().decode()}")
    return out


def filter_views(kind: str, sino: np.ndarray, filter: str, method: str, bw: float, inc: float,
                 param: float = -1.0, domain: str = "spatial", generation: str = "direct",
                 zeropad: int = 1, preinterp: int = 1, geometry: int = 0,
                 focal: float = 0.0, sdl: float = 0.0):
    """ProcessSignal(...).filterSignal on every row of ``sino``; returns
    (filtered [nView, n] f64, filter kernel f64, (min, max, inc))."""
    L = lib(kind)
    pre = "orc" if kind == "oracle" else "ref"
    sino = np.ascontiguousarray(sino, dtype=np.float32)
    nView, n = sino.shape
    out = np.zeros((nView, n), np.float64)
    cap = 1 << 16
    kern = np.zeros(cap, np.float64)
    nfp = C.c_int(0)
    mmi = np.zeros(3, np.float64)
    rc = getattr(L, f"{pre}_filter")(
        _b(filter), _b(method), bw, inc, n, param, _b(domain), _b(generation), zeropad,
        preinterp, geometry, focal, sdl, _ptr(sino), nView, _ptr(out), _ptr(kern), cap,
        C.byref(nfp), _ptr(mmi))
    if rc:
        raise RuntimeError(f"{kind} filter failed: {getattr(L, pre + '_last_error')().decode()}")
    return out, kern[: nfp.value].copy(), tuple(mmi)


def backproject(kind: str, phm: Phantom, scan: ScanSpec, backproj: str, interp: str, nx: int,
                ny: int, filtered: np.ndarray, angles: np.ndarray, post: bool = True,
                roi=None, interpFactor: int = 1):
    L = lib(kind)
    pre = "orc" if kind == "oracle" else "ref"
    t, p, n = phm.arrays()
    img = np.zeros((nx, ny), np.float32)
    filtered = np.ascontiguousarray(filtered, dtype=np.float64)
    angles = np.ascontiguousarray(angles, dtype=np.float64)
    roi_a = None if roi is None else np.ascontiguousarray(roi, dtype=np.float64)
    rc = getattr(L, f"{pre}_backproject")(
        _ptr(t), _ptr(p), n, _b(phm.builtin), _b(scan.geometry), scan.nDet, scan.nView,
        scan.offsetView, scan.nSample, scan.rotLen, scan.focalRatio, scan.centerRatio,
        scan.viewRatio, scan.scanRatio, _b(backproj), _b(interp), interpFactor, nx, ny,
        _ptr(roi_a), _ptr(filtered), _ptr(angles), int(post), _ptr(img))
    if rc:
        raise RuntimeError(f"{kind} backproject failed: "
                           f"{getattr(L, pre + '_last_error')().decode()}")
    return img


def stats(kind: str, ref_img: np.ndarray, img: np.ndarray):
    """Herman d, r, e with ``ref_img`` as the reference (imagefile.cpp:185-240)."""
    L = lib(kind)
    pre = "orc" if kind == "oracle" else "ref"
    a = np.ascontiguousarray(ref_img, dtype=np.float32)
    b = np.ascontiguousarray(img, dtype=np.float32)
    dre = np.zeros(3, np.float64)
    getattr(L, f"{pre}_stats")(a.shape[0], a.shape[1], _ptr(a), _ptr(b), _ptr(dre))
    return tuple(dre)


def _bind_helpers(L: C.CDLL) -> C.CDLL:
    d, i, vp, c = C.c_double, C.c_int, C.c_void_p, C.c_char_p
    L.orc_frequency_response.restype = d
    L.orc_frequency_response.argtypes = [i, d, d, d]
    L.orc_spatial_response_analytic.restype = d
    L.orc_spatial_response_analytic.argtypes = [i, d, d, d]
    L.orc_spatial_response_calc.restype = d
    L.orc_spatial_response_calc.argtypes = [i, d, d, d, i]
    L.orc_integrate_simpson.restype = d
    L.orc_integrate_simpson.argtypes = [d, d, vp, i]
    L.orc_cubic_interpolate.restype = d
    L.orc_cubic_interpolate.argtypes = [vp, i, d]
    L.orc_add_zeropad_factor.restype = i
    L.orc_add_zeropad_factor.argtypes = [i, i]
    L.orc_raysum_line.restype = d
    L.orc_raysum_line.argtypes = [vp, vp, i, c, d, d, d, d]
    L.orc_phantom_extent.restype = i
    L.orc_phantom_extent.argtypes = [vp, vp, i, c, vp]
    return L


def helpers() -> C.CDLL:
    """The oracle's building blocks (closed-form checks)."""
    L = lib("oracle")
    if not getattr(L, "_helpers_bound", False):
        _bind_helpers(L)
        L._helpers_bound = True
    return L


def raysum_line(phm: Phantom, x1, y1, x2, y2) -> float:
    t, p, n = phm.arrays()
    return helpers().orc_raysum_line(_ptr(t), _ptr(p), n, _b(phm.builtin), x1, y1, x2, y2)


def phantom_extent(phm: Phantom):
    t, p, n = phm.arrays()
    ext = np.zeros(4, np.float64)
    helpers().orc_phantom_extent(_ptr(t), _ptr(p), n, _b(phm.builtin), _ptr(ext))
    return ext


def fbp_sample(kind: str, phm: Phantom, geometry: str, nDet: int, nView: int, nSample: int,
               rotLen: float, focalRatio: float, centerRatio: float, filter: str, method: str,
               zeropad: int, interp: str, backproj: str, nx: int, ny: int, sino: np.ndarray,
               angles: np.ndarray, nUse: int):
    """Filter + backproject the first nUse views on the CPU; returns
    (image, filter seconds, backprojection seconds)."""
    L = lib(kind)
    pre = "orc" if kind == "oracle" else "ref"
    f = getattr(L, f"{pre}_fbp_sample")
    c, i, d, vp = C.c_char_p, C.c_int, C.c_double, C.c_void_p
    f.restype = i
    f.argtypes = [vp, vp, i, c, c, i, i, i, d, d, d, c, c, i, c, c, i, i, vp, vp, i, vp, vp]
    t, p, n = phm.arrays()
    img = np.zeros((nx, ny), np.float32)
    tim = np.zeros(2, np.float64)
    sino = np.ascontiguousarray(sino, np.float32)
    angles = np.ascontiguousarray(angles, np.float64)
    rc = f(_ptr(t), _ptr(p), n, _b(phm.builtin), _b(geometry), nDet, nView, nSample, rotLen,
           focalRatio, centerRatio, _b(filter), _b(method), zeropad, _b(interp), _b(backproj), nx,
           ny, _ptr(sino), _ptr(angles), nUse, _ptr(img), _ptr(tim))
    if rc:
        raise RuntimeError(f"{kind} fbp_sample failed: {getattr(L, pre + '_last_error')().decode()}")
    return img, tim[0], tim[1]


def fbp_threads(phm: Phantom, geometry: str, nDet: int, nView: int, nSample: int, rotLen: float,
                focalRatio: float, centerRatio: float, filter: str, method: str, zeropad: int,
                interp: str, backproj: str, nx: int, ny: int, sino: np.ndarray,
                angles: np.ndarray, nUse: int, nThreads: int):
    """The reference's all-cores reconstruction (per-thread Reconstructor over
    contiguous view blocks of views [0, nUse), float sum of the images);
    returns (image, wall seconds)."""
    L = lib("ref")
    f = L.ref_fbp_threads
    c, i, d, vp = C.c_char_p, C.c_int, C.c_double, C.c_void_p
    f.restype = i
    f.argtypes = [vp, vp, i, c, c, i, i, i, d, d, d, c, c, i, c, c, i, i, vp, vp, i, i, vp, vp]
    t, p, n = phm.arrays()
    img = np.zeros((nx, ny), np.float32)
    sec = np.zeros(1, np.float64)
    sino = np.ascontiguousarray(sino, np.float32)
    angles = np.ascontiguousarray(angles, np.float64)
    rc = f(_ptr(t), _ptr(p), n, _b(phm.builtin), _b(geometry), nDet, nView, nSample, rotLen,
           focalRatio, centerRatio, _b(filter), _b(method), zeropad, _b(interp), _b(backproj), nx,
           ny, _ptr(sino), _ptr(angles), nUse, nThreads, _ptr(img), _ptr(sec))
    if rc:
        raise RuntimeError(f"ref fbp_threads failed ({rc}): {L.ref_last_error().decode()}")
    return img, float(sec[0])


def ref_write_pj(path: str, builtin: str, geometry: str, nDet: int, nView: int, nSample: int,
                 rotLen: float, focalRatio: float, centerRatio: float, data: np.ndarray,
                 angles: np.ndarray, calcTime: float, remark: str) -> None:
    """Write a .pj file with the reference's Projections::write, geometry
    from the reference's Scanner on a built-in phantom."""
    L = lib("ref")
    f = L.ref_write_pj
    c, i, d, vp = C.c_char_p, C.c_int, C.c_double, C.c_void_p
    f.restype = i
    f.argtypes = [c, c, c, i, i, i, d, d, d, vp, vp, d, c]
    data = np.ascontiguousarray(data, np.float32)
    angles = np.ascontiguousarray(angles, np.float64)
    rc = f(_b(path), _b(builtin), _b(geometry), nDet, nView, nSample, rotLen, focalRatio,
           centerRatio, _ptr(data), _ptr(angles), calcTime, _b(remark))
    if rc:
        raise RuntimeError(f"ref_write_pj failed ({rc}): {L.ref_last_error().decode()}")


def ref_read_pj(path: str, cap_views: int, nDet: int):
    """Read a .pj file with the reference's Projections::read; returns
    (info[12], data[nView][nDet], angles)."""
    L = lib("ref")
    f = L.ref_read_pj
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    info = np.zeros(12, np.float64)
    data = np.zeros((cap_views, nDet), np.float32)
    ang = np.zeros(cap_views, np.float64)
    if f(_b(path), _ptr(info), _ptr(data), _ptr(ang), cap_views):
        raise RuntimeError(f"ref_read_pj failed: {L.ref_last_error().decode()}")
    return info, data, ang


def ref_write_if(path: str, image: np.ndarray, inc, extent, label: str, calcTime: float) -> None:
    """Write an [nx][ny] float32 image with the reference's ImageFile::fileWrite."""
    L = lib("ref")
    f = L.ref_write_if
    d = C.c_double
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_void_p, d, d, d, d, d, d, C.c_char_p, d]
    im = np.ascontiguousarray(image, np.float32)
    nx, ny = im.shape
    if f(_b(path), nx, ny, _ptr(im), inc[0], inc[1], *extent, _b(label), calcTime):
        raise RuntimeError("ref_write_if failed")


def ref_read_if(path: str, cap: int) -> np.ndarray:
    """Read a float32 .if image with the reference's ImageFile::fileRead; [nx][ny]."""
    L = lib("ref")
    f = L.ref_read_if
    f.restype = C.c_int
    f.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_int]
    nxny = np.zeros(2, np.int32)
    buf = np.zeros(cap, np.float32)
    if f(_b(path), _ptr(nxny), _ptr(buf), cap):
        raise RuntimeError(f"ref_read_if failed: {L.ref_last_error().decode()}")
    nx, ny = int(nxny[0]), int(nxny[1])
    return buf[: nx * ny].reshape(nx, ny)


def helical180li(geom: int, data: np.ndarray, rotInc: float, detInc: float, fanAngle: float,
                 sdl: float, interpView: int = -1, feather: bool = False):
    """Oracle Helical180LI (+ HalfScanFeather): returns (views, angles) or
    raises with the oracle's message.  geom 1 equiangular, 2 equilinear."""
    L = lib("oracle")
    f = L.orc_helical180li
    f.restype = C.c_int
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                  C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    data = np.ascontiguousarray(data, np.float32)
    nView, nDet = data.shape
    cap = int((np.pi + fanAngle) / rotInc) + 2
    out = np.zeros((cap, nDet), np.float32)
    ang = np.zeros(cap, np.float64)
    n = f(geom, nDet, nView, rotInc, detInc, fanAngle, sdl, interpView, _ptr(data), _ptr(out),
          _ptr(ang))
    if n < 0:
        raise RuntimeError(f"orc_helical180li: {L.orc_last_error().decode()}")
    out, ang = out[:n].copy(), ang[:n].copy()
    if feather:
        g = L.orc_halfscan_feather
        g.restype = C.c_int
        g.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                      C.c_void_p]
        rc = g(geom, nDet, n, rotInc, detInc, fanAngle, sdl, _ptr(out))
        if rc:
            raise RuntimeError(f"orc_halfscan_feather: {L.orc_last_error().decode()}")
    return out, ang


def ref_helical(builtin: str, geometry: str, nDet: int, nView: int, nSample: int, rotLen: float,
                focalRatio: float, centerRatio: float, data: np.ndarray, interpView: int = -1,
                feather: bool = False):
    """The reference's Helical180LI (+ HalfScanFeather) on ``data``; returns
    (views, angles, (rotInc, detInc, fanAngle, sdl))."""
    L = lib("ref")
    f = L.ref_helical
    c, i, d, vp = C.c_char_p, C.c_int, C.c_double, C.c_void_p
    f.restype = i
    f.argtypes = [c, c, i, i, i, d, d, d, vp, i, i, vp, vp, i, vp]
    data = np.ascontiguousarray(data, np.float32)
    cap = nView
    out = np.zeros((cap, nDet), np.float32)
    ang = np.zeros(cap, np.float64)
    info = np.zeros(4, np.float64)
    n = f(_b(builtin), _b(geometry), nDet, nView, nSample, rotLen, focalRatio, centerRatio,
          _ptr(data), interpView, int(feather), _ptr(out), _ptr(ang), cap, _ptr(info))
    if n < 0:
        raise RuntimeError(f"ref_helical failed ({n}): {L.ref_last_error().decode()}")
    return out[:n].copy(), ang[:n].copy(), tuple(info)


def interpolate_to_parallel(geom: int, data: np.ndarray, detStart: float, detInc: float,
                            focalLength: float, sdl: float, viewDiameter: float,
                            rotStart: float, angles: np.ndarray):
    """Oracle Projections::interpolateToParallel; returns (views, angles, info6)."""
    L = lib("oracle")
    f = L.orc_interpolate_to_parallel
    d, i, vp = C.c_double, C.c_int, C.c_void_p
    f.restype = i
    f.argtypes = [i, i, i, d, d, d, d, d, d, vp, vp, vp, vp, vp]
    data = np.ascontiguousarray(data, np.float32)
    nView, nDet = data.shape
    angles = np.ascontiguousarray(angles, np.float64)
    out = np.zeros_like(data)
    ang = np.zeros(nView, np.float64)
    info = np.zeros(6, np.float64)
    f(geom, nDet, nView, detStart, detInc, focalLength, sdl, viewDiameter, rotStart,
      _ptr(angles), _ptr(data), _ptr(out), _ptr(ang), _ptr(info))
    return out, ang, info


def ref_rebin(builtin: str, geometry: str, nDet: int, nView: int, rotLen: float,
              focalRatio: float, centerRatio: float, data: np.ndarray):
    """The reference's interpolateToParallel; returns (views, angles_in,
    angles_out, info8)."""
    L = lib("ref")
    f = L.ref_rebin
    c, i, d, vp = C.c_char_p, C.c_int, C.c_double, C.c_void_p
    f.restype = i
    f.argtypes = [c, c, i, i, d, d, d, vp, vp, vp, vp, vp]
    data = np.ascontiguousarray(data, np.float32)
    out = np.zeros_like(data)
    ain = np.zeros(nView, np.float64)
    aout = np.zeros(nView, np.float64)
    info = np.zeros(8, np.float64)
    rc = f(_b(builtin), _b(geometry), nDet, nView, rotLen, focalRatio, centerRatio, _ptr(data),
           _ptr(out), _ptr(ain), _ptr(aout), _ptr(info))
    if rc:
        raise RuntimeError(f"ref_rebin failed: {L.ref_last_error().decode()}")
    return out, ain, aout, info


def rasterize(kind: str, phm: Phantom, nx: int, ny: int, viewRatio: float = 1.0,
              nsample: int = 1) -> np.ndarray:
    """Phantom::convertToImagefile (phm2if): returns the [nx][ny] image."""
    L = lib(kind)
    pre = "orc" if kind == "oracle" else "ref"
    f = getattr(L, f"{pre}_rasterize")
    c, i, d, vp = C.c_char_p, C.c_int, C.c_double, C.c_void_p
    f.restype = i
    t, p, n = phm.arrays()
    img = np.zeros((nx, ny), np.float32)
    if kind == "oracle":
        f.argtypes = [vp, vp, i, c, i, i, d, i, vp, vp]
        ext = np.zeros(4, np.float64)
        f(_ptr(t), _ptr(p), n, _b(phm.builtin), nx, ny, viewRatio, nsample, _ptr(img), _ptr(ext))
    else:
        f.argtypes = [vp, vp, i, c, i, i, d, i, vp]
        f(_ptr(t), _ptr(p), n, _b(phm.builtin), nx, ny, viewRatio, nsample, _ptr(img))
    return img
