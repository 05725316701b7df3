"""Host helpers that chain the library's handles the way tools/phm2pj.cpp
and tools/pjrec.cpp chain the reference classes: simulate a slice's
sinogram on the GPU, reconstruct it on the FP64 or mixed single-slice path,
or describe it for a batch.  All compute runs in libctsim_amd.so."""
from __future__ import annotations

import os
import subprocess

import numpy as np

from . import api
from .workloads import HelicalSpec, SliceSpec

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bin")
TWO_PI = 6.28318530717958647692  # the reference's TWOPI (include/ctsupport.h)


def simulate(ctx: api.Context, spec: SliceSpec, precision: int = api.PREC_F64):
    """Scanner::collectProjections of the spec's phantom into a device
    sinogram float32 [nView, nDet]; returns (sinogram, angles, scanner info)."""
    import torch
    phm = api.Phantom(spec.builtin, spec.elems)
    sc = api.Scanner(phm, spec.geometry, spec.nDet, spec.nView, 0, spec.nSample, spec.rotLen,
                     spec.focalRatio, spec.centerRatio, spec.viewRatio, spec.scanRatio)
    sino = torch.empty((spec.nView, spec.nDet), dtype=torch.float32, device=f"cuda:{ctx.device}")
    ang = sc.collect(ctx, sino, precision)
    return sino, ang, sc.info


def plans(ctx: api.Context, spec: SliceSpec, info: dict):
    """The ProcessSignal and Backprojector of one slice (reconstruct.cpp:59-100)."""
    ps = api.ProcessSignal(ctx, spec.filter, spec.method, 1 / info["detInc"], info["detInc"],
                           spec.nDet, spec.filterParam, "spatial", spec.generation, spec.zeropad,
                           1, spec.geometry, info["focalLength"], info["sourceDetectorLength"])
    bp = api.Backprojector(ctx, spec.geometry, spec.nDet, info["detInc"], spec.nView,
                           info["viewDiameter"], info["focalLength"],
                           info["sourceDetectorLength"], spec.backproj, spec.interp, spec.nx,
                           spec.ny)
    return ps, bp


def reconstruct(ctx: api.Context, spec: SliceSpec, sino, angles, info, precision: int,
                image=None):
    """Reconstructor::reconstructAllViews on the device (filter, backproject,
    post-scale) into a float32 [nx, ny] image, as separate steps."""
    import torch
    dev = sino.device
    ps, bp = plans(ctx, spec, info)
    dt = torch.float64 if precision == api.PREC_F64 else torch.float32
    filt = torch.empty((spec.nView, spec.nDet), dtype=dt, device=dev)
    if image is None:
        image = torch.zeros((spec.nx, spec.ny), dtype=torch.float32, device=dev)
    else:
        image.zero_()
    ps.filter(sino, filt, precision)
    bp.backproject(filt, angles, image, precision)
    bp.postprocess(image)
    return image


def reconstruct_fused(ctx: api.Context, spec: SliceSpec, sino, angles, info, precision: int,
                      image=None):
    """The same reconstruction through the one-call slice entry point."""
    import torch
    ps, bp = plans(ctx, spec, info)
    if image is None:
        image = torch.empty((spec.nx, spec.ny), dtype=torch.float32, device=sino.device)
    bp.fbp(ps, sino, angles, image, precision)
    return image


def slice_desc(spec: SliceSpec, info: dict, angles) -> dict:
    """The batch description of one slice (ctsim_slice_desc)."""
    return {"geometry": spec.geometry, "nDet": spec.nDet, "nView": spec.nView,
            "detInc": info["detInc"], "viewDiameter": info["viewDiameter"],
            "focalLength": info["focalLength"],
            "sourceDetectorLength": info["sourceDetectorLength"], "filter": spec.filter,
            "method": spec.method, "filterParam": spec.filterParam,
            "generation": spec.generation, "zeropad": spec.zeropad, "backproj": spec.backproj,
            "interp": spec.interp, "nx": spec.nx, "ny": spec.ny,
            "angles": np.asarray(angles, np.float64)}


def read_pj(path: str):
    """A .pj projection file's views and view angles (the layout of
    Projections::write, projections.cpp:485-546, 816-889): (float32
    [nView, nDet], float64 [nView])."""
    b = open(path, "rb").read()
    hsize = int(np.frombuffer(b, "<u2", 1, 0)[0])
    nView, nDet = (int(x) for x in np.frombuffer(b, "<u4", 2, 4))
    data = np.zeros((nView, nDet), np.float32)
    ang = np.zeros(nView, np.float64)
    off = hsize
    for v in range(nView):
        ang[v] = np.frombuffer(b, "<f8", 1, off)[0]
        data[v] = np.frombuffer(b, "<f4", nDet, off + 12)
        off += 12 + 4 * nDet
    return data, ang


def helical_scan(ctx: api.Context, h: HelicalSpec, workdir: str, nTotal: int,
                 precision: int = api.PREC_F64):
    """The time-varying scan of ``h`` over ``nTotal`` views by the phm2helix
    tool (one phantom per view from the helixphm program, as
    tools/phm2helix.cpp:287-315 runs dynphm); returns (device sinogram
    float32 [nTotal, nDet], angles, info of the Scanner the tool builds)."""
    import torch
    base = os.path.join(workdir, "helix_base.txt")
    h.write_program_base(base)
    pj = os.path.join(workdir, "helix.pj")
    turns = repr(nTotal / h.viewsPerRotation)
    prog = (f"{os.path.join(BIN, 'helixphm')} --base {base} --period {h.viewsPerRotation} "
            f"--drift {h.drift!r}")
    subprocess.run([os.path.join(BIN, "phm2helix"), pj, str(h.nDet), str(nTotal), prog,
                    "--geometry", h.geometry, "--rotangle", turns, "--focal-length",
                    repr(h.focalRatio), "--center-detector-length", repr(h.centerRatio),
                    "--phmfile", os.path.join(workdir, "helix_phm.txt"), "--precision",
                    "f64" if precision == api.PREC_F64 else "mixed", "--device", str(ctx.device)],
                   check=True, capture_output=True, timeout=600)
    data, ang = read_pj(pj)
    # the tool's Scanner: the view-0 phantom, rotation turns * TWOPI
    phm0 = api.Phantom(None, h.elements(0, h.base()))
    sc = api.Scanner(phm0, h.geometry, h.nDet, nTotal, 0, 1, float(turns) * TWO_PI,
                     h.focalRatio, h.centerRatio)
    sino = torch.from_numpy(data).to(f"cuda:{ctx.device}")
    return sino, ang, sc.info
