"""Host helpers that chain the library's handles the way tools/phm2pj.cpp
and tools/pjrec.cpp chain the reference classes: simulate a slice's
sinogram on the GPU, reconstruct it on the FP64 or mixed single-slice path,
or describe it for a batch.  All compute runs in libctsim_amd.so."""
from __future__ import annotations

import numpy as np

from . import api
from .workloads import SliceSpec


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
