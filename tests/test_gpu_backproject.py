"""GPU parity of the backprojectors against the reference's golden images.

Golden images come from the reference compiled from source
(tests/golden/make_golden.py): one filtered sinogram per geometry and odd/even
detector count, backprojected by every backprojector x interpolation.

Bars
  F64, every family         bit-exact (integer indexing, host-tabulated
                            per-view constants, no FMA contraction; Trig,
                            Table and the fan families evaluate cos, sin and
                            atan per pixel with correctly rounded device
                            routines, which agree with the C library wherever
                            the latter is correctly rounded)
  MIXED                     Herman normalised distance d <= 1e-4 against the
                            FP64 reference image (nearest interpolation and
                            Diff take their detector index in FP64 in the
                            mixed path too, see backproject.hip)
"""
import itertools

import numpy as np
import pytest

from conftest import load_golden, record_parity

pytestmark = pytest.mark.gpu

GEOMS = [("parallel", 65), ("parallel", 66), ("equilinear", 65), ("equilinear", 66),
         ("equiangular", 65), ("equiangular", 66)]


def _dre(ref, img):
    ref = ref.astype(np.float64)
    img = img.astype(np.float64)
    d = np.sqrt(np.sum((ref - img) ** 2) / np.sum((ref - ref.mean()) ** 2))
    r = np.sum(np.abs(ref - img)) / np.sum(np.abs(ref))
    return d, r


def _setup(gpu_ctx, geom, nDet, bp, interp, nx=48, ny=48, roi=None):
    import ctsim_amd
    g = load_golden(f"tiny_{geom}_{nDet}")
    info = g["info"]
    b = ctsim_amd.Backprojector(gpu_ctx, geom, nDet, float(info[0]), 60, float(info[3]),
                                float(info[4]), float(info[5]), bp, interp, nx, ny, roi=roi)
    return g, b


def _run(gpu_ctx, b, g, precision, nx=48, ny=48, post=True):
    import torch
    import ctsim_amd
    dt = torch.float64 if precision == ctsim_amd.PREC_F64 else torch.float32
    f = torch.from_numpy(g["filtered"]).to("cuda", dt)
    img = torch.zeros((nx, ny), dtype=torch.float32, device="cuda")
    b.backproject(f, g["angles"], img, precision)
    if post:
        b.postprocess(img)
    torch.cuda.synchronize()
    return img.cpu().numpy()


def _cases():
    for geom, nDet in GEOMS:
        fams = ["trig", "table", "diff", "idiff"] if geom == "parallel" else ["idiff"]
        for bp, interp in itertools.product(fams, ["nearest", "linear", "cubic"]):
            yield geom, nDet, bp, interp


@pytest.mark.parametrize("geom,nDet,bp,interp", list(_cases()))
def test_f64_matches_reference(gpu_ctx, geom, nDet, bp, interp):
    import ctsim_amd
    print(f"case f64 {geom} {nDet} {bp} {interp}", flush=True)
    g, b = _setup(gpu_ctx, geom, nDet, bp, interp)
    out = _run(gpu_ctx, b, g, ctsim_amd.PREC_F64)
    ref = g[f"image_{bp}_{interp}"]
    n = record_parity(f"golden backproject {geom} {nDet} {bp} {interp}", out, ref)
    assert n == 0, f"{n} pixels differ"


@pytest.mark.parametrize("geom,nDet,bp,interp", list(_cases()))
def test_mixed_close_to_reference(gpu_ctx, geom, nDet, bp, interp):
    import ctsim_amd
    print(f"case mixed {geom} {nDet} {bp} {interp}", flush=True)
    g, b = _setup(gpu_ctx, geom, nDet, bp, interp)
    out = _run(gpu_ctx, b, g, ctsim_amd.PREC_MIXED)
    ref = g[f"image_{bp}_{interp}"]
    d, r = _dre(ref, out)
    print(f"  d={d:.3e} r={r:.3e}")
    assert d <= 1e-4


def test_f64_unscaled_and_roi(gpu_ctx):
    import ctsim_amd
    for geom, nDet in GEOMS:
        g, b = _setup(gpu_ctx, geom, nDet, "idiff", "linear")
        out = _run(gpu_ctx, b, g, ctsim_amd.PREC_F64, post=False)
        ref = g["image_idiff_linear_nopost"]
        assert record_parity(f"golden unscaled {geom} {nDet}", out, ref) == 0
        # golden roi is (xmin, xmax, ymin, ymax); the ABI takes ReconstructionROI
        # member order (xMin, yMin, xMax, yMax)
        x0, x1, y0, y1 = g["roi"]
        g2, b2 = _setup(gpu_ctx, geom, nDet, "idiff", "cubic", 40, 36, roi=[x0, y0, x1, y1])
        out2 = _run(gpu_ctx, b2, g2, ctsim_amd.PREC_F64, 40, 36)
        ref2 = g["image_idiff_cubic_roi"]
        assert record_parity(f"golden roi {geom} {nDet}", out2, ref2) == 0


def test_view_chunks_accumulate_in_order(gpu_ctx):
    """Backprojecting views in several calls equals one call (the image keeps
    view order across calls, like repeated BackprojectView)."""
    import torch
    import ctsim_amd
    g, b = _setup(gpu_ctx, "parallel", 66, "idiff", "linear")
    f = torch.from_numpy(g["filtered"]).to("cuda", torch.float64)
    img = torch.zeros((48, 48), dtype=torch.float32, device="cuda")
    for s in range(0, 60, 7):
        b.backproject(f[s:s + 7].contiguous(), g["angles"][s:s + 7], img, ctsim_amd.PREC_F64)
    b.postprocess(img)
    out = img.cpu().numpy()
    assert np.array_equal(out.view(np.uint32), g["image_idiff_linear"].view(np.uint32))
