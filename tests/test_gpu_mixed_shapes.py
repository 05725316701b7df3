"""The mixed backprojector under every launch shape the autotuner may pick
(lanes along y, rows per lane, view groups incl. the one-group-per-XCD
mapping, views per LDS chunk, windowed and full-row staging, the plain and
the rotationally symmetric kernel, the mirror kernel): each shape gives the
same image (to FP32 summation-order noise), within d <= 1e-4 of the FP64
path; and the autotuner's pick is one of them.  The symmetric kernel serves
square centred images whose views come in 4 quarter-turned copies (fan beam
over 2 pi, nView % 4 == 0), 2 quarter-turned copies (parallel over pi) or 2
half-turned copies (fan beam, nView % 4 == 2), in either rotation direction;
other view sets and odd image sizes fall back to the plain kernel.  The
mirror kernel serves centred grids of any size whose views have partners
under the grid's mirrors: all three mirrors (parallel beams over pi, fan
beams over 2 pi with an even view count), only x -> -x (fan beams with an
odd view count) or only the point mirror (views offset by a fraction of a
step); other view sets fall back to the plain kernel."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

SHAPES = [(8, 1, 1, 4), (16, 4, 8, 16), (32, 2, 2, 32), (16, 2, 3, 5), (64, 1, 16, 1),
          (16, 4, 1, 8),
          # the symmetric kernel: lanesY, P, split (of the base views), base views per
          # chunk, 1, staging KB
          (16, 1, 1, 4, 1, 0), (8, 2, 8, 16, 1, 24), (32, 2, 3, 5, 1, 0), (16, 2, 1, 1, 1, 16),
          (8, 1, 16, 2, 1, 32),
          # its software-pipelined form (windows of at most 128 cells)
          (16, 1, 1, 2, 2, 0), (8, 2, 8, 2, 2, 0), (8, 1, 16, 2, 2, 0), (32, 2, 3, 2, 2, 0),
          # the mirror kernel: lanesY, P, split, views per chunk, 3
          (16, 4, 8, 8, 3, 0), (8, 1, 1, 4, 3, 0), (32, 2, 3, 5, 3, 0), (16, 2, 16, 2, 3, 0),
          (64, 1, 1, 1, 3, 0), (8, 4, 2, 16, 3, 0)]


def _slice(ctx, geometry, nDet, nView, nx, interp, backproj, rot, filt="abs_hanning",
           method="fft", ny=None):
    import torch
    import ctsim_amd
    from ctsim_amd import workloads as W
    elems = W.seeded_elements(24, 0xC7511 + 77)
    phm = ctsim_amd.Phantom(None, elems)
    # a negative rotation length scans clockwise
    sc = ctsim_amd.Scanner(phm, geometry, nDet, nView, 0, 1, rot, 2.0, 2.0)
    info = sc.info
    sino = torch.empty((nView, nDet), dtype=torch.float32, device="cuda")
    ang = sc.collect(ctx, sino, ctsim_amd.PREC_F64)
    ps = ctsim_amd.ProcessSignal(ctx, filt, method, 1 / info["detInc"], info["detInc"], nDet,
                                 -1.0, "spatial", "direct", 1, 1, geometry, info["focalLength"],
                                 info["sourceDetectorLength"])
    bp = ctsim_amd.Backprojector(ctx, geometry, nDet, info["detInc"], nView, info["viewDiameter"],
                                 info["focalLength"], info["sourceDetectorLength"], backproj,
                                 interp, nx, ny or nx)
    f64 = torch.empty((nView, nDet), dtype=torch.float64, device="cuda")
    f32 = torch.empty((nView, nDet), dtype=torch.float32, device="cuda")
    ps.filter(sino, f64, ctsim_amd.PREC_F64)
    ps.filter(sino, f32, ctsim_amd.PREC_MIXED)
    return bp, f64, f32, ang


def _bp(ctx, bp, f, ang, nx, prec, ny=None):
    import torch
    img = torch.zeros((nx, ny or nx), dtype=torch.float32, device="cuda")
    bp.backproject(f, ang, img, prec)
    bp.postprocess(img)
    return img


CASES = [("equiangular", 512, 360, 256, "cubic", "idiff", 2 * np.pi),
         ("equilinear", 383, 300, 192, "linear", "idiff", 2 * np.pi),
         ("parallel", 367, 320, 256, "linear", "idiff", np.pi),
         ("parallel", 300, 256, 128, "cubic", "trig", np.pi),
         ("equiangular", 96, 64, 48, "linear", "table", 2 * np.pi),
         ("equiangular", 300, 198, 128, "cubic", "idiff", 2 * np.pi),   # half turns
         ("equilinear", 256, 240, 128, "cubic", "idiff", -2 * np.pi),   # clockwise
         ("equiangular", 256, 181, 96, "cubic", "idiff", 2 * np.pi),    # no symmetry
         ("equiangular", 256, 240, 95, "linear", "idiff", 2 * np.pi)]   # odd image


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[4]}-{c[5]}-{c[3]}" for c in CASES])
def test_shapes_agree(gpu_ctx, case):
    import ctsim_amd
    geometry, nDet, nView, nx, interp, backproj, rot = case
    bp, f64, f32, ang = _slice(gpu_ctx, geometry, nDet, nView, nx, interp, backproj, rot)
    ref = _bp(gpu_ctx, bp, f64, ang, nx, ctsim_amd.PREC_F64)
    scale = float(ref.abs().max())
    images = []
    for s in SHAPES:
        print(f"case {case[:6]} shape {s}", flush=True)
        try:
            bp.set_shape(*s)
        except ctsim_amd.CtsimError as exc:  # staging too large for this geometry
            assert "LDS" in str(exc)
            continue
        img = _bp(gpu_ctx, bp, f32, ang, nx, ctsim_amd.PREC_MIXED)
        d, r, e = ctsim_amd.image_stats(gpu_ctx, ref, img)
        print(f"  d={d:.3e} r={r:.3e} e={e:.3e}")
        assert d <= 1e-4
        images.append(img)
    assert len(images) >= 3
    for img in images[1:]:
        assert float((img - images[0]).abs().max()) <= 2e-5 * scale
    shape = bp.autotune(f32, ang)
    print(f"  autotuned {shape}")
    img = _bp(gpu_ctx, bp, f32, ang, nx, ctsim_amd.PREC_MIXED)
    assert float((img - images[0]).abs().max()) <= 2e-5 * scale


def test_bad_shape_rejected(gpu_ctx):
    import ctsim_amd
    g = load_golden("tiny_parallel_65")
    info = g["info"]
    b = ctsim_amd.Backprojector(gpu_ctx, "parallel", 65, float(info[0]), 60, float(info[3]),
                                float(info[4]), float(info[5]), "idiff", "linear", 48, 48)
    for s in [(12, 1, 1, 4), (16, 3, 1, 4), (16, 1, 0, 4), (16, 1, 1, 65),
              (16, 1, 1, 4, 4), (16, 1, 1, 4, 1, 63), (16, 1, 1, 4, 1, -1)]:
        with pytest.raises(ctsim_amd.CtsimError):
            b.set_shape(*s)


MIRROR_CASES = [
    # (geometry, nDet, nView, nx, ny, interp, rot, angle offset in steps, members)
    ("equiangular", 300, 181, 128, 96, "cubic", 2 * np.pi, 0.0, 2),   # Y only, not square
    ("parallel", 367, 161, 95, 128, "linear", np.pi, 0.0, 4),         # odd columns
    ("equilinear", 256, 240, 97, 97, "linear", 2 * np.pi, 0.3, 2),    # O only, odd size
    ("equiangular", 256, 120, 96, 128, "linear", 2 * np.pi, 0.0, 4),
    ("equilinear", 200, 91, 64, 64, "cubic", 2 * np.pi, 0.37, 1),     # no mirror partners
]


@pytest.mark.parametrize("case", MIRROR_CASES,
                         ids=[f"{c[0]}-{c[2]}v-{c[3]}x{c[4]}-off{c[7]}" for c in MIRROR_CASES])
def test_mirror_members(gpu_ctx, case):
    """The mirror kernel on grids that are not square or have odd sizes and
    on view sets with two, four or no mirror members (angles offset by a
    fraction of a step break the x -> -x partners): within d <= 1e-4 of the
    FP64 path and equal to the plain mixed kernel to FP32 summation noise."""
    import ctsim_amd
    geometry, nDet, nView, nx, ny, interp, rot, off, members = case
    print(f"case mirror {case}", flush=True)
    bp, f64, f32, ang = _slice(gpu_ctx, geometry, nDet, nView, nx, interp, "idiff", rot, ny=ny)
    ang = ang + off * rot / nView
    assert bp.mirror_members(ang) == members
    ref = _bp(gpu_ctx, bp, f64, ang, nx, ctsim_amd.PREC_F64, ny)
    scale = float(ref.abs().max())
    bp.set_shape(16, 2, 1, 8, 0, 0)
    plain = _bp(gpu_ctx, bp, f32, ang, nx, ctsim_amd.PREC_MIXED, ny)
    for s in [(16, 2, 1, 8, 3, 0), (8, 4, 8, 4, 3, 0), (32, 1, 3, 3, 3, 0)]:
        bp.set_shape(*s)
        img = _bp(gpu_ctx, bp, f32, ang, nx, ctsim_amd.PREC_MIXED, ny)
        d = ctsim_amd.image_stats(gpu_ctx, ref, img)[0]
        diff = float((img - plain).abs().max())
        print(f"  shape {s}: d={d:.3e} max|mirror - plain|={diff:.3e} (scale {scale:.3e})",
              flush=True)
        assert d <= 1e-4
        assert diff <= 2e-5 * scale
