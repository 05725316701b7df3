"""ctsim_fbp_slice (one call per slice) against the separate steps it
replaces (zeroed image, filter every view, backproject every view,
post-scale; reconstruct.cpp:156-239): bit-identical for F64 and for the
cases that run the separate steps (nearest, Diff), within FP32 summation-order
noise for the fused mixed path (FFT filters writing interval coefficients
from LDS, convolution filters through float rows), and within d <= 1e-4 of
the F64 image; the image is overwritten, so stale contents do not leak in."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [  # geometry, nDet, nView, nx, interp, backproj, rotation, filter, method
    ("equiangular", 512, 360, 256, "cubic", "idiff", 2 * np.pi, "abs_hanning", "fft"),
    ("equilinear", 383, 300, 192, "linear", "idiff", 2 * np.pi, "abs_cosine", "rfftw"),
    ("parallel", 367, 320, 256, "linear", "idiff", np.pi, "abs_bandlimit", "convolution"),
    ("parallel", 300, 256, 128, "cubic", "trig", np.pi, "abs_hamming", "fft"),
    ("equiangular", 96, 64, 48, "linear", "table", 2 * np.pi, "abs_bandlimit", "convolution"),
    ("parallel", 200, 180, 96, "nearest", "idiff", np.pi, "abs_bandlimit", "fft"),
    ("parallel", 160, 120, 64, "linear", "diff", np.pi, "abs_hanning", "fft"),
    # IntDiff's linear rule with its last cell included, through the FFT
    # coefficient kernel; 255 cells in 256-point transforms (more than N/2:
    # the first pass reads the whole row; 512 cells in 1024 points above skip
    # the zero half)
    ("parallel", 255, 200, 96, "linear", "idiff", np.pi, "abs_hamming", "fft"),
    # a fan-beam geometry runs its own backprojector whatever the name says
    ("equilinear", 160, 120, 64, "linear", "diff", 2 * np.pi, "abs_hanning", "fft"),
]


def _slice(ctx, case):
    import torch
    import ctsim_amd
    from ctsim_amd import workloads as W
    geometry, nDet, nView, nx, interp, backproj, rot, filt, method = case
    phm = ctsim_amd.Phantom(None, W.seeded_elements(24, 0xF0B5))
    sc = ctsim_amd.Scanner(phm, geometry, nDet, nView, 0, 1, rot, 2.0, 2.0)
    info = sc.info
    sino = torch.empty((nView, nDet), dtype=torch.float32, device="cuda")
    ang = sc.collect(ctx, sino, ctsim_amd.PREC_F64)
    ps = ctsim_amd.ProcessSignal(ctx, filt, method, 1 / info["detInc"], info["detInc"], nDet,
                                 -1.0, "spatial", "direct", 1, 1, geometry, info["focalLength"],
                                 info["sourceDetectorLength"])
    bp = ctsim_amd.Backprojector(ctx, geometry, nDet, info["detInc"], nView, info["viewDiameter"],
                                 info["focalLength"], info["sourceDetectorLength"], backproj,
                                 interp, nx, nx)
    return ps, bp, sino, ang


def _separate(ps, bp, sino, ang, nx, prec):
    import torch
    import ctsim_amd
    dt = torch.float64 if prec == ctsim_amd.PREC_F64 else torch.float32
    f = torch.empty(sino.shape, dtype=dt, device="cuda")
    ps.filter(sino, f, prec)
    img = torch.zeros((nx, nx), dtype=torch.float32, device="cuda")
    bp.backproject(f, ang, img, prec)
    bp.postprocess(img)
    return img


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[4]}-{c[5]}-{c[8]}" for c in CASES])
def test_fbp_slice_matches_steps(gpu_ctx, case):
    import torch
    import ctsim_amd
    nx = case[3]
    print(f"case {case}", flush=True)
    ps, bp, sino, ang = _slice(gpu_ctx, case)
    ref64 = _separate(ps, bp, sino, ang, nx, ctsim_amd.PREC_F64)
    img = torch.full((nx, nx), 7.0, dtype=torch.float32, device="cuda")  # stale contents
    bp.fbp(ps, sino, ang, img, ctsim_amd.PREC_F64)
    torch.cuda.synchronize()
    assert torch.equal(img, ref64)

    sep = _separate(ps, bp, sino, ang, nx, ctsim_amd.PREC_MIXED)
    for rep in range(2):  # a second call reuses the context's scratch
        img.fill_(-3.0)
        bp.fbp(ps, sino, ang, img, ctsim_amd.PREC_MIXED)
        torch.cuda.synchronize()
        if case[4] == "nearest" or (case[5] == "diff" and case[0] == "parallel"):
            assert torch.equal(img, sep)
        else:
            scale = float(sep.abs().max())
            assert float((img - sep).abs().max()) <= 2e-5 * scale
        d, r, e = ctsim_amd.image_stats(gpu_ctx, ref64, img)
        print(f"  rep {rep}: d={d:.3e} r={r:.3e} e={e:.3e}")
        assert d <= 1e-4


def test_fbp_slice_rejects_mismatch(gpu_ctx):
    import torch
    import ctsim_amd
    ps, _, _, _ = _slice(gpu_ctx, CASES[0])        # 512-point filter
    _, bp2, sino, ang = _slice(gpu_ctx, CASES[3])  # 300-detector backprojector
    img = torch.zeros((bp2.nx, bp2.ny), dtype=torch.float32, device="cuda")
    with pytest.raises(ctsim_amd.CtsimError, match="detector counts differ"):
        bp2.fbp(ps, sino, ang, img)
