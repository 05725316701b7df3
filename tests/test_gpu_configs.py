"""The benchmark configurations at their workloads, on the GPU.

  config 2   equiangular 1024 det x 1440 views, abs_hanning by FFT, cubic,
             512^2: the mixed one-call slice (autotuned) within d <= 1e-4 of
             the FP64 GPU path; at 128 det x 180 views, 96^2, the same
             chain against the oracle (FP64: the FFT filter is within 1e-12
             of the oracle's restatement of FFTW, so the image agrees to
             float rounding; mixed: d <= 1e-4)
  config 3   parallel 2048 det x 2880 views, 'shepp' convolution, IntDiff
             linear, 2048^2: the mixed path (split-BF16 Toeplitz MFMA filter)
             and the FFT-convolution batch within d <= 1e-4 of FP64; on 48
             views at full detector and image size, the FP64 path bit-exact
             against the oracle (sinogram, filtered views, image)
  config 4   the 256-slice uneven batch in one graph: every slice's image
             finite, sampled slices within d <= 1e-4 of their FP64 images
  config 5   the time-varying equilinear scan by the phm2helix drop-in (one
             phantom per view from the helixphm program), FP64 180LI and
             half-scan feathering at 64 positions, one batched FBP: position
             0 within d <= 1e-4 of FP64; phm2helix's views equal the
             library's one-view collects bit for bit
"""
import math
import tempfile

import numpy as np
import pytest

from conftest import record_parity

pytestmark = pytest.mark.gpu


def _d(ctx, ref, img):
    import ctsim_amd
    return ctsim_amd.image_stats(ctx, ref, img)[0]


def test_config2_full_mixed_vs_fp64(gpu_ctx):
    import torch
    import ctsim_amd
    from ctsim_amd import fbp, workloads as W
    print("case config 2 full", flush=True)
    spec = W.config2()
    sino, ang, info = fbp.simulate(gpu_ctx, spec, ctsim_amd.PREC_MIXED)
    ref = fbp.reconstruct(gpu_ctx, spec, sino, ang, info, ctsim_amd.PREC_F64)
    ps, bp = fbp.plans(gpu_ctx, spec, info)
    f32 = torch.empty(sino.shape, dtype=torch.float32, device="cuda")
    ps.filter(sino, f32, ctsim_amd.PREC_MIXED)
    bp.autotune(f32, ang)
    img = torch.full((spec.nx, spec.ny), 9.0, dtype=torch.float32, device="cuda")
    bp.fbp(ps, sino, ang, img, ctsim_amd.PREC_MIXED)
    d = _d(gpu_ctx, ref, img)
    print(f"  config 2 mixed vs FP64: d = {d:.3e}")
    assert d <= 1e-4


@pytest.mark.parametrize("prec", ["f64", "mixed"])
def test_config2_reduced_against_oracle(gpu_ctx, oracle_lib, prec):
    import torch
    import ctsim_amd
    from ctsim_amd import workloads as W
    O = oracle_lib
    print(f"case config 2 reduced vs oracle {prec}", flush=True)
    phm_o = O.Phantom(None, W.seeded_elements(64, W.config_seed(2)))
    # the oracle's fft is FFTW's definition evaluated as a direct transform
    nDet, nView, nx = 128, 180, 96
    scan = O.ScanSpec("equiangular", nDet, nView, 0, 1, 2 * np.pi)
    rec = O.ReconSpec(filter="abs_hanning", method="fft", interp="cubic", backproj="idiff",
                      nx=nx, ny=nx)
    want = O.pipeline("oracle", phm_o, scan, rec)
    info = want["info"]
    sino = torch.from_numpy(want["sino"]).cuda()
    ps = ctsim_amd.ProcessSignal(gpu_ctx, "abs_hanning", "fft", 1 / info[0], info[0], nDet, -1.0,
                                 "spatial", "direct", 1, 1, "equiangular", info[4], info[5])
    bp = ctsim_amd.Backprojector(gpu_ctx, "equiangular", nDet, info[0], nView, info[3], info[4],
                                 info[5], "idiff", "cubic", nx, nx)
    img = torch.empty((nx, nx), dtype=torch.float32, device="cuda")
    p = ctsim_amd.PREC_F64 if prec == "f64" else ctsim_amd.PREC_MIXED
    bp.fbp(ps, sino, want["angles"], img, p)
    out = img.cpu().numpy().astype(np.float64)
    ref = want["image"].astype(np.float64)
    d = float(np.sqrt(((out - ref) ** 2).sum() / ((ref - ref.mean()) ** 2).sum()))
    rel = float(np.abs(out - ref).max() / np.abs(ref).max())
    print(f"  {prec}: d = {d:.3e}, max rel {rel:.3e}")
    if prec == "f64":
        assert rel <= 4 * 2.0 ** -24 and d <= 1e-6
    else:
        assert d <= 1e-4


def test_config3_full_mixed_and_batch_vs_fp64(gpu_ctx):
    import torch
    import ctsim_amd
    from ctsim_amd import fbp, workloads as W
    print("case config 3 full", flush=True)
    spec = W.config3()
    sino, ang, info = fbp.simulate(gpu_ctx, spec, ctsim_amd.PREC_MIXED)
    ref = fbp.reconstruct(gpu_ctx, spec, sino, ang, info, ctsim_amd.PREC_F64)
    ps, bp = fbp.plans(gpu_ctx, spec, info)
    img = torch.empty((spec.nx, spec.ny), dtype=torch.float32, device="cuda")
    bp.fbp(ps, sino, ang, img, ctsim_amd.PREC_MIXED)
    d_t = _d(gpu_ctx, ref, img)
    b = ctsim_amd.Batch(gpu_ctx, [fbp.slice_desc(spec, info, ang)])
    imgb = torch.full((spec.nx, spec.ny), 4.0, dtype=torch.float32, device="cuda")
    b.run([sino], [imgb])
    b.status()
    d_b = _d(gpu_ctx, ref, imgb)
    print(f"  config 3: Toeplitz mixed d = {d_t:.3e}, FFT-convolution batch d = {d_b:.3e}")
    assert d_t <= 1e-4 and d_b <= 1e-4


def test_config3_view_subset_fp64_against_oracle(gpu_ctx, oracle_lib):
    """2048 det, 48 views over pi, 2048^2: the config-3 phantom, filter and
    backprojector at full detector and image size on a subset of views."""
    import torch
    import ctsim_amd
    from ctsim_amd import workloads as W
    O = oracle_lib
    print("case config 3 view subset vs oracle", flush=True)
    spec = W.config3()
    phm_o = O.Phantom(None, spec.elems)
    scan = O.ScanSpec("parallel", 2048, 48, 0, 1, np.pi)
    rec = O.ReconSpec(filter="shepp", method="convolution", interp="linear", backproj="idiff",
                      nx=2048, ny=2048)
    want = O.pipeline("oracle", phm_o, scan, rec)
    info = want["info"]
    phm = ctsim_amd.Phantom(None, spec.elems)
    sc = ctsim_amd.Scanner(phm, "parallel", 2048, 48, 0, 1, np.pi)
    sino = torch.empty((48, 2048), dtype=torch.float32, device="cuda")
    ang = sc.collect(gpu_ctx, sino, ctsim_amd.PREC_F64)
    assert record_parity("oracle config 3 raysum 2048x48", sino.cpu().numpy(), want["sino"]) == 0
    ps = ctsim_amd.ProcessSignal(gpu_ctx, "shepp", "convolution", 1 / info[0], info[0], 2048)
    filt = torch.empty((48, 2048), dtype=torch.float64, device="cuda")
    ps.filter(sino, filt, ctsim_amd.PREC_F64)
    assert record_parity("oracle config 3 shepp convolution 2048x48", filt.cpu().numpy(),
                         want["filtered"]) == 0
    bp = ctsim_amd.Backprojector(gpu_ctx, "parallel", 2048, info[0], 48, info[3], info[4],
                                 info[5], "idiff", "linear", 2048, 2048)
    img = torch.zeros((2048, 2048), dtype=torch.float32, device="cuda")
    bp.backproject(filt, ang, img, ctsim_amd.PREC_F64)
    bp.postprocess(img)
    assert record_parity("oracle config 3 IntDiff linear 2048^2 x 48 views", img.cpu().numpy(),
                         want["image"]) == 0


def test_config4_batch_of_256(gpu_ctx):
    import torch
    import ctsim_amd
    from ctsim_amd import fbp, workloads as W
    print("case config 4 batch of 256", flush=True)
    specs = W.config4(256)
    sims = [fbp.simulate(gpu_ctx, s, ctsim_amd.PREC_MIXED) for s in specs]
    b = ctsim_amd.Batch(gpu_ctx, [fbp.slice_desc(s, info, ang)
                                  for s, (_, ang, info) in zip(specs, sims)])
    sinos = [x[0] for x in sims]
    images = [torch.full((s.nx, s.ny), 7.0, dtype=torch.float32, device="cuda") for s in specs]
    shape = b.autotune(sinos, images)
    b.run(sinos, images)
    b.status()
    torch.cuda.synchronize()
    print(f"  autotuned shape {shape}")
    assert all(bool(torch.isfinite(im).all()) for im in images)
    ds = []
    for k in range(0, 256, 32):
        ref = fbp.reconstruct(gpu_ctx, specs[k], *sims[k], ctsim_amd.PREC_F64)
        ds.append(_d(gpu_ctx, ref, images[k]))
        print(f"  slice {k}: {specs[k].geometry} {specs[k].nDet}x{specs[k].nView} "
              f"{specs[k].nx}^2: d = {ds[-1]:.3e}")
    assert max(ds) <= 1e-4


def test_config5_helical_chain(gpu_ctx):
    import torch
    import ctsim_amd
    from ctsim_amd import fbp, workloads as W
    print("case config 5 helical chain", flush=True)
    h = W.config5()
    base = h.base()
    phm0 = ctsim_amd.Phantom(None, h.elements(0, base))
    probe = ctsim_amd.Scanner(phm0, h.geometry, h.nDet, h.viewsPerRotation, 0, 1, 2 * np.pi,
                              h.focalRatio, h.centerRatio)
    L = int((np.pi + probe.info["fanBeamAngle"]) / (2 * np.pi / h.viewsPerRotation))
    nTotal = 2 * L + 1 + (h.positions - 1) * h.stride() + 2
    with tempfile.TemporaryDirectory() as wd:
        sino, ang, info = fbp.helical_scan(gpu_ctx, h, wd, nTotal, ctsim_amd.PREC_F64)
    # phm2helix's views are the library's one-view collects of each view's phantom
    turns = float(repr(nTotal / h.viewsPerRotation))
    sc = ctsim_amd.Scanner(phm0, h.geometry, h.nDet, nTotal, 0, 1, turns * fbp.TWO_PI,
                           h.focalRatio, h.centerRatio)
    one = torch.empty((nTotal, h.nDet), dtype=torch.float32, device="cuda")
    for v in (0, 1, nTotal // 2, nTotal - 1):
        sc.collect(gpu_ctx, one, ctsim_amd.PREC_F64, iStartView=v, iNumViews=1,
                   phantom=ctsim_amd.Phantom(None, h.elements(v, base)))
        assert torch.equal(one[v], sino[v]), f"view {v}"
    fan = info["fanBeamAngle"]
    planar = []
    for p in range(h.positions):
        o, a = ctsim_amd.helical180li(gpu_ctx, h.geometry, sino, info["rotInc"], info["detInc"],
                                      fan, info["sourceDetectorLength"], L + 1 + p * h.stride())
        planar.append((ctsim_amd.halfscan_feather(gpu_ctx, h.geometry, o, info["rotInc"],
                                                  info["detInc"], fan,
                                                  info["sourceDetectorLength"]), a))
    spec = W.SliceSpec(geometry=h.geometry, nDet=h.nDet, nView=planar[0][0].shape[0],
                       filter=h.filter, method=h.method, interp=h.interp, backproj=h.backproj,
                       nx=h.nx, ny=h.ny)
    b = ctsim_amd.Batch(gpu_ctx, [fbp.slice_desc(spec, info, a) for _, a in planar])
    images = [torch.empty((h.nx, h.ny), dtype=torch.float32, device="cuda") for _ in planar]
    b.run([p for p, _ in planar], images)
    b.status()
    ref = fbp.reconstruct(gpu_ctx, spec, planar[0][0], planar[0][1], info, ctsim_amd.PREC_F64)
    d = _d(gpu_ctx, ref, images[0])
    print(f"  config 5: {nTotal} views, {len(planar)} positions; position 0 d = {d:.3e}")
    assert d <= 1e-4
    assert all(math.isfinite(float(im.abs().max())) for im in images)
