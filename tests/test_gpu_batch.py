"""Batched mixed FBP (one persistent launch in a HIP graph) against the
single-slice paths.  Bars: every slice within d <= 1e-5 of the FP64
single-slice image (Herman's d; the batch filters convolutions by FP32 FFT
convolution, measured d ~ 1e-6) and within d <= 1e-4 of the single-slice
mixed image (which filters convolutions by the BF16x3 Toeplitz product);
every launch shape, with and without the rotationally symmetric body and
the mirror body, gives the same images to FP32 summation-order noise;
graph replays are bitwise identical; new image buffers are picked up."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _specs():
    from ctsim_amd import workloads as W
    out = []
    rows = [("parallel", 183, 90, 128, "linear", "convolution", "abs_bandlimit"),
            ("equilinear", 367, 300, 256, "linear", "convolution", "shepp"),
            ("equiangular", 512, 360, 256, "cubic", "fft", "abs_hanning"),
            ("parallel", 401, 200, 128, "cubic", "fourier", "abs_hamming"),
            ("equiangular", 255, 181, 384, "linear", "convolution", "abs_cosine"),
            ("equilinear", 300, 97, 128, "cubic", "fft", "abs_cosine"),
            ("parallel", 1024, 720, 512, "linear", "convolution", "abs_bandlimit")]
    for k, (g, nd, nv, nx, interp, meth, filt) in enumerate(rows):
        out.append(W.SliceSpec(elems=W.seeded_elements(16 + 4 * k, 0xC7511 + 900 + k), geometry=g,
                               nDet=nd, nView=nv, rotLen=np.pi if g == "parallel" else 2 * np.pi,
                               filter=filt, method=meth, interp=interp, backproj="idiff",
                               filterParam=0.54 if "hamming" in filt else -1.0, nx=nx, ny=nx))
    return out


@pytest.fixture(scope="module")
def batch_case(gpu_ctx):
    import torch
    import ctsim_amd
    from ctsim_amd import fbp
    specs = _specs()
    sims = [fbp.simulate(gpu_ctx, s) for s in specs]
    ref64 = [fbp.reconstruct(gpu_ctx, s, sino, ang, info, ctsim_amd.PREC_F64)
             for s, (sino, ang, info) in zip(specs, sims)]
    refmx = [fbp.reconstruct(gpu_ctx, s, sino, ang, info, ctsim_amd.PREC_MIXED)
             for s, (sino, ang, info) in zip(specs, sims)]
    b = ctsim_amd.Batch(gpu_ctx, [fbp.slice_desc(s, info, ang)
                                  for s, (_, ang, info) in zip(specs, sims)])
    sinos = [x[0] for x in sims]
    images = [torch.full((s.nx, s.ny), 7.0, dtype=torch.float32, device="cuda") for s in specs]
    torch.cuda.synchronize()
    return specs, b, sinos, images, ref64, refmx


def _check(ctx, specs, images, ref64, refmx):
    import ctsim_amd
    for s, img, r64, rmx in zip(specs, images, ref64, refmx):
        d = ctsim_amd.image_stats(ctx, r64, img)[0]
        dm = ctsim_amd.image_stats(ctx, rmx, img)[0]
        print(f"  {s.geometry} {s.nDet}x{s.nView} {s.nx}^2 {s.method} {s.interp}: d={d:.2e} "
              f"vs single-slice mixed d={dm:.2e}", flush=True)
        assert d <= 1e-5
        assert dm <= 1e-4


def test_batch_matches_single_slice(gpu_ctx, batch_case):
    specs, b, sinos, images, ref64, refmx = batch_case
    print("case batch default shape", b.info(), flush=True)
    b.run(sinos, images)
    b.status()
    _check(gpu_ctx, specs, images, ref64, refmx)


# sym: slices whose views come in rotated copies (here all but the 181- and
# 97-view ones) run the rotationally symmetric body when P <= 2 (in two-phase
# mode with P = 4 those slices take P = 2 in their own launches), with each
# quadrant tile's views in sym groups (1, 2, 4) added after the launch
@pytest.mark.parametrize("P,lanesY,chunk,fused,sym",
                         [(2, 16, 32768, True, 1), (4, 8, 16384, True, 0),
                          (2, 32, 16384, False, 1), (4, 16, 32768, False, 0),
                          (1, 16, 8192, True, 1), (2, 16, 32768, True, 0),
                          (1, 8, 16384, False, 1), (2, 8, 16384, True, 4),
                          (1, 16, 16384, False, 2), (2, 8, 32768, False, 4),
                          (4, 8, 16384, False, 2),
                          # + 8: the other slices (here 181 and 97 views: one
                          # mirror, the rest four) run the mirror body
                          (4, 8, 16384, False, 8), (2, 16, 32768, False, 9),
                          (1, 8, 16384, False, 12), (4, 16, 32768, False, 10),
                          (2, 32, 8192, False, 8)])
def test_batch_shapes(gpu_ctx, batch_case, P, lanesY, chunk, fused, sym):
    import torch
    specs, b, sinos, images, ref64, refmx = batch_case
    print(f"case batch shape P={P} lanesY={lanesY} chunk={chunk} fused={fused} sym={sym}",
          flush=True)
    b.set_shape(P, lanesY, chunk, fused, sym)
    imgs = [torch.zeros_like(i) for i in images]
    b.run(sinos, imgs)
    b.status()
    _check(gpu_ctx, specs, imgs, ref64, refmx)
    b.set_shape(2, 16, 32768, True, False)
    base = [torch.zeros_like(i) for i in images]
    b.run(sinos, base)
    b.status()
    for a, c in zip(imgs, base):
        assert float((a - c).abs().max()) <= 2e-5 * float(c.abs().max())


def test_batch_graph_replays(gpu_ctx, batch_case):
    import torch
    specs, b, sinos, images, ref64, refmx = batch_case
    b.set_shape(2, 16, 32768, True, True)
    b.run(sinos, images)
    b.status()
    first = [i.clone() for i in images]
    for rep in range(3):
        for i in images:
            i.fill_(float(rep))  # other stream work between replays
        torch.cuda.synchronize()
        b.run(sinos, images)
        b.status()
        for a, f in zip(images, first):
            assert torch.equal(a, f)
    shape = b.autotune(sinos, images)
    print("  autotuned", shape, flush=True)
    b.run(sinos, images)
    b.status()
    _check(gpu_ctx, specs, images, ref64, refmx)


@pytest.mark.parametrize("P,fused,sym", [(2, False, 4), (1, False, 1), (2, True, 2)])
def test_batch_all_symmetric(gpu_ctx, P, fused, sym):
    """A batch whose slices all take the symmetric body (two-phase: no plain
    backprojection items at all, one launch per symmetric class) against the
    FP64 single-slice images."""
    import torch
    import ctsim_amd
    from ctsim_amd import fbp
    from ctsim_amd import workloads as W
    rows = [("equiangular", 512, 360, 256, "cubic", "fft", "abs_hanning"),
            ("parallel", 300, 200, 128, "linear", "fft", "abs_bandlimit"),
            ("equilinear", 367, 300, 256, "linear", "convolution", "shepp"),
            ("equiangular", 400, 240, 128, "cubic", "fft", "abs_hanning"),
            # a 4096-point FFT convolution (2 nDet - 1 > 2048)
            ("parallel", 1100, 200, 256, "linear", "convolution", "abs_bandlimit")]
    specs = [W.SliceSpec(elems=W.seeded_elements(20 + k, 0xC7511 + 950 + k), geometry=g, nDet=nd,
                         nView=nv, rotLen=np.pi if g == "parallel" else 2 * np.pi, filter=filt,
                         method=meth, interp=interp, backproj="idiff", filterParam=-1.0, nx=nx, ny=nx)
             for k, (g, nd, nv, nx, interp, meth, filt) in enumerate(rows)]
    print(f"case all-symmetric batch P={P} fused={fused} sym={sym}", flush=True)
    sims = [fbp.simulate(gpu_ctx, s) for s in specs]
    ref64 = [fbp.reconstruct(gpu_ctx, s, sino, ang, info, ctsim_amd.PREC_F64)
             for s, (sino, ang, info) in zip(specs, sims)]
    b = ctsim_amd.Batch(gpu_ctx, [fbp.slice_desc(s, info, ang) for s, (_, ang, info) in zip(specs, sims)])
    b.set_shape(P, 8, 16384, fused, sym)
    images = [torch.full((s.nx, s.ny), 7.0, dtype=torch.float32, device="cuda") for s in specs]
    for rep in range(2):
        b.run([x[0] for x in sims], images)
        b.status()
        for s, img, r64 in zip(specs, images, ref64):
            d = ctsim_amd.image_stats(gpu_ctx, r64, img)[0]
            print(f"  rep {rep} {s.geometry} {s.nDet}x{s.nView} {s.nx}^2: d={d:.2e}", flush=True)
            assert d <= 1e-5


@pytest.mark.parametrize("P,sym", [(1, 1), (4, 1), (4, 8), (2, 9)])
def test_batch_one_plain_class(gpu_ctx, P, sym):
    """A batch whose slices share one plain-body class (no rotated view copies:
    odd view counts) runs that class's own kernel in two-phase mode (with
    sym + 8 the mirror body: the views' only mirror is x -> -x, theta ->
    -theta); each slice against its FP64 single-slice image."""
    import torch
    import ctsim_amd
    from ctsim_amd import fbp
    from ctsim_amd import workloads as W
    specs = [W.SliceSpec(elems=W.seeded_elements(18 + k, 0xC7511 + 970 + k), geometry="equilinear",
                         nDet=300 + 40 * k, nView=181 + 20 * k, rotLen=2 * np.pi, filter="abs_hanning",
                         method="fft", interp="linear", backproj="idiff", nx=128 + 64 * k, ny=128 + 64 * k)
             for k in range(3)]
    print(f"case one-plain-class batch P={P} sym={sym}", flush=True)
    sims = [fbp.simulate(gpu_ctx, s) for s in specs]
    ref64 = [fbp.reconstruct(gpu_ctx, s, sino, ang, info, ctsim_amd.PREC_F64)
             for s, (sino, ang, info) in zip(specs, sims)]
    b = ctsim_amd.Batch(gpu_ctx, [fbp.slice_desc(s, info, ang) for s, (_, ang, info) in zip(specs, sims)])
    b.set_shape(P, 16, 16384, False, sym)
    images = [torch.full((s.nx, s.ny), 7.0, dtype=torch.float32, device="cuda") for s in specs]
    b.run([x[0] for x in sims], images)
    b.status()
    for s, img, r64 in zip(specs, images, ref64):
        d = ctsim_amd.image_stats(gpu_ctx, r64, img)[0]
        print(f"  {s.nDet}x{s.nView} {s.nx}^2: d={d:.2e}", flush=True)
        assert d <= 1e-5


def test_failed_layout_leaves_nothing_to_run(gpu_ctx, batch_case):
    """A shape whose layout fails (here a 1 KiB chunk budget, narrower than
    the view window of the 401-detector 128^2 cubic slice) leaves the batch
    with nothing to replay: run and status report the failed layout and no
    kernel touches the images; a supported shape then runs normally."""
    import torch
    import ctsim_amd
    specs, b, sinos, images, ref64, refmx = batch_case
    with pytest.raises(ctsim_amd.CtsimError, match="exceeds the chunk budget"):
        b.set_shape(2, 16, 1024, False, 0)
    imgs = [torch.full_like(i, 5.0) for i in images]
    with pytest.raises(ctsim_amd.CtsimError, match="last layout failed"):
        b.run(sinos, imgs)
    with pytest.raises(ctsim_amd.CtsimError, match="last layout failed"):
        b.status()
    torch.cuda.synchronize()
    assert all(float(i.min()) == 5.0 and float(i.max()) == 5.0 for i in imgs)
    b.set_shape(2, 16, 32768, True, 1)
    b.run(sinos, imgs)
    b.status()
    _check(gpu_ctx, specs, imgs, ref64, refmx)
