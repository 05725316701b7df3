"""ctsim_fbp_slice_static, the entry point that allocates nothing: called
through the C ABI, its image must equal ctsim_fbp_slice's at MIXED precision
bit for bit (same kernels, same launch shape, same inputs), on two calls in a
row, the second with a smaller input; device memory must not change across
the calls; a slice whose scratch exceeds the static array is rejected before
any launch.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(ctx, geometry, nDet, nView, nx, interp, rot, filt, method, seed):
    import torch
    import ctsim_amd
    from ctsim_amd import workloads as W
    phm = ctsim_amd.Phantom(None, W.seeded_elements(32, seed))
    sc = ctsim_amd.Scanner(phm, geometry, nDet, nView, 0, 1, rot, 2.0, 2.0)
    info = sc.info
    sino = torch.empty((nView, nDet), dtype=torch.float32, device="cuda")
    ang = sc.collect(ctx, sino, ctsim_amd.PREC_MIXED)
    ps = ctsim_amd.ProcessSignal(ctx, filt, method, 1 / info["detInc"], info["detInc"], nDet,
                                 -1.0, "spatial", "direct", 1, 1, geometry, info["focalLength"],
                                 info["sourceDetectorLength"])
    bp = ctsim_amd.Backprojector(ctx, geometry, nDet, info["detInc"], nView, info["viewDiameter"],
                                 info["focalLength"], info["sourceDetectorLength"], "idiff",
                                 interp, nx, nx)
    return ps, bp, sino, ang


def _static(ctx, ps, bp, sino, ang, img):
    import ctsim_amd
    a = np.ascontiguousarray(ang, np.float64)
    return ctsim_amd.lib.ctsim_fbp_slice_static(ctx.handle, ps.handle, bp.handle,
                                                C.c_void_p(sino.data_ptr()),
                                                a.ctypes.data_as(C.c_void_p), a.size,
                                                C.c_void_p(img.data_ptr()))


def test_static_entry_matches_fbp_slice(gpu_ctx):
    import torch
    import ctsim_amd
    cases = [  # a config-2 slice (tuned shape, view-group partials), then a smaller one
        ("equiangular", 1024, 1440, 512, "cubic", 2 * np.pi, "abs_hanning", "fft", True),
        ("equilinear", 383, 300, 192, "linear", 2 * np.pi, "abs_bandlimit", "convolution", False),
    ]
    plans, refs, outs = [], [], []
    for c in cases:
        print(f"case no-alloc {c[:8]}", flush=True)
        ps, bp, sino, ang = _plan(gpu_ctx, *c[:8], seed=0xA110C + c[1])
        if c[8]:
            bp.autotune(torch.zeros(sino.shape, dtype=torch.float32, device="cuda"), ang)
        ref = torch.empty((c[3], c[3]), dtype=torch.float32, device="cuda")
        bp.fbp(ps, sino, ang, ref, ctsim_amd.PREC_MIXED)
        plans.append((ps, bp, sino, ang))
        refs.append(ref)
        outs.append(torch.full((c[3], c[3]), 5.0, dtype=torch.float32, device="cuda"))
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for (ps, bp, sino, ang), out in zip(plans, outs):
        rc = _static(gpu_ctx, ps, bp, sino, ang, out)
        assert rc == 0, ctsim_amd.lib.ctsim_last_error().decode()
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, "the static entry point changed the free device memory"
    for c, ref, out in zip(cases, refs, outs):
        nd = int((out != ref).sum())
        print(f"  {c[0]} {c[1]}x{c[2]} {c[3]}^2: {nd} pixels differ from ctsim_fbp_slice")
        assert torch.equal(out, ref)


def test_static_entry_rejects_oversized_slice(gpu_ctx):
    import torch
    import ctsim_amd
    # 2048 det x 2880 views, linear: 47 MB of coefficients + 24 MB of rows
    nDet, nView, nx = 2048, 2880, 2048
    ps = ctsim_amd.ProcessSignal(gpu_ctx, "abs_bandlimit", "convolution", 1.0, 1.0, nDet)
    bp = ctsim_amd.Backprojector(gpu_ctx, "parallel", nDet, 1.0, nView, 2048 * 1.41421356,
                                 2.0, 4.0, "idiff", "linear", nx, nx)
    sino = torch.zeros((nView, nDet), dtype=torch.float32, device="cuda")
    img = torch.full((nx, nx), 3.0, dtype=torch.float32, device="cuda")
    ang = np.arange(nView) * (np.pi / nView)
    rc = _static(gpu_ctx, ps, bp, sino, ang, img)
    msg = ctsim_amd.lib.ctsim_last_error().decode()
    print(f"  rejected: {msg}")
    assert rc == -1  # CTSIM_ERR_INVALID
    assert "MiB of scratch" in msg
    torch.cuda.synchronize()
    assert float(img.min()) == 3.0 and float(img.max()) == 3.0  # nothing ran
