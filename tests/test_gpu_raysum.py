"""GPU ray sums (Scanner::collectProjections) against the reference's golden
sinograms, and the whole phm2pj -> pjrec chain of config 1.

Bars
  F64     bit-exact sinograms (the clip chain's atan2, cos and sin are
          correctly rounded device routines, which agree with the C library
          wherever the latter is correctly rounded)
  MIXED   closed-form FP64 chords: max |error| <= 4 ulp(max|sino|)
  end to end (config 1, F64): phantom -> sinogram -> filter -> IntDiff image
          equals the reference's image bit for bit
"""
import itertools

import numpy as np
import pytest

from conftest import load_golden, record_parity
from ctsim_amd import workloads as W

pytestmark = pytest.mark.gpu


def _sino(ctx, phantom_args, geom, nDet, nView, offset, nSample, rotLen, fr=2.0, cr=2.0,
          precision=0):
    import torch
    import ctsim_amd
    builtin, elems = phantom_args
    phm = ctsim_amd.Phantom(builtin, elems)
    sc = ctsim_amd.Scanner(phm, geom, nDet, nView, offset, nSample, rotLen, fr, cr)
    sino = torch.zeros((nView, nDet), dtype=torch.float32, device="cuda")
    ang = sc.collect(ctx, sino, precision)
    torch.cuda.synchronize()
    return sino.cpu().numpy(), ang, sc


def _check_f64(case, out, ref):
    assert record_parity(case, out, ref) == 0


@pytest.mark.parametrize("geom,ns", list(itertools.product(["parallel", "equilinear", "equiangular"],
                                                           [1, 3])))
def test_raysums_match_reference(gpu_ctx, geom, ns):
    import ctsim_amd
    print(f"case raysum {geom} nSample={ns}", flush=True)
    g = load_golden("raysums")
    elems = W.mixed_elements(23) + W.seeded_elements(14, 24)
    nDet = 128 if ns == 1 else 127
    rl = np.pi if geom == "parallel" else 2 * np.pi
    out, ang, sc = _sino(gpu_ctx, (None, elems), geom, nDet, 90, 3, ns, rl, 2.5, 1.7)
    ref = g[f"{geom}_{ns}_sino"]
    assert np.array_equal(ang.view(np.uint64), g[f"{geom}_{ns}_angles"].view(np.uint64))
    info = g[f"{geom}_{ns}_info"]
    for k, i in (("detInc", 0), ("detStart", 1), ("rotInc", 2), ("viewDiameter", 3),
                 ("focalLength", 4), ("sourceDetectorLength", 5), ("fanBeamAngle", 6)):
        assert sc.info[k] == info[i], k
    _check_f64(f"golden raysum {geom} nSample={ns}", out, ref)
    mixed, _, _ = _sino(gpu_ctx, (None, elems), geom, nDet, 90, 3, ns, rl, 2.5, 1.7,
                        ctsim_amd.PREC_MIXED)
    assert np.abs(mixed - ref).max() <= 4 * np.spacing(np.float32(np.abs(ref).max()))


def test_builtin_phantoms(gpu_ctx):
    g = load_golden("raysums")
    out, _, _ = _sino(gpu_ctx, ("unit-pulse", ()), "parallel", 33, 8, 0, 1, np.pi)
    assert np.array_equal(out, g["unit_pulse_sino"])
    out, _, _ = _sino(gpu_ctx, ("herman", ()), "equiangular", 95, 40, 0, 2, 2 * np.pi)
    _check_f64("golden raysum herman equiangular", out, g["herman_equiangular_sino"])


def test_view_ranges_and_storage_offsets(gpu_ctx):
    """collectProjections over a partial range stores rows where asked and
    equals the full run (the chunked-work API, scanner.cpp:296-300)."""
    import torch
    import ctsim_amd
    elems = W.seeded_elements(9, 5)
    full, _, sc = _sino(gpu_ctx, (None, elems), "equilinear", 64, 30, 0, 1, 2 * np.pi)
    sino = torch.zeros((30, 64), dtype=torch.float32, device="cuda")
    for s in range(0, 30, 8):
        sc.collect(gpu_ctx, sino, ctsim_amd.PREC_F64, iStartView=s, iNumViews=min(8, 30 - s))
    torch.cuda.synchronize()
    assert np.array_equal(sino.cpu().numpy(), full)


def test_config1_phm2pj_pjrec_end_to_end(gpu_ctx):
    import torch
    import ctsim_amd
    g = load_golden("config1_shepp_parallel")
    s = W.config1()
    sino, ang, sc = _sino(gpu_ctx, (s.builtin, s.elems), s.geometry, s.nDet, s.nView, 0,
                          s.nSample, s.rotLen)
    _check_f64("golden raysum config 1", sino, g["sino"])
    assert np.array_equal(ang.view(np.uint64), g["angles"].view(np.uint64))
    i = sc.info
    ps = ctsim_amd.ProcessSignal(gpu_ctx, s.filter, s.method, 1 / i["detInc"], i["detInc"], s.nDet,
                                 geometry=s.geometry, focalLength=i["focalLength"],
                                 sourceDetectorLength=i["sourceDetectorLength"])
    bp = ctsim_amd.Backprojector(gpu_ctx, s.geometry, s.nDet, i["detInc"], s.nView,
                                 i["viewDiameter"], i["focalLength"], i["sourceDetectorLength"],
                                 s.backproj, s.interp, s.nx, s.ny)
    d_sino = torch.from_numpy(sino).cuda()
    filt = torch.zeros(d_sino.shape, dtype=torch.float64, device="cuda")
    ps.filter(d_sino, filt, ctsim_amd.PREC_F64)
    img = torch.zeros((s.nx, s.ny), dtype=torch.float32, device="cuda")
    bp.backproject(filt, ang, img, ctsim_amd.PREC_F64)
    bp.postprocess(img)
    torch.cuda.synchronize()
    out = img.cpu().numpy()
    assert record_parity("golden config 1 phm2pj + pjrec image", out, g["image"]) == 0
