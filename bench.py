#!/usr/bin/env python3
"""FBP slices/s at 512^2 equiangular fan beam on one MI355X: mixed-precision
path versus the plain FP64 GPU path, with the mixed image's Herman distance d
against the FP64 image.

Workload (one slice, inputs resident in HBM before timing starts): 64 seeded
ellipses, equiangular fan beam, 1024 detectors x 1440 views over 2 pi,
focal-length ratio 2, Hanning-windowed ramp ("abs_hanning") by FFT with
zeropad 1 (2048-point transforms), cubic interpolation into 512 x 512.

One step = one full FBP of one slice: clear the image, filter all 1440 views,
backproject them, post-scale.  The sinogram is simulated on the GPU by the
library's own ray-sum kernel before timing.

Prints ONE JSON line on stdout; progress goes to stderr.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--cpu-seconds 15]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FBP slices/s at 512² fan-beam: mixed vs FP64 GPU path at d ≤1e-4 vs FP64"
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6    # MI355X published FP64 vector peak
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
# Algorithmic FP32 flops per pixel-view update of the mixed equiangular cubic
# backprojector with Kahan accumulation: cubic interpolation 28 + equiangular
# geometry 16 (3 special: divide, atan, reciprocal) + Kahan 3.
FLOPS_PER_PVU_EQUIANGULAR_CUBIC_KAHAN = 47


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "bp_mixed_equiangular_cubic_512_traffic.txt")


def measured_traffic():
    """HBM bytes per launch of the main backprojection kernel, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or
    None when the summary is absent."""
    try:
        with open(TRAFFIC_SUMMARY) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(spec, sino_h, ang_h, seconds):
    """The reference's own CPU path (compiled from source into oracle/_ref when
    present, else the C restatement) on the first views of the same sinogram."""
    import numpy as np
    import oracle as O
    kind = "reference" if O.have_ref() else "port"
    lib_kind = "ref" if kind == "reference" else "oracle"
    phm = O.Phantom(None, spec.elems)
    # Without FFTW the reference's "fft" method refuses to run
    # (procsignal.cpp:160-168); its Fourier path is the direct DFT "fourier".
    args = dict(geometry=spec.geometry, nDet=spec.nDet, nView=spec.nView, nSample=spec.nSample,
                rotLen=spec.rotLen, focalRatio=spec.focalRatio, centerRatio=spec.centerRatio,
                filter=spec.filter, method="fourier", zeropad=spec.zeropad, interp=spec.interp,
                backproj=spec.backproj, nx=spec.nx, ny=spec.ny, sino=sino_h, angles=ang_h)
    _, tf1, tb1 = O.fbp_sample(lib_kind, phm, nUse=1, **args)
    per_view = max(tf1 + tb1, 1e-6)
    n = int(max(1, min(spec.nView, seconds / per_view)))
    _, tf, tb = O.fbp_sample(lib_kind, phm, nUse=n, **args)
    per_view_f, per_view_b = tf / n, tb / n
    slices_per_s = 1.0 / ((per_view_f + per_view_b) * spec.nView)
    return {
        "value": slices_per_s,
        "unit": "slices/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"first {n} of {spec.nView} views of the same sinogram, filter "
                   f"'fourier' (direct DFT; the reference's fft needs FFTW) + equiangular "
                   f"cubic backprojection, single thread, scaled to one slice"),
        "filter_s_per_view": per_view_f,
        "backproject_s_per_view": per_view_b,
        "backproject_only_slices_per_s": 1.0 / (per_view_b * spec.nView),
        "cpu_model": cpu_model(),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after the benchmark, run three device copies of this many MiB "
                         "(a kernel of known traffic for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import ctsim_amd
    from ctsim_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    spec = W.config2()
    ctx = ctsim_amd.Context(0)
    dev = "cuda:0"
    log(f"[bench] {torch.cuda.get_device_name(0)}; workload config 2: {spec.geometry} "
        f"{spec.nDet} det x {spec.nView} views, {spec.nx}^2, {spec.filter}/{spec.method}, {spec.interp}")

    # --- inputs: simulated on the GPU (FP64 ray sums), resident in HBM
    phm = ctsim_amd.Phantom(None, spec.elems)
    sc = ctsim_amd.Scanner(phm, spec.geometry, spec.nDet, spec.nView, 0, spec.nSample, spec.rotLen,
                           spec.focalRatio, spec.centerRatio)
    info = sc.info
    sino = torch.empty((spec.nView, spec.nDet), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ang = sc.collect(ctx, sino, ctsim_amd.PREC_F64)
    torch.cuda.synchronize()
    t_raysum64 = time.perf_counter() - t0
    sino_m = torch.empty_like(sino)
    t0 = time.perf_counter()
    sc.collect(ctx, sino_m, ctsim_amd.PREC_MIXED)
    torch.cuda.synchronize()
    t_raysum_m = time.perf_counter() - t0
    n_re = spec.nView * spec.nDet * spec.nSample * len(spec.elems)
    log(f"[bench] ray sums: FP64 {t_raysum64 * 1e3:.1f} ms, closed-form {t_raysum_m * 1e3:.1f} ms "
        f"({n_re / 1e6:.1f} M ray-element pairs)")

    ps = ctsim_amd.ProcessSignal(ctx, spec.filter, spec.method, 1 / info["detInc"], info["detInc"],
                                 spec.nDet, spec.filterParam, "spatial", spec.generation,
                                 spec.zeropad, 1, spec.geometry, info["focalLength"],
                                 info["sourceDetectorLength"])
    bp = ctsim_amd.Backprojector(ctx, spec.geometry, spec.nDet, info["detInc"], spec.nView,
                                 info["viewDiameter"], info["focalLength"],
                                 info["sourceDetectorLength"], spec.backproj, spec.interp,
                                 spec.nx, spec.ny)
    f64 = torch.empty((spec.nView, spec.nDet), dtype=torch.float64, device=dev)
    f32 = torch.empty((spec.nView, spec.nDet), dtype=torch.float32, device=dev)
    img64 = torch.empty((spec.nx, spec.ny), dtype=torch.float32, device=dev)
    img32 = torch.empty((spec.nx, spec.ny), dtype=torch.float32, device=dev)

    def fbp(prec, filt, img):
        img.zero_()
        ps.filter(sino, filt, prec)
        bp.backproject(filt, ang, img, prec)
        bp.postprocess(img)

    # --- plain FP64 GPU path
    fbp(ctsim_amd.PREC_F64, f64, img64)
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        fbp(ctsim_amd.PREC_F64, f64, img64)
    torch.cuda.synchronize()
    t_f64 = (time.perf_counter() - t0) / args.fp64_steps
    k64_ms, k64_n = ctx.profile_read()
    ctx.profile(False)
    log(f"[bench] FP64 path: {t_f64 * 1e3:.2f} ms/slice (backprojection kernel {k64_ms / k64_n:.3f} ms)")

    # --- mixed path: launch shape from the autotuner (setup, untimed)
    ps.filter(sino, f32, ctsim_amd.PREC_MIXED)
    shape = bp.autotune(f32, ang)
    log(f"[bench] autotuned mixed shape: lanesY={shape[0]} P={shape[1]} split={shape[2]} "
        f"chunk={shape[3]} window={shape[4]} ({shape[5]:.3f} ms)")

    # --- mixed path: warmup, then exactly K timed steps
    for _ in range(args.warmup):
        fbp(ctsim_amd.PREC_MIXED, f32, img32)
    torch.cuda.synchronize()
    ctx.profile(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        fbp(ctsim_amd.PREC_MIXED, f32, img32)
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    kms, kn = ctx.profile_read()
    ctx.profile(False)
    bp_ms = kms / kn
    log(f"[bench] mixed path: {t_mixed * 1e3:.3f} ms/slice (backprojection kernel {bp_ms:.3f} ms)")

    d, r, e = ctsim_amd.image_stats(ctx, img64, img32)
    log(f"[bench] mixed vs FP64: d={d:.3e} r={r:.3e} e={e:.3e}")

    pvu = spec.nx * spec.ny * spec.nView
    flops = FLOPS_PER_PVU_EQUIANGULAR_CUBIC_KAHAN * pvu
    achieved = flops / (bp_ms * 1e-3) / 1e12
    hbm_bytes = 4 * spec.nView * spec.nDet + 2 * 4 * spec.nx * spec.ny
    traffic = measured_traffic()

    out = {
        "metric": METRIC,
        "value": 1.0 / t_mixed,
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (64 seeded ellipses, GPU-simulated sinogram)",
        "config": {"workload": "config 2: equiangular fan beam, 1024 det x 1440 views over 2pi, "
                               "abs_hanning by FFT (2048-point), cubic, 512x512",
                   "slices_per_step": 1, "precision": "mixed"},
        "fp64_gpu": {"slices_per_s": 1.0 / t_f64, "ms_per_slice": t_f64 * 1e3,
                     "backprojection_kernel_ms": k64_ms / k64_n,
                     "speedup_mixed_over_fp64": t_f64 / t_mixed},
        "accuracy_vs_fp64": {"d": d, "r": r, "e": e, "target_d": 1e-4},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic,
                     "kernel": "bp_mixed_kernel<equiangular, cubic>", "kernel_ms": bp_ms,
                     "flops_per_pvu": FLOPS_PER_PVU_EQUIANGULAR_CUBIC_KAHAN,
                     "pvu_per_launch": pvu,
                     "hbm_algorithmic_bytes": hbm_bytes,
                     "hbm_frac": hbm_bytes / (bp_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                     "pvu_per_s": pvu / (bp_ms * 1e-3),
                     "launch_shape": {"lanesY": shape[0], "P": shape[1], "split": shape[2],
                                      "chunk": shape[3], "window": shape[4]}},
        "raysum": {"fp64_ms": t_raysum64 * 1e3, "closed_form_ms": t_raysum_m * 1e3,
                   "ray_element_pairs": n_re},
    }
    if not args.no_cpu:
        log(f"[bench] CPU baseline (~{args.cpu_seconds:.0f} s of single-thread work) ...")
        sino_h = sino.cpu().numpy()
        out["cpu_baseline"] = cpu_baseline(spec, sino_h, ang, args.cpu_seconds)
        log(f"[bench] CPU: {out['cpu_baseline']['value']:.3e} slices/s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copies: 3 x {args.copy_calibration_mib} MiB read + write")


if __name__ == "__main__":
    main()
