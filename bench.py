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
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FBP slices/s at 512² fan-beam: mixed vs FP64 GPU path at d ≤1e-4 vs FP64"
PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_FP64_TFLOPS = 78.6    # MI355X published FP64 vector peak
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
PEAK_BF16_TFLOPS = 2500.0   # MI355X published dense BF16 MFMA peak
# Algorithmic FP32 flops per pixel-view update of the mixed equiangular cubic
# backprojector with Kahan accumulation: cubic interpolation 28 + equiangular
# geometry 16 (3 special: divide, atan, reciprocal) + Kahan 3.
FLOPS_PER_PVU_EQUIANGULAR_CUBIC_KAHAN = 47


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "bp_sym_pipe_equiangular_cubic_512_traffic.txt")
# the kernel a mixed launch shape runs (its last field: 0 plain, 1 chunked
# symmetric, 2 pipelined symmetric)
KERNEL_BY_SYM = {0: "bp_mixed_kernel<equiangular, cubic>",
                 1: "bp_sym_kernel<equiangular, cubic, 4 quarter turns>",
                 2: "bp_sym_pipe_kernel<equiangular, cubic, 4 quarter turns>",
                 3: "bp_mirror_kernel<equiangular, cubic, mirror members>"}


def measured_traffic():
    """HBM bytes per launch of the pipelined symmetric backprojection kernel
    on this workload, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (copy-calibrated), and the launch shape that run pinned
    (lanesY,P,split,chunk,sym,KB); (None, None) when the summary is absent."""
    traffic, shape = None, None
    try:
        with open(TRAFFIC_SUMMARY) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    traffic = float(line.split(":", 1)[1])
                elif "--shape " in line and shape is None:
                    shape = line.split("--shape ", 1)[1].split()[0]
    except OSError:
        pass
    return traffic, shape


COUNTER_SUMMARY = os.path.join(ROOT, "profiles", "bp_sym_pipe_equiangular_cubic_512_counters.md")


def measured_issue():
    """Executed-instruction figures of the pipelined symmetric kernel on this
    workload from the committed rocprofv3 counter summary: VALU issue
    utilisation (wave VALU instructions x 4 cycles over all SIMD cycles) and
    lane instructions per pixel-view update; None where absent."""
    out = {"valu_issue_utilisation": None, "lane_instructions_per_pvu": None,
           "source": os.path.relpath(COUNTER_SUMMARY, ROOT)}
    try:
        with open(COUNTER_SUMMARY) as fh:
            for line in fh:
                cells = [c.strip().strip("*") for c in line.split("|")]
                if len(cells) < 3:
                    continue
                if cells[1].startswith("VALU issue utilisation"):
                    out["valu_issue_utilisation"] = float(cells[2])
                elif cells[1].startswith("lane instructions per pixel-view update"):
                    out["lane_instructions_per_pvu"] = float(cells[2])
    except (OSError, ValueError):
        pass
    return out


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
    present, else the C restatement) on the first views of the same sinogram:
    one thread (Reconstructor::reconstructView view by view), then all host
    cores as the reference's GUI reconstructs (src/backgroundsupr.cpp:111-152,
    src/threadrecon.cpp:175-260: a Reconstructor per thread over contiguous
    view blocks, the partial images summed in float)."""
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
    n = int(max(1, min(spec.nView, 0.5 * seconds / per_view)))
    _, tf, tb = O.fbp_sample(lib_kind, phm, nUse=n, **args)
    per_view_f, per_view_b = tf / n, tb / n
    one_thread = 1.0 / ((per_view_f + per_view_b) * spec.nView)
    out = {
        "value": one_thread,
        "unit": "slices/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"first {n} of {spec.nView} views of the same sinogram, single thread, "
                   f"scaled to one slice"),
        "filter_method": ("fourier: the reference's fft needs FFTW, absent here, so the CPU "
                          "path filters by its direct O(N^2) DFT of the same filter "
                          "(procsignal.cpp:160-168, 565-581); the filter then dominates the "
                          "CPU time, backproject_only_slices_per_s isolates the comparable stage"),
        "single_thread": {"slices_per_s": one_thread, "views": n,
                          "filter_s_per_view": per_view_f,
                          "backproject_s_per_view": per_view_b,
                          "backproject_only_slices_per_s": 1.0 / (per_view_b * spec.nView)},
        "cpu_model": cpu_model(),
    }
    if kind == "reference":
        threads = max(1, min(16, os.cpu_count() or 1))
        nt = int(max(threads, min(spec.nView, 0.5 * seconds * threads * 0.8 / per_view)))
        _, wall = O.fbp_threads(phm, nUse=nt, nThreads=threads,
                                **{k: v for k, v in args.items()})
        all_cores = nt / wall / spec.nView
        out.update(value=all_cores, cores=threads,
                   sample=(f"first {nt} of {spec.nView} views of the same sinogram on {threads} "
                           f"threads (contiguous view blocks, a Reconstructor per thread, float "
                           f"sum of the partial images), wall time scaled to one slice"),
                   all_cores={"slices_per_s": all_cores, "threads": threads, "views": nt,
                              "wall_s": wall})
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fp64-steps", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", default="config2",
                    choices=["config2", "config3", "config4", "config5"],
                    help="config2 (default): the headline single-slice measurement; the others "
                         "are the large-size, batched and helical configurations")
    ap.add_argument("--slices", type=int, default=256, help="config4: slices in the batch")
    ap.add_argument("--shape", default="",
                    help="config2: mixed launch shape lanesY,P,split,chunk[,sym[,stagingKB]] "
                         "instead of autotuning")
    ap.add_argument("--batch", type=int, default=8,
                    help="config2: also time this many copies of the slice as one batch (0: skip)")
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

    if args.workload != "config2":
        assert torch.cuda.is_available(), "bench.py needs a GPU"
        ctx = ctsim_amd.Context(0)
        fn = {"config3": bench_config3, "config4": bench_config4, "config5": bench_config5}
        out = fn[args.workload](ctx, args)
        print(json.dumps(out), flush=True)
        return

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
    sino_m = torch.empty_like(sino)
    # a first call of each precision uploads the elements; the second is timed
    ang = sc.collect(ctx, sino, ctsim_amd.PREC_F64)
    sc.collect(ctx, sino_m, ctsim_amd.PREC_MIXED)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ang = sc.collect(ctx, sino, ctsim_amd.PREC_F64)
    torch.cuda.synchronize()
    t_raysum64 = time.perf_counter() - t0
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
    t0 = time.perf_counter()
    for _ in range(args.fp64_steps):
        fbp(ctsim_amd.PREC_F64, f64, img64)
    torch.cuda.synchronize()
    t_f64 = (time.perf_counter() - t0) / args.fp64_steps
    # kernel duration from HIP events recorded around the backprojection
    # kernel on its stream, in one more step (event records are barrier
    # packets that add a few microseconds between kernels, so they stay out
    # of the timed steps)
    ctx.profile(True)
    fbp(ctsim_amd.PREC_F64, f64, img64)
    torch.cuda.synchronize()
    k64_ms, k64_n = ctx.profile_read()
    ctx.profile(False)
    log(f"[bench] FP64 path: {t_f64 * 1e3:.2f} ms/slice (backprojection kernel {k64_ms / k64_n:.3f} ms)")

    # --- mixed path: launch shape from the autotuner (setup, untimed), or
    # the one given with --shape (profiling runs pin the timed run's shape)
    ps.filter(sino, f32, ctsim_amd.PREC_MIXED)
    if args.shape:
        vals = [int(x) for x in args.shape.split(",")]
        lanesY, P, split, chunk = vals[:4]
        sym = vals[4] if len(vals) > 4 else 0
        kb = vals[5] if len(vals) > 5 else 0
        bp.set_shape(lanesY, P, split, chunk, sym, kb)
        shape = (lanesY, P, split, chunk, -1, sym, kb, float("nan"))
    else:
        shape = bp.autotune(f32, ang)
    log(f"[bench] autotuned mixed shape: lanesY={shape[0]} P={shape[1]} split={shape[2]} "
        f"chunk={shape[3]} window={shape[4]} sym={shape[5]} stagingKB={shape[6]} ({shape[7]:.3f} ms)")

    # --- mixed path as separate steps (zero, filter, backproject, scale)
    for _ in range(2):
        fbp(ctsim_amd.PREC_MIXED, f32, img32)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        fbp(ctsim_amd.PREC_MIXED, f32, img32)
    torch.cuda.synchronize()
    t_sep = (time.perf_counter() - t0) / 5
    log(f"[bench] mixed path as separate steps: {t_sep * 1e3:.3f} ms/slice")

    # --- mixed path, one call per slice (ctsim_fbp_slice: the filter writes
    # interval coefficients, the backprojection overwrites the scaled image):
    # warmup, then exactly K timed steps
    for _ in range(args.warmup):
        bp.fbp(ps, sino, ang, img32, ctsim_amd.PREC_MIXED)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        bp.fbp(ps, sino, ang, img32, ctsim_amd.PREC_MIXED)
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    # the same K steps again with HIP events around every backprojection
    # kernel (on the library's stream) for the kernel's mean duration
    ctx.profile(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        bp.fbp(ps, sino, ang, img32, ctsim_amd.PREC_MIXED)
    torch.cuda.synchronize()
    t_instr = (time.perf_counter() - t0) / args.steps
    kms, kn = ctx.profile_read()
    ctx.profile(False)
    bp_ms = kms / kn
    log(f"[bench] mixed path: {t_mixed * 1e3:.3f} ms/slice (instrumented replay {t_instr * 1e3:.3f} "
        f"ms/slice, backprojection kernel {bp_ms:.3f} ms)")

    d, r, e = ctsim_amd.image_stats(ctx, img64, img32)
    log(f"[bench] mixed vs FP64: d={d:.3e} r={r:.3e} e={e:.3e}")

    # --- the same slice as a batch of B copies: filtering, coefficients and
    # backprojection of all B in one persistent launch replayed from a graph
    batch = None
    if args.batch > 0:
        import dataclasses
        from ctsim_amd import fbp
        # B distinct slices: the workload's phantom first, then phantoms of the
        # same distribution from the following seeds (each simulated by the
        # FP64 ray sums, each with its own scanner geometry and plan)
        bspecs = [spec] + [dataclasses.replace(spec, elems=W.seeded_elements(64, W.config_seed(2) + 1000 + k))
                           for k in range(1, args.batch)]
        bsims = [(sino, ang, info)] + [fbp.simulate(ctx, s) for s in bspecs[1:]]
        b = ctsim_amd.Batch(ctx, [fbp.slice_desc(s, inf, a) for s, (_, a, inf) in zip(bspecs, bsims)])
        sinos = [x[0] for x in bsims]
        bimgs = [torch.empty((spec.nx, spec.ny), dtype=torch.float32, device=dev)
                 for _ in range(args.batch)]
        bshape = b.autotune(sinos, bimgs)
        for _ in range(args.warmup):
            b.run(sinos, bimgs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.run(sinos, bimgs)
        torch.cuda.synchronize()
        t_batch = (time.perf_counter() - t0) / args.steps
        b.status()
        dmax = 0.0
        for s_k, (sino_k, ang_k, info_k), img_k in zip(bspecs, bsims, bimgs):
            ref_k = fbp.reconstruct(ctx, s_k, sino_k, ang_k, info_k, ctsim_amd.PREC_F64)
            dmax = max(dmax, ctsim_amd.image_stats(ctx, ref_k, img_k)[0])
        log(f"[bench] batch of {args.batch} distinct slices: {t_batch * 1e3:.3f} ms "
            f"({args.batch / t_batch:.1f} slices/s), shape {bshape}, max d={dmax:.3e}")
        batch = {"slices": args.batch, "distinct_slices": True, "ms_per_batch": t_batch * 1e3,
                 "slices_per_s": args.batch / t_batch, "d_max": dmax,
                 "speedup_over_fp64": t_f64 * args.batch / t_batch,
                 "launch_shape": {"P": bshape[0], "lanesY": bshape[1], "chunk_bytes": bshape[2],
                                  "fused": bool(bshape[3]), "symmetric": bool(bshape[4])}}

    pvu = spec.nx * spec.ny * spec.nView
    flops = FLOPS_PER_PVU_EQUIANGULAR_CUBIC_KAHAN * pvu
    achieved = flops / (bp_ms * 1e-3) / 1e12
    # algorithmic HBM bytes of one slice: the float sinogram read once and the
    # float image written once
    hbm_bytes = 4 * spec.nView * spec.nDet + 4 * spec.nx * spec.ny
    traffic, traffic_shape = measured_traffic()

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
                   "slices_per_step": 1, "precision": "mixed", "entry": "ctsim_fbp_slice"},
        "mixed_separate_steps_ms": t_sep * 1e3,
        "fp64_gpu": {"slices_per_s": 1.0 / t_f64, "ms_per_slice": t_f64 * 1e3,
                     "backprojection_kernel_ms": k64_ms / k64_n,
                     "speedup_mixed_over_fp64": t_f64 / t_mixed},
        "accuracy_vs_fp64": {"d": d, "r": r, "e": e, "target_d": 1e-4},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic if shape[5] == 2 else None,
                     # the partial images of the view groups scale the writes:
                     # the profile's shape may differ from the autotuned one
                     "traffic_launch_shape": traffic_shape if shape[5] == 2 else None,
                     "kernel": KERNEL_BY_SYM[shape[5]], "kernel_ms": bp_ms,
                     "kernel_timing": "HIP events around each backprojection launch on the "
                                      "library's stream, over a replay of the K timed steps "
                                      f"({t_instr * 1e3:.3f} ms/step with the events)",
                     "flops_per_pvu": FLOPS_PER_PVU_EQUIANGULAR_CUBIC_KAHAN,
                     "flops_counting": "algorithmic: one position, weight and cubic interpolation "
                                       "per update; the symmetric kernels evaluate a position "
                                       "once per 4 updates and issue fewer instructions",
                     "executed": measured_issue() if shape[5] == 2 else None,
                     "pvu_per_launch": pvu,
                     "hbm_algorithmic_bytes": hbm_bytes,
                     "hbm_frac": hbm_bytes / (bp_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                     "pvu_per_s": pvu / (bp_ms * 1e-3),
                     "launch_shape": {"lanesY": shape[0], "P": shape[1], "split": shape[2],
                                      "chunk": shape[3], "window": shape[4],
                                      "symmetric": shape[5], "staging_kb": shape[6]}},
        "raysum": {"fp64_ms": t_raysum64 * 1e3, "closed_form_ms": t_raysum_m * 1e3,
                   "ray_element_pairs": n_re},
        "batched": batch,
    }
    if not args.no_cpu:
        log(f"[bench] CPU baseline (~{args.cpu_seconds:.0f} s: one thread, then all cores) ...")
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


def _line(metric, value, unit, args, ms, dtype, data, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": data,
           "config": config}
    out.update(extra)
    return out


def _time(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def bench_config3(ctx, args):
    """Config 3: parallel 2048 det x 2880 views, 'shepp' convolution, IntDiff
    linear into 2048^2: FP64 path vs mixed with the MFMA Toeplitz filter vs
    mixed with an FFT convolution (a one-slice batch)."""
    import torch
    import ctsim_amd
    from ctsim_amd import fbp, workloads as W
    spec = W.config3()
    log(f"[bench] config 3: {spec.geometry} {spec.nDet} det x {spec.nView} views, {spec.nx}^2, "
        f"{spec.filter}/{spec.method}, {len(spec.elems)} elements")
    sino, ang, info = fbp.simulate(ctx, spec, ctsim_amd.PREC_MIXED)
    img64 = torch.zeros((spec.nx, spec.ny), dtype=torch.float32, device="cuda")
    img32 = torch.zeros_like(img64)
    imgb = torch.zeros_like(img64)
    # plans (filter generation, geometry tables) are set-up, outside the timing
    ps, bp = fbp.plans(ctx, spec, info)
    t64 = _time(lambda: bp.fbp(ps, sino, ang, img64, ctsim_amd.PREC_F64), 1, 1)
    log(f"[bench] FP64 path: {t64 * 1e3:.1f} ms/slice")
    f32 = torch.empty((spec.nView, spec.nDet), dtype=torch.float32, device="cuda")
    t_toe = _time(lambda: ps.filter(sino, f32, ctsim_amd.PREC_MIXED), args.steps, args.warmup)
    shape = bp.autotune(f32, ang)
    log(f"[bench] autotuned mixed shape {shape}")
    t_mix = _time(lambda: bp.fbp(ps, sino, ang, img32, ctsim_amd.PREC_MIXED), args.steps, args.warmup)
    b = ctsim_amd.Batch(ctx, [fbp.slice_desc(spec, info, ang)])
    b.autotune([sino], [imgb])
    t_fft = _time(lambda: b.run([sino], [imgb]), args.steps, args.warmup)
    b.status()
    d_t = ctsim_amd.image_stats(ctx, img64, img32)
    d_f = ctsim_amd.image_stats(ctx, img64, imgb)
    log(f"[bench] mixed (Toeplitz MFMA filter {t_toe * 1e3:.3f} ms): {t_mix * 1e3:.2f} ms/slice, "
        f"d={d_t[0]:.2e}; FFT-convolution batch: {t_fft * 1e3:.2f} ms/slice, d={d_f[0]:.2e}")
    macs = spec.nView * spec.nDet * spec.nDet
    return _line("FBP slices/s at 2048^2 parallel (config 3): mixed path", 1 / t_mix, "slices/s",
                 args, t_mix * 1e3, "f32", "synthetic (256 ellipses + 32 rectangles)",
                 {"workload": "config 3: parallel 2048 det x 2880 views over pi, shepp by "
                              "convolution, IntDiff linear, 2048x2048"},
                 fp64_gpu={"slices_per_s": 1 / t64, "ms_per_slice": t64 * 1e3},
                 toeplitz_filter={"ms": t_toe * 1e3, "macs": macs,
                                  "bf16_products_per_mac": 6,
                                  "mfma_tflops": 12 * macs / t_toe / 1e12,
                                  "mfma_frac": 12 * macs / t_toe / 1e12 / PEAK_BF16_TFLOPS},
                 launch_shape=list(shape),
                 fft_convolution_batch={"ms_per_slice": t_fft * 1e3, "d": d_f[0]},
                 accuracy_vs_fp64={"d": d_t[0], "r": d_t[1], "e": d_t[2]})


def bench_config4(ctx, args):
    """Config 4: a batch of uneven slices in one persistent graph launch."""
    import torch
    import ctsim_amd
    from ctsim_amd import fbp, workloads as W
    specs = W.config4(args.slices)
    log(f"[bench] config 4: {len(specs)} slices")
    t0 = time.perf_counter()
    sims = [fbp.simulate(ctx, s, ctsim_amd.PREC_MIXED) for s in specs]
    torch.cuda.synchronize()
    log(f"[bench] sinograms simulated in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    b = ctsim_amd.Batch(ctx, [fbp.slice_desc(s, info, ang) for s, (_, ang, info) in zip(specs, sims)])
    log(f"[bench] batch plan in {time.perf_counter() - t0:.1f} s: {b.info()}")
    sinos = [x[0] for x in sims]
    images = [torch.zeros((s.nx, s.ny), dtype=torch.float32, device="cuda") for s in specs]
    shape = b.autotune(sinos, images)
    log(f"[bench] autotuned batch shape {shape}")
    t = _time(lambda: b.run(sinos, images), args.steps, args.warmup)
    b.status()
    pvu = b.info()["pvu"]
    # accuracy on a sample of slices against the FP64 single-slice path
    ds = []
    for k in range(0, len(specs), max(1, len(specs) // 4)):
        ref = fbp.reconstruct(ctx, specs[k], sims[k][0], sims[k][1], sims[k][2], ctsim_amd.PREC_F64)
        ds.append(ctsim_amd.image_stats(ctx, ref, images[k])[0])
    log(f"[bench] batch: {t * 1e3:.2f} ms, {len(specs) / t:.1f} slices/s, {pvu / t / 1e12:.3f} T PVU/s; "
        f"sampled d vs FP64: max {max(ds):.2e}")
    return _line("FBP slices/s, uneven batch of 256 slices (config 4), mixed", len(specs) / t,
                 "slices/s", args, t * 1e3, "f32", "synthetic (seeded phantoms per slice)",
                 {"workload": f"config 4: {len(specs)} seeded uneven slices, abs_bandlimit, linear",
                  "slices_per_step": len(specs)},
                 pvu_per_s=pvu / t, launch_shape=shape, accuracy_vs_fp64={"d_max_sampled": max(ds)})


def bench_config5(ctx, args):
    """Config 5: time-varying phantom scanned over several rotations, FP64
    180LI interpolation to 64 planar half-scan sinograms, batched FBP."""
    import numpy as np
    import torch
    import ctsim_amd
    from ctsim_amd import fbp, workloads as W
    h = W.config5()
    base = h.base()
    phm0 = ctsim_amd.Phantom(None, h.elements(0, base))
    # geometry from the view-0 phantom, as tools/phm2helix.cpp:287-296
    probe = ctsim_amd.Scanner(phm0, h.geometry, h.nDet, h.viewsPerRotation, 0, 1, 2 * np.pi,
                              h.focalRatio, h.centerRatio)
    fan, rotInc = probe.info["fanBeamAngle"], 2 * np.pi / h.viewsPerRotation
    L = int((np.pi + fan) / rotInc)
    nTotal = 2 * L + 1 + (h.positions - 1) * h.stride() + 2
    log(f"[bench] config 5: {nTotal} views ({nTotal / h.viewsPerRotation:.2f} rotations), "
        f"{h.positions} positions, L={L}")
    t0 = time.perf_counter()
    # the scan, one phantom per view, by the phm2helix drop-in (C++)
    with tempfile.TemporaryDirectory() as wd:
        sino, _, info = fbp.helical_scan(ctx, h, wd, nTotal, ctsim_amd.PREC_F64)
    torch.cuda.synchronize()
    log(f"[bench] time-varying scan simulated by phm2helix in {time.perf_counter() - t0:.1f} s")
    fan = info["fanBeamAngle"]
    spec = W.SliceSpec(geometry=h.geometry, nDet=h.nDet, nView=L + 1, filter=h.filter,
                       method=h.method, interp=h.interp, backproj=h.backproj, nx=h.nx, ny=h.ny)
    ivs = [L + 1 + p * h.stride() for p in range(h.positions)]

    def interpolate():
        outs = []
        for iv in ivs:
            o, a = ctsim_amd.helical180li(ctx, h.geometry, sino, info["rotInc"], info["detInc"],
                                          fan, info["sourceDetectorLength"], iv)
            outs.append((ctsim_amd.halfscan_feather(ctx, h.geometry, o, info["rotInc"],
                                                    info["detInc"], fan,
                                                    info["sourceDetectorLength"]), a))
        return outs

    planar = interpolate()
    b = ctsim_amd.Batch(ctx, [fbp.slice_desc(spec, info, a) for _, a in planar])
    sinos = [p for p, _ in planar]
    images = [torch.zeros((h.nx, h.ny), dtype=torch.float32, device="cuda") for _ in ivs]
    shape = b.autotune(sinos, images)
    t_i = _time(interpolate, max(1, args.steps // 4), 1)
    t_b = _time(lambda: b.run(sinos, images), args.steps, args.warmup)
    b.status()
    ref = fbp.reconstruct(ctx, spec, sinos[0], planar[0][1], info, ctsim_amd.PREC_F64)
    d = ctsim_amd.image_stats(ctx, ref, images[0])
    total = t_i + t_b
    log(f"[bench] helical interpolation {t_i * 1e3:.2f} ms, batched FBP {t_b * 1e3:.2f} ms "
        f"({h.positions / total:.1f} slices/s); d vs FP64 (position 0) {d[0]:.2e}")
    return _line("Helical slices/s (config 5): FP64 180LI + mixed batched FBP at 512^2",
                 h.positions / total, "slices/s", args, total * 1e3, "f32",
                 "synthetic (32 drifting ellipses, one phantom per view)",
                 {"workload": f"config 5: equilinear {h.nDet} det, {h.viewsPerRotation} views per "
                              f"rotation, {h.positions} positions, 512x512",
                  "slices_per_step": h.positions},
                 interpolation_ms=t_i * 1e3, batch_fbp_ms=t_b * 1e3, launch_shape=shape,
                 accuracy_vs_fp64={"d": d[0], "r": d[1], "e": d[2]})


if __name__ == "__main__":
    main()
