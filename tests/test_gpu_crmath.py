"""The FP64 kernels' device cos, sin, atan and atan2 (ctsim_amd/csrc/crmath.h)
against values computed by mpmath at 160 bits and rounded to double: every
result must be the correctly rounded one, bit for bit.

Arguments: the ranges the kernels use (theta - phi within a few turns, the
clip chain's directions, tan of fan angles), log-uniform magnitudes, points
near multiples of pi/2 and special values.  The C library's values (Python's
math module) are compared too and their misroundings only counted: they are
why device and host can differ in the last bit where the C library is not
correctly rounded (about 1 call in 1200 in glibc 2.35).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 12000


def _args(fn, rng):
    if fn in ("cos", "sin"):
        a = [rng.uniform(-10, 10, N // 4), rng.uniform(-4 * np.pi, 4 * np.pi, N // 4),
             np.ldexp(1 + rng.uniform(size=N // 4), rng.integers(-60, 20, N // 4)) *
             rng.choice([-1.0, 1.0], N // 4),
             np.nextafter(rng.integers(-1000, 1000, N // 4) * (np.pi / 2), 0) +
             (rng.uniform(size=N // 4) - 0.5) * np.ldexp(1.0, -rng.integers(0, 40, N // 4))]
        x = np.concatenate(a + [np.array([0.0, -0.0, 1e-300, -5e-324, np.pi / 4, 1e6, -1.5e6,
                                          np.pi / 2, np.pi, 3 * np.pi / 2])])
        return x, None
    if fn == "atan":
        a = [rng.uniform(-4, 4, N // 3),
             np.ldexp(1 + rng.uniform(size=N // 3), rng.integers(-60, 60, N // 3)) *
             rng.choice([-1.0, 1.0], N // 3),
             (1 + (rng.uniform(size=N // 3) - 0.5) * 1e-3) * rng.choice([-1.0, 1.0], N // 3)]
        x = np.concatenate(a + [np.array([0.0, -0.0, 1.0, -1.0, 1e300, -1e300, np.inf, -np.inf,
                                          2.0 ** 54, 2.0 ** 55, 1e-310])])
        return x, None
    y = rng.uniform(-3, 3, N) * np.where(rng.uniform(size=N) < 0.25, np.ldexp(1.0, -rng.integers(0, 60, N)), 1)
    x = rng.uniform(-3, 3, N) * np.where(rng.uniform(size=N) < 0.25, np.ldexp(1.0, -rng.integers(0, 60, N)), 1)
    sy = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 1.0, np.inf, -np.inf, np.inf, 2.0, 1e-300])
    sx = np.array([1.0, 1.0, -1.0, -1.0, 0.0, -0.0, -np.inf, np.inf, -np.inf, 1.0, -1e-300, 1.0])
    return np.concatenate([x, sx]), np.concatenate([y, sy])


def _reference(fn, x, y):
    import mpmath as mp
    mp.mp.prec = 160
    f = {"cos": mp.cos, "sin": mp.sin, "atan": mp.atan}.get(fn)
    out = np.empty_like(x)
    for i in range(x.size):
        if fn == "atan2":
            # the sign of atan2 is the sign of y, zeros included (mpmath has no
            # -0); infinite arguments give multiples of pi/4 (C99 F.9.1.4)
            xi, yi = float(x[i]), float(y[i])
            if math.isinf(xi) and math.isinf(yi):
                v = float(mp.pi / 4) if xi > 0 else float(3 * mp.pi / 4)
            elif math.isinf(xi):
                v = 0.0 if xi > 0 else float(mp.pi)
            elif math.isinf(yi):
                v = float(mp.pi / 2)
            else:
                v = float(mp.atan2(mp.mpf(yi), mp.mpf(xi)))
                if yi == 0 and math.copysign(1.0, xi) < 0:
                    v = float(mp.pi)  # atan2(+-0, -x) = +-pi, -0 included
            out[i] = math.copysign(abs(v), yi)
        elif fn in ("cos", "sin") and abs(x[i]) < 1e-300:
            out[i] = 1.0 if fn == "cos" else x[i]  # cos rounds to 1, sin to x
        else:
            v = float(f(mp.mpf(float(x[i]))))
            out[i] = math.copysign(0.0, x[i]) if v == 0 and fn != "cos" else v
    return out


@pytest.mark.parametrize("fn", ["cos", "sin", "atan", "atan2"])
def test_device_functions_are_correctly_rounded(gpu_ctx, fn):
    import torch
    import ctsim_amd
    print(f"case crmath {fn}", flush=True)
    rng = np.random.default_rng(20261203 + len(fn))
    x, y = _args(fn, rng)
    dx = torch.from_numpy(x).cuda()
    dy = torch.from_numpy(y).cuda() if y is not None else None
    out = ctsim_amd.crmath_eval(gpu_ctx, fn, dx, dy).cpu().numpy()
    ref = _reference(fn, x, y)
    bad = np.flatnonzero(out.view(np.uint64) != ref.view(np.uint64))
    for i in bad[:5]:
        print(f"  {fn}({x[i]!r}{'' if y is None else ', y=' + repr(y[i])}) = {out[i]!r}, "
              f"correctly rounded {ref[i]!r}")
    lm = {"cos": math.cos, "sin": math.sin, "atan": math.atan}.get(fn)
    libc = np.array([math.atan2(y[i], x[i]) if fn == "atan2" else lm(x[i]) for i in range(x.size)])
    libc_wrong = int((libc.view(np.uint64) != ref.view(np.uint64)).sum())
    print(f"  {fn}: {bad.size} of {x.size} device results not correctly rounded; "
          f"C library misrounds {libc_wrong}")
    assert bad.size == 0
