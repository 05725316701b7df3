"""Device comparative statistics (ImageFile::comparativeStatistics,
libctsim/imagefile.cpp:185-240) against the reference's golden d, r, e and
the oracle.  Bars: e bit-identical (same float block sums); d and r within
max(1e-13, n * 2.3e-16) relative for n pixels: the device sums in a tree
order, the reference sequentially, whose rounding error bound grows as
n times the FP64 unit roundoff."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _stats(ctx, a, b):
    import torch
    import ctsim_amd
    ta = torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()
    tb = torch.from_numpy(np.ascontiguousarray(b, np.float32)).cuda()
    return ctsim_amd.image_stats(ctx, ta, tb)


def _close(got, want, n):
    d, r, e = got
    tol = max(1e-13, n * 2.3e-16)
    assert d == pytest.approx(want[0], rel=tol, abs=0)
    assert r == pytest.approx(want[1], rel=tol, abs=0)
    assert np.float64(e).view(np.uint64) == np.float64(want[2]).view(np.uint64)


def test_stats_golden(gpu_ctx):
    g = load_golden("stats")
    _close(_stats(gpu_ctx, g["a"], g["b"]), g["dre"], g["a"].size)


@pytest.mark.parametrize("nx,ny", [(512, 512), (1023, 257), (7, 5), (1, 9), (2048, 2048)])
def test_stats_against_oracle(gpu_ctx, oracle_lib, nx, ny):
    rng = np.random.default_rng(nx * 7 + ny)
    a = rng.standard_normal((nx, ny)).astype(np.float32)
    b = (a + 1e-3 * rng.standard_normal((nx, ny))).astype(np.float32)
    print(f"case {nx}x{ny}", flush=True)
    _close(_stats(gpu_ctx, a, b), oracle_lib.stats("oracle", a, b), nx * ny)


def test_stats_identical_images(gpu_ctx):
    a = np.random.default_rng(3).standard_normal((64, 48)).astype(np.float32)
    assert _stats(gpu_ctx, a, a) == (0.0, 0.0, 0.0)
