"""Helical 180LI interpolation and half-scan feathering on the GPU
(Projections::Helical180LI / HalfScanFeather, libctsim/projections.cpp:
200-439).  Bars: equiangular bit-exact against the reference's golden
outputs; both geometries bit-exact against the oracle on seeded sinograms
(the equilinear rules are this build's derivation, pinned by the oracle's
static-object invariant test)."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _run(ctx, geom, data, rotInc, detInc, fan, sdl, iv, feather):
    import torch
    import ctsim_amd
    d = torch.from_numpy(np.ascontiguousarray(data, np.float32)).cuda()
    out, ang = ctsim_amd.helical180li(ctx, geom, d, rotInc, detInc, fan, sdl, iv)
    if feather:
        out = ctsim_amd.halfscan_feather(ctx, geom, out, rotInc, detInc, fan, sdl)
    torch.cuda.synchronize()
    return out.cpu().numpy(), ang


def test_helical_golden(gpu_ctx):
    g = load_golden("helical")
    for tag in "abc":
        nDet, nView, iv, feather = (int(x) for x in g[f"{tag}_args"])
        rotInc, detInc, fan, sdl = (float(x) for x in g[f"{tag}_info"])
        print(f"case golden {tag}", flush=True)
        out, ang = _run(gpu_ctx, 1, g[f"{tag}_in"], rotInc, detInc, fan, sdl, iv, bool(feather))
        assert np.array_equal(out.view(np.uint32), g[f"{tag}_out"].view(np.uint32))
        assert np.array_equal(ang, g[f"{tag}_angles"])


@pytest.mark.parametrize("geom", [1, 2])
@pytest.mark.parametrize("nDet,nView,turns,iv", [(888, 2700, 2.3, -1), (255, 1000, 1.7, 600),
                                                  (64, 420, 1.7, -1)])
def test_helical_against_oracle(gpu_ctx, oracle_lib, geom, nDet, nView, turns, iv):
    rng = np.random.default_rng(nDet + geom)
    data = rng.standard_normal((nView, nDet)).astype(np.float32)
    rotInc = 2 * np.pi * turns / nView
    fan, sdl = np.pi / 3, 4.0
    detInc = fan / nDet if geom == 1 else 2 * sdl * np.tan(fan / 2) / nDet
    for feather in (False, True):
        print(f"case geom={geom} nDet={nDet} nView={nView} iv={iv} feather={feather}", flush=True)
        ref, rang = oracle_lib.helical180li(geom, data, rotInc, detInc, fan, sdl, iv, feather)
        out, ang = _run(gpu_ctx, geom, data, rotInc, detInc, fan, sdl, iv, feather)
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(ang, rang)


def test_helical_rejects_short_scans(gpu_ctx):
    import ctsim_amd
    data = np.zeros((100, 32), np.float32)
    with pytest.raises(ctsim_amd.CtsimError):
        _run(gpu_ctx, 1, data, 2 * np.pi / 100, 0.01, np.pi / 3, 4.0, -1, False)
    with pytest.raises(ctsim_amd.CtsimError):
        _run(gpu_ctx, 0, data, 2 * np.pi / 100, 0.01, np.pi / 3, 4.0, -1, False)


def test_helical_alternating_scans(gpu_ctx, oracle_lib):
    """Two scans that differ only in detector spacing, alternated on one
    context (its cached detector angles are rebuilt whenever the scan
    changes): every call matches the oracle."""
    rng = np.random.default_rng(77)
    nDet, nView, turns, fan, sdl = 96, 600, 1.7, np.pi / 3, 4.0
    data = rng.standard_normal((nView, nDet)).astype(np.float32)
    rotInc = 2 * np.pi * turns / nView
    base = 2 * sdl * np.tan(fan / 2) / nDet
    for k, detInc in enumerate((base, 0.9 * base, base, 0.9 * base)):
        print(f"case alternating scan {k} detInc={detInc:.6g}", flush=True)
        ref, rang = oracle_lib.helical180li(2, data, rotInc, detInc, fan, sdl, -1, True)
        out, ang = _run(gpu_ctx, 2, data, rotInc, detInc, fan, sdl, -1, True)
        assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
        assert np.array_equal(ang, rang)
