"""The drop-in's .pj projection files and .if image files against the
reference's own writers and readers (Projections::write/read,
projections.cpp:486-546 and 817-889; Array2dFile::fileWrite/fileRead,
array2dfile.cpp:443-633), and the drop-in Scanner's geometry as it lands in
the .pj header.  Host-only: the driver makes no GPU call.

Bars: files byte-identical apart from the creation-date fields; each side
reads the other's file back exactly.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

PJ_DATE = slice(88, 100)  # six u16 date fields after 3 u32 and 9 f64 header fields


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    lib = os.path.join(ROOT, "ctsim_amd", "libctsim_amd.so")
    if not os.path.exists(lib):
        pytest.skip("libctsim_amd.so is not built")
    exe = str(tmp_path_factory.mktemp("fmt") / "file_formats")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
                           f"-I{ROOT}/include", "-I/opt/rocm/include",
                           os.path.join(ROOT, "tests", "cxx", "file_formats.cpp"), "-o", exe,
                           f"-L{ROOT}/ctsim_amd", "-lctsim_amd", f"-Wl,-rpath,{ROOT}/ctsim_amd",
                           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    return exe


@pytest.fixture(scope="module")
def ref():
    import oracle as O
    if not O.have_ref():
        pytest.skip("reference library not built")
    return O


def _run(exe, *args):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.parametrize("geometry,nDet,nView,rot", [("parallel", 37, 11, np.pi),
                                                    ("equiangular", 40, 9, 2 * np.pi),
                                                    ("equilinear", 33, 7, 2 * np.pi)])
def test_pj_matches_reference(driver, ref, tmp_path, geometry, nDet, nView, rot):
    rng = np.random.default_rng(nDet)
    data = rng.standard_normal((nView, nDet)).astype(np.float32)
    ang = np.cumsum(np.full(nView, rot / nView)) - rot / nView
    data.tofile(tmp_path / "d.f32")
    ang.tofile(tmp_path / "a.f64")
    remark = f"phm2pj: NDet={nDet}, {geometry}"
    ours, theirs = tmp_path / "ours.pj", tmp_path / "ref.pj"
    _run(driver, "pj-write", ours, "shepp-logan", geometry, nDet, nView, 2, rot, 2.0, 2.0,
         tmp_path / "d.f32", tmp_path / "a.f64", 1.25, remark)
    ref.ref_write_pj(str(theirs), "shepp-logan", geometry, nDet, nView, 2, rot, 2.0, 2.0, data,
                     ang, 1.25, remark)
    a, b = bytearray(ours.read_bytes()), bytearray(theirs.read_bytes())
    assert len(a) == len(b)
    a[PJ_DATE] = b[PJ_DATE] = bytes(12)
    assert a == b
    # the reference reads ours, we read the reference's
    info, d2, a2 = ref.ref_read_pj(str(ours), nView, nDet)
    assert (int(info[0]), int(info[1])) == (nView, nDet)
    assert np.array_equal(d2, data) and np.array_equal(a2, ang)
    out = _run(driver, "pj-read", theirs, tmp_path / "r.f32", tmp_path / "r.f64").splitlines()
    assert out[1] == remark
    hdr = out[0].split()
    assert (int(hdr[0]), int(hdr[1])) == (nView, nDet)
    assert [float(x) for x in hdr[3:11]] == [float(x) for x in info[3:11]]
    assert np.array_equal(np.fromfile(tmp_path / "r.f32", np.float32).reshape(nView, nDet), data)
    assert np.array_equal(np.fromfile(tmp_path / "r.f64", np.float64), ang)


def test_if_matches_reference(driver, ref, tmp_path):
    nx, ny = 13, 8
    img = np.random.default_rng(1).standard_normal((nx, ny)).astype(np.float32)
    img.tofile(tmp_path / "i.f32")
    inc, ext = (0.125, 0.25), (-0.8125, 0.8125, -1.0, 1.0)
    ours, theirs = tmp_path / "ours.if", tmp_path / "ref.if"
    _run(driver, "if-write", ours, nx, ny, tmp_path / "i.f32", *inc, *ext, "pjrec: 13x8", 0.5)
    ref.ref_write_if(str(theirs), img, inc, ext, "pjrec: 13x8", 0.5)
    a, b = bytearray(ours.read_bytes()), bytearray(theirs.read_bytes())
    assert len(a) == len(b) == 88 + 4 * nx * ny + 2 + 12 + 8 + 2 + len("pjrec: 13x8")
    date = slice(88 + 4 * nx * ny + 2, 88 + 4 * nx * ny + 14)
    a[date] = b[date] = bytes(12)
    assert a == b
    assert np.array_equal(ref.ref_read_if(str(ours), nx * ny), img)
    assert _run(driver, "if-read", theirs, tmp_path / "r.f32").split() == [str(nx), str(ny)]
    assert np.array_equal(np.fromfile(tmp_path / "r.f32", np.float32).reshape(nx, ny), img)
