"""The drop-in header against the calls the reference's tools make.

tests/cxx/tools_api.cpp makes the calls of tools/pjrec.cpp (default names
from the static ID API, trace names, Projections::read + printScanInfo, the
image's history labels) and tools/phm2pj.cpp (geometry names,
createFromFile, Phantom::print, printScanInfo, setNView, Timer).  It is
compiled twice: against include/ctsim/ctsim.h and libctsim_amd.so, and
against the reference's own headers and the reference library built from
its sources (oracle/_ref).  Both must compile, and both must print the same
text.  Host-only: no call needs a GPU.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cxx", "tools_api.cpp")
REF_INCLUDE = "/root/reference/include"
# the -D values the reference's configure probes on LP64 Linux (oracle/Makefile; timer.h
# also needs the sys/time.h probe)
REF_DEFS = ["-DSIZEOF_SHORT=2", "-DSIZEOF_INT=4", "-DSIZEOF_LONG=8", "-DSIZEOF_FLOAT=4",
            "-DSIZEOF_DOUBLE=8", "-DHAVE_SSTREAM=1", "-DHAVE_VSNPRINTF=1",
            "-DHAVE_GETTIMEOFDAY=1", "-DHAVE_UNISTD_H=1", "-DHAVE_STRING_H=1",
            "-DHAVE_SYS_TIME_H=1"]

PHANTOM = """ellipse 0 0 0.69 0.92 0 1
ellipse 0 -0.0184 0.6624 0.874 0 -0.98
rectangle 0.25 0.1 0.2 0.05 30 0.5
triangle -0.3 -0.2 0.15 0.1 -45 0.25
sector 0.1 0.4 0.1 0.2 10 0.2
segment -0.1 0.5 0.12 0.3 200 -0.1
"""
# the second record lacks its last three fields; the scan resumes at the
# next record's type
MALFORMED = """ellipse 0 0 0.5 0.3 10 1
ellipse 0.1 0.2 0.3
rectangle 0 0 0.2 0.2 0 0.5
"""


def _build(out, reference):
    if reference:
        ref_lib = os.path.join(ROOT, "oracle", "_ref")
        cmd = ["g++", "-O1", "-std=c++17", "-w", "-DCTSIM_REFERENCE", *REF_DEFS,
               f"-I{REF_INCLUDE}", SRC, "-o", out, f"-L{ref_lib}", "-lctsimref",
               f"-Wl,-rpath,{ref_lib}"]
    else:
        cmd = ["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
               f"-I{ROOT}/include", "-I/opt/rocm/include", SRC, "-o", out,
               f"-L{ROOT}/ctsim_amd", "-lctsim_amd", f"-Wl,-rpath,{ROOT}/ctsim_amd",
               "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.check_call(cmd)


@pytest.fixture(scope="module")
def inputs(tmp_path_factory):
    import oracle as O
    if not os.path.exists(os.path.join(ROOT, "ctsim_amd", "libctsim_amd.so")):
        pytest.skip("libctsim_amd.so is not built")
    if not O.have_ref() or not os.path.isdir(REF_INCLUDE):
        pytest.skip("the reference sources and library are needed")
    d = tmp_path_factory.mktemp("tools")
    rng = np.random.default_rng(7)
    nView, nDet = 9, 40
    data = rng.standard_normal((nView, nDet)).astype(np.float32)
    ang = np.arange(nView) * (2 * np.pi / nView)
    O.ref_write_pj(str(d / "scan.pj"), "shepp-logan", "equiangular", nDet, nView, 2, 2 * np.pi,
                   2.0, 2.0, data, ang, 3.5, "phm2pj: NDet=40, Nview=9, equiangular")
    (d / "phm.txt").write_text(PHANTOM)
    (d / "bad.txt").write_text(MALFORMED)
    exes = {}
    for kind in ("ours", "ref"):
        exes[kind] = str(d / f"tools_api_{kind}")
        _build(exes[kind], kind == "ref")
    return d, exes


def _run(exe, d):
    r = subprocess.run([exe, str(d / "scan.pj"), str(d / "phm.txt"), str(d / "bad.txt")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.splitlines()


def test_tools_calls_compile_and_match_the_reference(inputs):
    d, exes = inputs
    ours = _run(exes["ours"], d)
    ref = _run(exes["ref"], d)
    assert len(ours) == len(ref)
    for a, b in zip(ours, ref):
        assert a == b
    # the short record is reported on standard output, as sys_error does
    assert any("Insufficient fields reading phantom file" in x for x in ours)


def test_malformed_phantom_file_keeps_reading(inputs):
    d, exes = inputs
    out = _run(exes["ours"], d)
    line = next(x for x in out if x.startswith("createFromFile(malformed)"))
    # false, and all three records became elements
    assert line.split()[1:] == ["0", "nPElem", "3"]


def test_pjrec_defaults_and_shepp_quirk(inputs):
    d, exes = inputs
    out = _run(exes["ours"], d)
    assert "pjrec defaults: abs_bandlimit convolution direct linear idiff" in out
    assert "shepp -> 4 (FILTER_ABS_SINC 4)" in out
