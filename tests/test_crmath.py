"""Host build of the FP64 kernels' correctly rounded cos, sin, atan, atan2
(ctsim_amd/csrc/crmath.h) against a 113-bit reference (libquadmath, with
x - k pi/2 formed exactly from a 172-bit pi/2) and the C library.

Bars
  - no result differs from the correctly rounded value;
  - every difference from the C library is one of the C library's own
    misroundings;
  - each stage's measured error stays below a quarter of the bound its
    rounding test uses (the test is only sound while the bound holds), and
    the argument reduction's error below a tenth of its bound.

The same source runs on the device (tests/test_gpu_crmath.py checks the
device build against mpmath).  Host-only.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cxx", "crmath_check.cpp")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    exe = str(tmp_path_factory.mktemp("crm") / "crmath_check")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-mfma", "-msse4.1",
                        f"-I{ROOT}/ctsim_amd/csrc", SRC, "-o", exe, "-lquadmath", "-lm"],
                       capture_output=True, text=True)
    if r.returncode != 0 and "quadmath" in r.stderr:
        pytest.skip("libquadmath is not available")
    assert r.returncode == 0, r.stderr
    return exe


def _fields(line):
    w = line.split()
    return {w[i]: w[i + 1] for i in range(1, len(w) - 1, 2)}


@pytest.mark.parametrize("fn", ["cos", "sin", "atan", "atan2"])
def test_correctly_rounded(checker, fn):
    out = subprocess.run([checker, fn, "400000", "7"], capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr)
    f = _fields(out.stdout)
    assert int(f["wrong"]) == 0
    assert int(f["undecided"]) == 0
    assert int(f["differ"]) == int(f["libc_wrong"])


@pytest.mark.parametrize("fn", ["cos", "sin", "atan"])
def test_error_bounds_hold(checker, fn):
    out = subprocess.run([checker, "bounds", fn, "300000", "9"], capture_output=True, text=True,
                         timeout=300)
    print(out.stdout)
    w = out.stdout.split()
    assert float(w[w.index("stage1") + 3]) < 0.25
    assert float(w[w.index("stage2") + 3]) < 0.25
    assert float(w[w.index("reduction") + 1]) < 0.1


def test_host_and_file_evaluation_agree(checker, tmp_path):
    """The eval mode (used to compare builds) returns the routines' values."""
    import numpy as np
    x = np.array([0.5, -2.0, 100.0, 1e-5], np.float64)
    x.tofile(tmp_path / "x.f64")
    subprocess.check_call([checker, "eval", "cos", str(tmp_path / "x.f64"), str(tmp_path / "c.f64")])
    c = np.fromfile(tmp_path / "c.f64", np.float64)
    assert np.allclose(c, np.cos(x), rtol=2e-16, atol=0)
