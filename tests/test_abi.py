"""The C ABI: every entry point include/*.h declares is exported by
libctsim_amd.so and bound by the Python loader (CPU only, no kernel runs)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ctsim_\w+)\s*\(", text, flags=re.M))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "ctsim_ctx_create" in names and "ctsim_backproject_views" in names


def test_library_exports_every_declared_symbol():
    from ctsim_amd._lib import LIB_PATH
    lib = C.CDLL(LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_loader_binds_every_declared_symbol():
    from ctsim_amd import _lib
    assert declared_symbols() == set(_lib.SIGNATURES)


def test_errors_are_reported_not_thrown():
    import ctsim_amd
    h = C.c_void_p()
    rc = ctsim_amd.lib.ctsim_ctx_create(C.byref(h), -5, None)
    assert rc != 0
    assert ctsim_amd.lib.ctsim_last_error()
    assert ctsim_amd.lib.ctsim_backproject_views(None, None, 0, None, None, 0, None) != 0
    assert b"NULL" in ctsim_amd.lib.ctsim_last_error()


def test_product_does_not_touch_the_oracle():
    pkg = os.path.join(ROOT, "ctsim_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text \
                    and "libctsimref" not in text, f
