"""Build libctsim_amd.so in-tree.

Host sources (*.cpp) are compiled by g++ with -O2 -ffp-contract=off, the
reference's own code-generation class, so the FP64 geometry, filter and trig
tables they produce are bit-identical to the reference's.  Device sources
(*.hip) are compiled by hipcc for gfx950 only, with packed-FP32 instructions
disabled.  Everything links into ctsim_amd/libctsim_amd.so; the command-line tools of
ctsim_amd/tools are linked against it into ctsim_amd/bin.

    python -m ctsim_amd.build [-j N] [--verbose]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libctsim_amd.so")
TOOLS = os.path.join(PKG, "tools")
BIN = os.path.join(PKG, "bin")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

INCLUDES = [f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}", f"-I{ROCM}/include"]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
              "-D__HIP_PLATFORM_AMD__"]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-munsafe-fp-atomics"]


def _sources():
    host = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    dev = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    return host, dev


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose and r.stderr.strip():
        print(r.stderr, file=sys.stderr)


def build(jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    host, dev = _sources()
    cmds, objs = [], []
    for f in host:
        o = os.path.join(OBJ, f[:-4] + ".o")
        cmds.append(["g++", *HOST_FLAGS, *INCLUDES, "-c", os.path.join(CSRC, f), "-o", o])
        objs.append(o)
    for f in dev:
        o = os.path.join(OBJ, f[:-4] + ".hip.o")
        cmds.append(["hipcc", *HIP_FLAGS, *INCLUDES, "-c", os.path.join(CSRC, f), "-o", o])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    tmp = LIB + ".tmp"
    _run(["hipcc", "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs,
          f"-L{ROCM}/lib", "-lamdhip64", "-Wl,-z,defs", "-lm", "-lpthread"], verbose)
    os.replace(tmp, LIB)
    _build_tools(verbose)
    return LIB


def _build_tools(verbose: bool) -> None:
    """Command-line drop-ins (pjrec, phm2pj, if2) linked to the library with
    an rpath of $ORIGIN/.., so they run from the tree."""
    os.makedirs(BIN, exist_ok=True)
    for f in sorted(os.listdir(TOOLS)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(BIN, f[:-4])
        _run(["g++", *HOST_FLAGS, *INCLUDES, os.path.join(TOOLS, f), "-o", exe, f"-L{PKG}",
              "-lctsim_amd", "-Wl,-rpath,$ORIGIN/..", f"-L{ROCM}/lib", "-lamdhip64",
              f"-Wl,-rpath,{ROCM}/lib"], verbose)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.verbose))
