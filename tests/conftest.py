import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")


def record_parity(case, out, ref):
    """Bitwise comparison of a GPU FP64 result with the reference's.

    Prints, and appends as one JSON line to $CTSIM_PARITY_LOG (default
    gpurun_out/parity_fp64.jsonl when that folder exists), the number of
    values whose bits differ and max|out - ref| / max|ref|; returns the count.
    """
    import json
    out = np.asarray(out)
    ref = np.asarray(ref)
    bits = {4: np.uint32, 8: np.uint64}[ref.dtype.itemsize]
    diff = int((out.view(bits) != ref.view(bits)).sum())
    scale = float(np.abs(ref).max()) or 1.0
    rel = float(np.abs(out.astype(np.float64) - ref.astype(np.float64)).max()) / scale
    print(f"  parity {case}: {diff} of {ref.size} values differ, max rel {rel:.3e}", flush=True)
    path = os.environ.get("CTSIM_PARITY_LOG")
    if not path and os.path.isdir(os.path.join(ROOT, "gpurun_out")):
        path = os.path.join(ROOT, "gpurun_out", "parity_fp64.jsonl")
    if path:
        with open(path, "a") as f:
            f.write(json.dumps({"case": case, "values": int(ref.size), "differ": diff,
                                "max_rel": rel}) + "\n")
    return diff


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle as O
    if not os.path.exists(O.ORACLE_SO):
        O.build(ref=False)
    return O


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import ctsim_amd
    return ctsim_amd.Context(0)
