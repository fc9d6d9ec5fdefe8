"""Writes tests/golden/reference_fp32_errors.json: the reference's own fp32
error against FP64 on every parity case (see tests/reference_fp32.py), from
which the GPU parity tests derive their tolerances.  CPU only:

    python tests/golden/make_reference_fp32_errors.py            # every case
    python tests/golden/make_reference_fp32_errors.py --missing  # only cases the table lacks
"""

from __future__ import annotations

import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402
import torch  # noqa: E402

from gs_tmi_amd.tile_voxelizer import voxel_space_params  # noqa: E402
from tests import cases  # noqa: E402
from tests import reference_fp32 as rf  # noqa: E402

FUNCTION_BAND = 2e-5  # the Function-level tests' cutoff ambiguity band


def function_inputs(c):
    """fp32 centre and cov_inv of the torch prologue (as the Function-level tests form them)."""
    t = torch.from_numpy
    center, _, cov, _ = voxel_space_params(t(c["pos"]), t(c["scales"]), t(c["rot"]), c["shape"])
    return center.numpy(), cov.numpy()


def main():
    torch.set_num_threads(8)
    missing = "--missing" in sys.argv[1:]
    out = {"note": "relative L2 error against FP64 of the reference's fp32 CPU composition; "
                   "made by tests/golden/make_reference_fp32_errors.py", "function": {}, "call": {}, "param": {}}
    if missing:
        out = rf.load_table()
    for name in cases.FUNCTION_CASES:
        if name in out["function"]:
            continue
        t0 = time.time()
        c = cases.case_by_name(name)
        center, cov = function_inputs(c)
        g_r, g_i = cases.function_upstream(c["shape"])
        out["function"][name] = rf.function_errors(c, center, cov, g_r, g_i, FUNCTION_BAND)
        print(f"function {name}: {out['function'][name]} ({time.time() - t0:.1f} s)", flush=True)
    for name in cases.CALL_CASES:
        if name in out["call"]:
            continue
        t0 = time.time()
        c = cases.case_by_name(name)
        g_r, g_i = cases.call_upstream(c["shape"])
        out["call"][name] = rf.call_errors(c, g_r, g_i)
        print(f"call {name}: {out['call'][name]} ({time.time() - t0:.1f} s)", flush=True)
    for reading, seed in (("voxel", 0), ("literal", 1)):
        if reading in out["param"]:
            continue
        c = cases.adjoint_check_case(seed, reading)
        g_r, g_i = cases.call_upstream(c["shape"], seed=3)
        out["param"][reading] = rf.call_errors(c, g_r, g_i)
        print(f"param {reading}: {out['param'][reading]}", flush=True)
    rf.TABLE.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {rf.TABLE}")


if __name__ == "__main__":
    main()
