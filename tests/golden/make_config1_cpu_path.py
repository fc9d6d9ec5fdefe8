"""Writes tests/golden/config1_cpu_path.json: configuration 1 reconstructed
through the reference's CPU tile path (oracle/cpu_tile_torch.py, restating
gaussian/tile_voxelizer.py:210-244 with torch autograd) -- 64^3 phantom,
8 coils, R=4 variable-density mask, 20k Gaussians, 200 Adam steps at lr 1e-2.
The GPU test of configuration 1 compares the HIP path's final PSNR with this
run's (within 0.1 dB).  It also records a 2-step run with an evaluation after
every step, which a CPU test recomputes to pin this file to the code.  CPU
only (a few minutes):

    python tests/golden/make_config1_cpu_path.py
"""

from __future__ import annotations

import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import torch  # noqa: E402

from gs_tmi_amd.recon import Reconstructor  # noqa: E402
from gs_tmi_amd.recon.configs import config1_gaussians, config1_problem, config1_recon_config  # noqa: E402
from oracle.cpu_tile_torch import CpuTileVoxelizer  # noqa: E402

OUT = Path(__file__).resolve().parent / "config1_cpu_path.json"


def run(steps, eval_every):
    p = config1_problem("cpu")
    g = config1_gaussians(p)
    rec = Reconstructor(p, g, CpuTileVoxelizer(p.shape), config1_recon_config())
    t0 = time.perf_counter()
    hist = rec.run(steps, eval_every=eval_every,
                   log=lambda r: print(json.dumps(r), flush=True))
    rows = [{k: r[k] for k in ("step", "loss", "nmse", "psnr", "n")} for r in hist.rows]
    return rows, time.perf_counter() - t0


def main():
    torch.set_num_threads(os.cpu_count() or 1)
    short, _ = run(2, 1)
    full, wall = run(200, 25)
    OUT.write_text(json.dumps({
        "workload": "configuration 1: 64^3, 20 ellipsoids, 8 coils, R=4 variable density (16x16 centre), "
                    "noise 1e-3, 20k Gaussians, 200 Adam steps lr 1e-2, reference CPU tile path (fp32)",
        "made_by": "tests/golden/make_config1_cpu_path.py", "cpu_threads": torch.get_num_threads(),
        "wall_s": round(wall, 1), "pin_two_steps": short, "history": full}, indent=1) + "\n")
    print(f"wrote {OUT}: final {full[-1]}")


if __name__ == "__main__":
    main()
