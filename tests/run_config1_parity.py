"""Configuration 1 end to end, HIP path on the GPU against the reference's CPU
path on the host cores: same phantom, coils, mask, noise and initial Gaussians,
200 Adam steps at lr 1e-2.  Writes a JSON summary.  Run on a GPU machine:

    python tests/run_config1_parity.py gpurun_out/config1_parity.json
"""

from __future__ import annotations

import json
import os
import sys
import time
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))

import torch  # noqa: E402

from gs_tmi_amd import TileVoxelizer  # noqa: E402
from gs_tmi_amd.recon import Reconstructor  # noqa: E402
from gs_tmi_amd.recon.configs import config1_gaussians, config1_problem, config1_recon_config  # noqa: E402
from oracle.cpu_tile_torch import CpuTileVoxelizer  # noqa: E402


def main(out_path):
    torch.set_num_threads(os.cpu_count() or 1)
    results = {}
    for dev in ("cuda:0", "cpu"):
        print(f"[config1] {dev}", flush=True)
        p = config1_problem(dev)
        g = config1_gaussians(p)
        vox = TileVoxelizer(p.shape, device=dev) if dev != "cpu" else CpuTileVoxelizer(p.shape)
        rec = Reconstructor(p, g, vox, config1_recon_config())
        t0 = time.perf_counter()
        hist = rec.run(200, eval_every=25, log=lambda r: print(json.dumps(r), flush=True))
        results[dev] = {"wall_s": time.perf_counter() - t0, "history": hist.rows}
    g_final, c_final = results["cuda:0"]["history"][-1], results["cpu"]["history"][-1]
    summary = {"psnr_gpu": g_final["psnr"], "psnr_cpu": c_final["psnr"], "nmse_gpu": g_final["nmse"],
               "nmse_cpu": c_final["nmse"], "psnr_diff_db": g_final["psnr"] - c_final["psnr"],
               "train_s_gpu": g_final["train_s"], "train_s_cpu": c_final["train_s"],
               "cpu_threads": torch.get_num_threads()}
    results["summary"] = summary
    Path(out_path).parent.mkdir(parents=True, exist_ok=True)
    Path(out_path).write_text(json.dumps(results))
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "gpurun_out/config1_parity.json")
