"""Reconstruction harness on the GPU against the reference's CPU path (GPU tests)."""

from __future__ import annotations

import pytest
import torch

from gs_tmi_amd import recon
from oracle.cpu_tile_torch import CpuTileVoxelizer

pytestmark = pytest.mark.gpu


def _small_problem(device):
    shape = (32, 32, 32)
    mask = recon.variable_density_mask((32, 32), 4, 8, 0)
    return recon.make_problem(shape, 10, 4, mask, seed=5, noise_sigma=1e-3, device=device)


def test_gpu_reconstruction_matches_cpu_reference_path():
    """Same data, same initial Gaussians, 60 Adam steps: final PSNR of the HIP
    path within 0.1 dB of the reference's CPU tile path."""
    from gs_tmi_amd import TileVoxelizer

    print("[case] 32^3, 4 coils, R=4, 2000 Gaussians, 60 steps: GPU vs CPU path", flush=True)
    res = {}
    for dev in ("cuda:0", "cpu"):
        p = _small_problem(dev)
        g = recon.GaussianSet.from_image(p.zero_filled, 2000, seed=0)
        vox = TileVoxelizer(p.shape, device=dev) if dev != "cpu" else CpuTileVoxelizer(p.shape)
        r = recon.Reconstructor(p, g, vox, recon.ReconConfig.single_lr(1e-2, gamma=1.0, max_grad_norm=0.0))
        h = r.run(60, eval_every=20)
        res[dev] = h.rows
        print(f"  {dev}: " + ", ".join(f"step {x['step']}: psnr {x['psnr']:.3f} nmse {x['nmse']:.4f}" for x in h.rows))
    assert res["cuda:0"][0]["psnr"] == pytest.approx(res["cpu"][0]["psnr"], abs=1e-3)
    assert res["cuda:0"][-1]["psnr"] > res["cuda:0"][0]["psnr"] + 1.0
    assert abs(res["cuda:0"][-1]["psnr"] - res["cpu"][-1]["psnr"]) < 0.1


def test_gpu_densify_and_prune_keep_training():
    from gs_tmi_amd import TileVoxelizer

    print("[case] densify/prune on the GPU", flush=True)
    p = _small_problem("cuda:0")
    g = recon.GaussianSet.from_image(p.zero_filled, 1500, seed=1)
    cfg = recon.ReconConfig(densify=True, densify_from=10, densify_every=10, densify_until=40, grad_threshold=1e-4,
                            scale_threshold=0.0, opacity_threshold=1e-3, max_points=4000, clone=False,
                            max_grad_norm=0.0)
    r = recon.Reconstructor(p, g, TileVoxelizer(p.shape, device="cuda:0"), cfg)
    counts = []
    losses = []
    for _ in range(50):
        losses.append(float(r.step()))
        counts.append(r.g.num_points)
    print(f"  points {counts[0]} -> {max(counts)} -> {counts[-1]}, loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    assert max(counts) > 1500 and max(counts) <= 4000
    assert losses[-1] < losses[0]
    assert all(torch.isfinite(t).all() for t in r.g.param_list())
