"""The reconstruction workloads of the benchmark configurations.

config 1  64^3 phantom (20 ellipsoids, magnitudes U[0.2,1], smooth linear phase),
          8 Gaussian-profile coils on a sphere, 2D variable-density random
          undersampling R=4 over the two phase-encode axes with a fully sampled
          16x16 centre, k-space noise sigma 1e-3, 20k Gaussians at above-threshold
          voxels of the zero-filled image (isotropic 1 voxel), 200 Adam steps at
          lr 1e-2, fp32.
config 4  256x256x128 phantom (60 ellipsoids), 16 coils, R=8 Poisson-disc
          undersampling with a 24x24 calibration centre, 300k initial Gaussians,
          densify/prune every 100 steps up to 1M, 2000 Adam steps.
"""

from __future__ import annotations

import argparse
import json
import sys

import torch

from .data import make_problem, poisson_disc_mask, variable_density_mask
from .model import GaussianSet
from .trainer import ReconConfig, Reconstructor


def config1_problem(device, seed=0):
    shape = (64, 64, 64)
    mask = variable_density_mask((64, 64), 4, 16, seed)
    return make_problem(shape, 20, 8, mask, seed, noise_sigma=1e-3, device=device)


def config1_gaussians(problem, seed=0):
    return GaussianSet.from_image(problem.zero_filled, 20_000, mode="importance", scale_vox=1.0, seed=seed)


def config1_recon_config():
    return ReconConfig.single_lr(1e-2, gamma=1.0, max_grad_norm=0.0)


def config4_problem(device, seed=0, shape=(256, 256, 128), n_ellipsoids=60, n_coils=16):
    mask = poisson_disc_mask((shape[1], shape[2]), 8, 24, seed)
    return make_problem(shape, n_ellipsoids, n_coils, mask, seed, noise_sigma=1e-3, device=device)


def config4_recon_config(max_points=1_000_000):
    return ReconConfig(lr_position=1e-3, lr_scale=5e-3, lr_rotation=1e-3, lr_density=1e-2, gamma=0.999,
                       max_grad_norm=0.0, densify=True, densify_from=100, densify_every=100, densify_until=1500,
                       grad_threshold=2e-4, scale_threshold=0.0, opacity_threshold=1e-3, max_scale=0.5,
                       max_points=max_points, clone=False)


def run_plain(problem, gaussians, cfg, steps, eval_every, target_nmse=None, tile_size=8, log=None):
    from ..tile_voxelizer import TileVoxelizer

    vox = TileVoxelizer(problem.shape, tile_size=tile_size, device=str(problem.target.device))
    rec = Reconstructor(problem, gaussians, vox, cfg)
    hist = rec.run(steps, eval_every=eval_every, target_nmse=target_nmse, log=log)
    return rec, hist


def main(argv=None):
    ap = argparse.ArgumentParser(description="plain-initialised 3DGSMR reconstruction on the GPU")
    ap.add_argument("config", choices=["1", "4"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--eval-every", type=int, default=50)
    ap.add_argument("--target-nmse", type=float, default=None)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    log = lambda r: print(json.dumps(r), file=sys.stderr, flush=True)  # noqa: E731
    if args.config == "1":
        p = config1_problem(dev)
        g = config1_gaussians(p)
        _, hist = run_plain(p, g, config1_recon_config(), args.steps or 200, args.eval_every, args.target_nmse, log=log)
    else:
        p = config4_problem(dev)
        g = GaussianSet.from_image(p.zero_filled, 300_000, mode="importance", scale_vox=None, seed=0)
        _, hist = run_plain(p, g, config4_recon_config(), args.steps or 2000, args.eval_every, args.target_nmse,
                            log=log)
    out = {"config": args.config, "history": hist.rows, "time_to_nmse_1e-2": hist.time_to("nmse", 1e-2)}
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f)
    print(json.dumps({k: v for k, v in out.items() if k != "history"} | {"final": hist.rows[-1]}))


if __name__ == "__main__":
    main()
