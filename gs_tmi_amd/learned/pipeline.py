"""Train the initialiser and compare learned against plain initialisation.

    python -m gs_tmi_amd.learned.pipeline train    --phantoms 200 --weights W.safetensors
    python -m gs_tmi_amd.learned.pipeline evaluate --weights W.safetensors --out R.json
    python -m gs_tmi_amd.learned.pipeline all      --phantoms 200 --weights W --out R.json

Training acquisitions use seeds 1000 + i (evaluation acquisitions use seeds
below 1000, by default 0, so their phantoms are not among them).  Evaluation runs the config-4 workload
(256 x 256 x 128, 60 ellipsoids, 16 coils, Poisson-disc R=8 with a 24 x 24
calibration centre, noise 1e-3) from both initialisations with identical
optimiser settings under two schedules:

  config4  300k plain Gaussians, default learning rates, densify/prune every
           100 steps up to 1M, 2000 steps;
  tuned    600k plain Gaussians, 3x learning rates, no densification.

Wall time is initialisation time plus training time (evaluations excluded).
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import torch

TRAIN_SEED0 = 1000


def _log_to(path):
    f = open(path, "a") if path else None

    def log(row):
        line = json.dumps(row)
        print(line, file=sys.stderr, flush=True)
        if f:
            f.write(line + "\n")
            f.flush()
    return log


def schedules():
    from ..recon.configs import config4_recon_config

    base = config4_recon_config()
    tuned = config4_recon_config()
    tuned.lr_position *= 3
    tuned.lr_scale *= 3
    tuned.lr_rotation *= 3
    tuned.lr_density *= 3
    tuned.densify = False
    return {"config4": (300_000, base), "tuned": (600_000, tuned)}


def _time_to(rows, init_s, threshold):
    for r in rows:
        if r["nmse"] <= threshold:
            return init_s + r["train_s"]
    return None


def run_train(args, log):
    from .training import build_dataset, save_weights, train_network

    dev = torch.device("cuda:0")
    seeds = [TRAIN_SEED0 + i for i in range(args.phantoms)]
    t0 = time.perf_counter()
    data = build_dataset(seeds, dev, fit_steps=args.fit_steps, log=log, labels=args.labels)
    log({"dataset": len(data), "occupied": float(data.targets[..., 8].mean()), "build_s": time.perf_counter() - t0})
    model, hist = train_network(data, epochs=args.epochs, batch=args.batch, lr=args.lr,
                                val_phantoms=max(1, args.phantoms // 20), log=log)
    Path(args.weights).parent.mkdir(parents=True, exist_ok=True)
    save_weights(model, args.weights)
    log({"saved": str(args.weights)})
    return model


def run_evaluate(args, log, model=None):
    from ..recon.data import make_problem, poisson_disc_mask
    from ..recon.trainer import Reconstructor
    from ..tile_voxelizer import TileVoxelizer
    from .init import LearnedInitialiser, initialise

    dev = torch.device("cuda:0")
    shape = tuple(args.shape)
    vox = TileVoxelizer(shape, tile_size=8, device="cuda:0")
    learned_init = LearnedInitialiser(model=model, weights=None if model is not None else args.weights, device=dev,
                                      threshold=args.threshold)
    results = {"workload": {"shape": shape, "coils": 16, "accel": 8, "calib": 24 * shape[1] // 256},
               "threshold_nmse": args.target_nmse, "seeds": args.seeds, "stop_at_target": args.stop_at_target,
               "runs": {}}
    for seed in args.seeds:
        mask = poisson_disc_mask((shape[1], shape[2]), 8, 24 * shape[1] // 256, seed)
        problem = make_problem(shape, 60, 16, mask, seed, noise_sigma=1e-3, device=dev, phantom_on_device=True)
        for name, (n_plain, cfg) in schedules().items():
            if args.schedules and name not in args.schedules:
                continue
            for kind in ("learned", "plain"):
                g, info = initialise(problem, vox, learned_init if kind == "learned" else None, plain_points=n_plain)
                rec = Reconstructor(problem, g, vox, cfg)
                hist = rec.run(args.steps, eval_every=args.eval_every,
                               target_nmse=args.target_nmse if args.stop_at_target else None,
                               log=lambda r, k=kind, n=name, sd=seed: log({"seed": sd, "schedule": n, "init": k, **r}))
                run = {"init": info, "rows": hist.rows, "final": hist.rows[-1],
                       "time_to_target_s": _time_to(hist.rows, info["init_s"], args.target_nmse),
                       "train_time_to_target_s": hist.time_to("nmse", args.target_nmse)}
                tag = f"{name}/{kind}" if len(args.seeds) == 1 else f"seed{seed}/{name}/{kind}"
                results["runs"][tag] = run
                log({"seed": seed, "schedule": name, "init": kind, "used": info["used"], "init_s": info["init_s"],
                     "time_to_target_s": run["time_to_target_s"], "final": hist.rows[-1]})
                del rec, g
                torch.cuda.empty_cache()
            pre = "" if len(args.seeds) == 1 else f"seed{seed}/"
            a, b = results["runs"].get(f"{pre}{name}/learned"), results["runs"].get(f"{pre}{name}/plain")
            if a and b:
                ta, tb = a["time_to_target_s"], b["time_to_target_s"]
                results[f"{pre}{name}_ratio"] = (ta / tb) if (ta is not None and tb is not None) else None
                results[f"{pre}{name}_final_psnr_diff_db"] = a["final"]["psnr"] - b["final"]["psnr"]
        del problem
        torch.cuda.empty_cache()
    if args.out:
        Path(args.out).parent.mkdir(parents=True, exist_ok=True)
        with open(args.out, "w") as f:
            json.dump(results, f)
    summary = {k: v for k, v in results.items() if k != "runs"}
    summary.update({k: {"time_to_target_s": v["time_to_target_s"], "final": v["final"], "used": v["init"]["used"],
                        "init": {kk: vv for kk, vv in v["init"].items() if kk in ("learned", "plain")}}
                    for k, v in results["runs"].items()})
    print(json.dumps(summary), flush=True)
    return results


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("command", choices=["train", "evaluate", "all"])
    ap.add_argument("--phantoms", type=int, default=200)
    ap.add_argument("--fit-steps", type=int, default=500)
    ap.add_argument("--labels", choices=["plain", "grid"], default="grid",
                    help="slot targets from slot-grid fits (the shipped weights) or from the plain method's fits, "
                         "moment-matched into the slots")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--log", default=None)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--eval-every", type=int, default=25)
    ap.add_argument("--target-nmse", type=float, default=1e-2)
    ap.add_argument("--threshold", type=float, default=0.5)
    ap.add_argument("--shape", type=int, nargs=3, default=[256, 256, 128])
    ap.add_argument("--schedules", nargs="*", default=None)
    ap.add_argument("--seeds", type=int, nargs="+", default=[0], help="evaluation acquisitions (not training seeds)")
    ap.add_argument("--stop-at-target", action="store_true", help="stop each run at the target NMSE")
    args = ap.parse_args(argv)
    log = _log_to(args.log)
    if args.weights is None:
        from .init import DEFAULT_WEIGHTS

        args.weights = str(DEFAULT_WEIGHTS)
    model = None
    if args.command in ("train", "all"):
        model = run_train(args, log)
    if args.command in ("evaluate", "all"):
        run_evaluate(args, log, model)


if __name__ == "__main__":
    main()
