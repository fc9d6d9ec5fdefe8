"""Benchmark of the HIP tile voxelizer on one MI355X.

One step = one full voxelizer call and its backward at the voxelizer
microbenchmark workload: TileVoxelizer.forward (fp32 prologue, GPU binning,
forward tile kernel) on a 256x256x128 volume with 1M Gaussians (means uniform
in the volume, log-normal voxel sigma with median 1.5 and sigma 0.5, random
rotations, N(0,1)+iN(0,1) amplitudes, 8^3 tiles), then autograd of
<G, vol> with a fixed random upstream gradient G down to positions,
log-scales, quaternions and complex amplitudes (G goes to autograd directly;
no loss reduction is computed).  Inputs are resident in HBM before the timed
region.

Prints one JSON line (rank 0).  The roofline entry is the dominant tile
kernel's FP32 rate on the work it executes: the reference kernel's FLOP count
per Gaussian-voxel evaluation (26 for every evaluation, +50 where q <= 9 in
the adjoint) times the evaluations the kernel performs on this input (exact
counts from the library), over its average duration, timed with HIP events
recorded by the library on the launch stream during the timed steps.  The
same FLOPs on the reference kernel's evaluation count (every voxel of every
binned tile) are reported beside it as the reference-basis rate; their ratio
is the work culling saves.  HBM traffic per launch comes from the committed
rocprofv3 counter summary of the same kernel source (null when the source
has changed since).  The CPU baseline is a PyTorch CPU restatement of the
reference's per-tile voxelizer, timed on a bounded sample of tiles and scaled
by the sample's share of the evaluations.

The second half of the metric, time to NMSE 1e-2 of the learned- against the
plain-initialised reconstruction, is the "learned_init" object: both runs of
the config-4 reconstruction (256x256x128, 16 coils, R=8) with the same Adam
settings, each stopped at the target.
"""

from __future__ import annotations

import argparse
import contextlib
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "voxelizer fwd+bwd ms (1M Gauss,256²×128); time-to-NMSE 1e-2 learned vs plain"
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector, published (packed FP32)
SHAPE = (256, 256, 128)
N_GAUSS = 1_000_000
TILE = 8

# Counter summary of the tile kernels at this workload (rocprofv3 FETCH_SIZE /
# WRITE_SIZE passes scaled by a copy kernel of known size, and the FP32 VALU
# instruction counters), keyed by the SHA-256 of the kernel source it was
# measured on.  Counters cannot be read from inside this process.
COUNTERS = ROOT / "profiles" / "tile_kernels_voxelizer_microbench_counters.json"
KERNEL_SOURCE = ROOT / "gs_tmi_amd" / "csrc" / "gstv.hip"

# Algorithmic FLOPs per evaluation, counted from the reference kernel
# (FMA = 2, mul/add/sub/div = 1, exp = 1): forward 26 ungated + 6 gated,
# backward 26 ungated + 50 gated.
FWD_UNGATED, FWD_GATED = 26, 6
BWD_UNGATED, BWD_GATED = 26, 50


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def counter_summary():
    """The committed counter summary when it was measured on the current
    kernel source, else None."""
    import hashlib

    try:
        summary = json.loads(COUNTERS.read_text())
    except (OSError, ValueError):
        return None
    digest = hashlib.sha256(KERNEL_SOURCE.read_bytes()).hexdigest()
    return summary if summary.get("gstv_hip_sha256") == digest else None


def make_inputs(dev, seed=0):
    import numpy as np
    import torch

    from gs_tmi_amd.synthetic import make_gaussians

    pos, scales, rot, rr, ri = make_gaussians(N_GAUSS, SHAPE, seed, sigma_median=1.5, sigma_log_sd=0.5)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    params = dict(pos=t(pos), log_s=t(np.log(scales)), rot=t(rot), dr=t(rr), di=t(ri))
    for v in params.values():
        v.requires_grad_()
    g = torch.Generator(device="cpu").manual_seed(seed + 1)
    grad_r = torch.randn(SHAPE, generator=g).to(dev)
    grad_i = torch.randn(SHAPE, generator=g).to(dev)
    return params, grad_r, grad_i


def cpu_baseline(params, budget_s=15.0):
    """Times the PyTorch CPU restatement of the reference's per-tile forward +
    backward on a sample of tiles; returns (extrapolated ms, sample note, cores)."""
    import numpy as np
    import torch

    from oracle import cpu_tile_torch as cpu
    from oracle import voxelizer_oracle as vo

    cores = os.cpu_count() or 1
    torch.set_num_threads(cores)
    pos = params["pos"].detach().cpu().numpy()
    sc = np.exp(params["log_s"].detach().cpu().numpy())
    rot = params["rot"].detach().cpu().numpy()
    c32, r32, _, hs = vo.prologue(pos, sc, rot, SHAPE)
    cov = vo.fp32_cov_inv(sc, rot)
    ranges, pids = vo.bin_tiles(c32, r32, SHAPE, TILE)
    pairs = (ranges[:, 1] - ranges[:, 0]).astype(np.int64)
    e_per_tile = pairs * TILE ** 3  # all tiles are full at this shape
    tc = torch.from_numpy(c32)
    tm = torch.from_numpy(cov)
    trr = params["dr"].detach().cpu()
    tri = params["di"].detach().cpu()
    ths = torch.from_numpy(hs)
    tr = torch.from_numpy(ranges)
    tp = torch.from_numpy(pids)
    g = torch.complex(torch.randn(SHAPE), torch.randn(SHAPE))
    rng = np.random.default_rng(0)
    order = rng.permutation(ranges.shape[0])
    # Pilot on 4 tiles to size the sample.
    t0 = time.perf_counter()
    cpu.tile_forward_backward(tc, tm, trr, tri, ths, tr, tp, g, SHAPE, TILE, order[:4])
    pilot = max(time.perf_counter() - t0, 1e-3)
    ntiles = int(max(8, min(ranges.shape[0], 4 * budget_s / pilot)))
    sample = order[4:4 + ntiles]
    t0 = time.perf_counter()
    cpu.tile_forward_backward(tc, tm, trr, tri, ths, tr, tp, g, SHAPE, TILE, sample)
    dt = time.perf_counter() - t0
    share = e_per_tile[sample].sum() / e_per_tile.sum()
    ms = dt / share * 1e3
    note = (f"per-tile fp32 torch forward+autograd backward on {len(sample)} random tiles of {ranges.shape[0]} "
            f"({share * 100:.2f}% of evaluations, {dt:.1f} s), extrapolated by evaluation share; "
            f"{cores} threads; CPU {cpu_model()}")
    return ms, note, cores


def _nmse_at(rows, init_s, wall_s):
    """NMSE of the last evaluation at or before wall time `wall_s` (init + training)."""
    best = None
    for r in rows:
        if init_s + r["train_s"] <= wall_s:
            best = r["nmse"]
    return best


def learned_init_leg(dev, max_steps=1500, eval_every=25, target=1e-2, config4_steps=2000):
    """Time to NMSE `target` of the config-4 reconstruction (256x256x128, 60
    ellipsoids, 16 coils, Poisson-disc R=8 with a 24x24 calibration centre,
    noise 1e-3) from the learned and from the plain initialisation with the
    same Adam settings, under two schedules:

    tuned_schedule    3x default learning rates, no densification, plain =
                      600k Gaussians at above-threshold voxels; each run
                      stops at the target;
    config4_schedule  configuration 4 as specified: default learning rates,
                      300k plain Gaussians, densify / prune every 100 steps up
                      to 1M, `config4_steps` steps (the plain start does not
                      reach the target there, so the plain NMSE at the learned
                      start's time to target is reported too).

    Wall time = initialisation + training, evaluations excluded.  Every
    initialisation runs once untimed first, so that one-time costs (FFT
    plans, first kernel launches, allocator growth) fall on no timed leg."""
    import torch

    from gs_tmi_amd import TileVoxelizer
    from gs_tmi_amd.learned.init import LearnedInitialiser, initialise
    from gs_tmi_amd.learned.pipeline import schedules
    from gs_tmi_amd.recon.data import make_problem, poisson_disc_mask
    from gs_tmi_amd.recon.trainer import Reconstructor

    shape = (256, 256, 128)
    problem = make_problem(shape, 60, 16, poisson_disc_mask((256, 128), 8, 24, 0), 0, noise_sigma=1e-3, device=dev,
                           phantom_on_device=True)
    with contextlib.redirect_stdout(sys.stderr):
        vox = TileVoxelizer(shape, tile_size=8, device=str(dev))
    learned = LearnedInitialiser(device=dev)
    out = {"workload": "config-4 reconstruction 256x256x128, 16 coils, R=8 Poisson-disc, noise 1e-3, seed 0",
           "target_nmse": target}
    for sched, steps, stop in (("tuned", max_steps, True), ("config4", config4_steps, False)):
        n_plain, cfg = schedules()[sched]
        for init in (learned, None):
            g, _ = initialise(problem, vox, init, plain_points=n_plain)
            del g
        torch.cuda.synchronize()
        leg = {"schedule": {"tuned": "Adam, 3x default learning rates, no densification, plain 600k",
                            "config4": "configuration 4 as specified: default learning rates, plain 300k, "
                                       "densify/prune every 100 steps up to 1M"}[sched],
               "steps": steps}
        for kind, init in (("learned", learned), ("plain", None)):
            log(f"[bench] learned-init leg ({sched}): {kind}")
            g, info = initialise(problem, vox, init, plain_points=n_plain)
            hist = Reconstructor(problem, g, vox, cfg).run(steps, eval_every=eval_every,
                                                          target_nmse=target if stop else None)
            t = hist.time_to("nmse", target)
            last = hist.rows[-1]
            leg[kind] = {"used": info["used"], "gaussians": last["n"], "init_s": round(info["init_s"], 3),
                         "time_to_target_s": None if t is None else round(info["init_s"] + t, 3),
                         "steps": last["step"], "nmse": last["nmse"], "psnr": last["psnr"],
                         "init_nmse": hist.rows[0]["nmse"], "_rows": hist.rows, "_init_s": info["init_s"]}
            del g, hist
            torch.cuda.empty_cache()
        a, b = leg["learned"]["time_to_target_s"], leg["plain"]["time_to_target_s"]
        leg["ratio"] = None if (a is None or b is None) else round(a / b, 4)
        if a is not None and b is None:
            # The plain start did not reach the target within its steps, so its
            # time to target exceeds its whole wall time: the ratio is below
            # learned time to target / plain wall time.
            plain_wall = leg["plain"]["_init_s"] + leg["plain"]["_rows"][-1]["train_s"]
            leg["plain"]["wall_s"] = round(plain_wall, 3)
            leg["ratio_upper_bound"] = round(a / plain_wall, 4)
        if a is not None:
            leg["plain_nmse_at_learned_time_to_target"] = _nmse_at(leg["plain"]["_rows"], leg["plain"]["_init_s"], a)
        for kind in ("learned", "plain"):
            del leg[kind]["_rows"], leg[kind]["_init_s"]
        out[f"{sched}_schedule"] = leg
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain time-to-NMSE leg")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="after timing, copy a buffer of this many MiB once (a known-size kernel that "
                         "calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested); the multi-GPU design is "
              "described, not run", file=sys.stderr)
        sys.exit(2)

    import torch

    from gs_tmi_amd import TileVoxelizer, _native, build_tile_pairs, voxel_space_params

    dev = torch.device("cuda:0")
    log("[bench] building inputs")
    params, grad_r, grad_i = make_inputs(dev)
    with contextlib.redirect_stdout(sys.stderr):  # keep stdout to the one JSON line
        vox = TileVoxelizer(SHAPE, tile_size=TILE, device=str(dev))

    def step():
        for p in params.values():
            p.grad = None
        vr, vi = vox(params["pos"], torch.exp(params["log_s"]), params["rot"],
                     torch.complex(params["dr"], params["di"]))
        # Upstream gradient G of the loss <G, vol>, handed straight to autograd.
        torch.autograd.backward((vr, vi), (grad_r, grad_i))

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # Exact evaluation counts of this input (outside the timed region).
    with torch.no_grad():
        center, radius, cov, hs = voxel_space_params(params["pos"], torch.exp(params["log_s"]), params["rot"],
                                                     SHAPE)
        ranges, pids = build_tile_pairs(center, radius, SHAPE, TILE)
        e_total, e_active, e_culled_ref, e_adjoint_ref = _native.count_evaluations(
            center, cov.contiguous(), params["dr"].detach(), params["di"].detach(), ranges, pids, hs, TILE, *SHAPE)
        pairs = int(pids.numel())
        # Binning alone, for the breakdown.
        bt = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            build_tile_pairs(center, radius, SHAPE, TILE)
            e1.record()
            torch.cuda.synchronize()
            bt.append(e0.elapsed_time(e1))
        # The fused call bins only the pairs whose tile meets a Gaussian's
        # extent box (records first, then the two binning phases), into the
        # capacity its voxelizer carries from earlier calls: the total stays
        # in device memory.
        dv = torch.stack([params["dr"].detach(), params["di"].detach()], 1).contiguous()
        kept = _native.keep_prepare(center, cov.contiguous(), dv[:, 0], dv[:, 1], hs, radius, SHAPE, TILE)
        pairs_kept = int(kept.total.item())
        cap = _native.PairCapacity()
        cap.update(N_GAUSS, pairs_kept)
        kbt = []
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            kept = _native.keep_prepare(center, cov.contiguous(), dv[:, 0], dv[:, 1], hs, radius, SHAPE, TILE)
            k_ranges, k_pids = _native.bin_fill(center, radius, SHAPE, TILE, kept.offsets, cap.capacity(N_GAUSS),
                                                kept=kept, capped=True)
            e1.record()
            torch.cuda.synchronize()
            kbt.append(e0.elapsed_time(e1))
        # The evaluations the tile kernels perform on the fused call's lists
        # (the reference's lists hold pairs that add nothing; E_active is the
        # same on both).
        _, e_active_kept, e_culled, e_adjoint = _native.count_evaluations(
            center, cov.contiguous(), params["dr"].detach(), params["di"].detach(), k_ranges, k_pids, hs, TILE,
            *SHAPE)
        if e_active_kept != e_active:
            log(f"[bench] active evaluations differ between the lists: {e_active_kept} vs {e_active}")
        del kept, dv, k_ranges, k_pids
    log(f"[bench] pairs {pairs}, E_total {e_total:.4e}, E_active {e_active:.4e}, E_culled {e_culled:.4e}, "
        f"E_adjoint {e_adjoint:.4e}")

    log(f"[bench] timing {args.steps} steps")
    torch.cuda.synchronize()
    _native.profile_begin()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    prof = _native.profile_end()
    ms_per_step = t_total * 1e3 / args.steps

    # Per-step spread (each step synchronised on its own, outside the timed loop).
    per_step = []
    for _ in range(min(args.steps, 10)):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        per_step.append((time.perf_counter() - t0) * 1e3)

    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    fwd_ms = prof["forward"][0] / max(prof["forward"][1], 1)
    bwd_ms = prof["backward"][0] / max(prof["backward"][1], 1)
    # FLOPs on the evaluations each kernel executes, and on the reference's.
    fwd_flop = FWD_UNGATED * e_culled + FWD_GATED * e_active
    bwd_flop = BWD_UNGATED * e_adjoint + BWD_GATED * e_active
    fwd_flop_ref = FWD_UNGATED * e_total + FWD_GATED * e_active
    bwd_flop_ref = BWD_UNGATED * e_total + BWD_GATED * e_active
    dominant = "backward" if bwd_ms >= fwd_ms else "forward"
    k_flop, k_flop_ref, k_ms, k_evals = ((bwd_flop, bwd_flop_ref, bwd_ms, e_adjoint) if dominant == "backward"
                                         else (fwd_flop, fwd_flop_ref, fwd_ms, e_culled))
    achieved = k_flop / (k_ms * 1e-3) / 1e12
    counters = counter_summary()
    kc = (counters or {}).get("kernels", {}).get(dominant)
    traffic = None if kc is None else int(kc["fetch_bytes"] + kc["write_bytes"])

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline")
        cpu_ms, note, cores = cpu_baseline(params, args.cpu_budget)
        cpu = {"value": round(cpu_ms, 1), "unit": "ms", "cores": cores, "kind": "port", "sample": note}

    learned = None
    if not args.no_learned:
        del params, grad_r, grad_i
        torch.cuda.empty_cache()
        learned = learned_init_leg(dev)

    out = {
        "metric": METRIC,
        "value": round(ms_per_step, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "voxelizer microbenchmark: 256x256x128 complex64 volume, 1M Gaussians, 8^3 tiles, "
                               "fwd+bwd with random upstream gradient", "volume": list(SHAPE),
                   "gaussians": N_GAUSS, "tile": TILE, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "kernel": f"k_{dominant}", "achieved": round(achieved, 3),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic,
                     "basis": "reference FLOPs per evaluation (26, +50 with q<=9 in the adjoint, +6 in the "
                              "forward) x the evaluations this kernel executes",
                     "evaluations": k_evals,
                     "reference_basis_tflops": round(k_flop_ref / (k_ms * 1e-3) / 1e12, 3),
                     "algorithmic_speedup": round((fwd_flop_ref + bwd_flop_ref) / (fwd_flop + bwd_flop), 3),
                     "hw_tflops": None if kc is None else round(kc["fp32_flop"] / (k_ms * 1e-3) / 1e12, 3),
                     # The kernel's shader clock under load (GRBM_GUI_ACTIVE over its duration) and the
                     # fraction of the peak scaled to that clock (157.3 TFLOP/s is quoted at 2.4 GHz).
                     "clock_ghz": None if kc is None or kc.get("clock_ghz") is None else round(kc["clock_ghz"], 3),
                     "frac_at_clock": None if kc is None or not kc.get("clock_ghz") else
                     round(achieved / (PEAK_FP32_TFLOPS * kc["clock_ghz"] / 2.4), 4),
                     "counters": None if counters is None else COUNTERS.name},
        "cpu_baseline": cpu,
        "learned_init": learned,
        "detail": {
            "median_synchronised_step_ms": round(statistics.median(per_step), 4),
            "forward_kernel_ms": round(fwd_ms, 4),
            "backward_kernel_ms": round(bwd_ms, 4),
            "binning_ms": round(statistics.median(bt), 4),
            "keep_binning_ms": round(statistics.median(kbt), 4),
            "pairs": pairs,
            "pairs_kept": pairs_kept,
            "E_total": e_total,
            "E_active": e_active,
            "E_culled": e_culled,
            "E_adjoint": e_adjoint,
            "E_active_fused_lists": e_active_kept,
            "E_culled_reference_lists": e_culled_ref,
            "E_adjoint_reference_lists": e_adjoint_ref,
            "forward_executed_tflops": fwd_flop / (fwd_ms * 1e-3) / 1e12,
            "forward_roofline_frac": fwd_flop / (fwd_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
            "reference_evals_per_s_fwd": e_total / (fwd_ms * 1e-3),
            "reference_evals_per_s_bwd": e_total / (bwd_ms * 1e-3),
            "kernels_executed_frac": (fwd_flop + bwd_flop) / ((fwd_ms + bwd_ms) * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
            "step_reference_basis_tflops": (fwd_flop_ref + bwd_flop_ref) / (ms_per_step * 1e-3) / 1e12,
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
