"""Training data and training of the initialiser network.

Training data: seeded synthetic acquisitions of the config-4 family at
128 x 128 x 64 (60-ellipsoid phantom with smooth phase, 16 coils, Poisson-disc
R=8 with a 12 x 12 calibration centre, k-space noise 1e-3).  Each is fitted by
the plain reconstruction loop (HIP tile voxelizer, Adam), and the fit becomes
the slot targets of its 8^3 patches.  Labels "plain": the plain method's own
fit -- plain initialisation (gaussian/gaussian_model.py:308-355) with one
Gaussian per slot, refined by the loop -- moment-matched into the slots
(``plain_slot_targets``).  Labels "grid" (the shipped weights): a fit that
starts from the slot grid (one Gaussian per cell, rotations fixed, no
densification), so fitted Gaussian i is slot i.  A network trained on plain
labels proposes about 80k broad Gaussians for a 256 x 256 x 128 volume
(merged cells): it reaches the NMSE target faster under a fixed-size schedule,
but densification every 100 steps splits nearly all of them three times over
and the run does not recover; the grid labels' network proposes about 300k
and converges under both schedules.

Training: the network (``InitMLP``, BF16 MFMA GEMMs) minimises ``slot_loss``
with Adam and a cosine learning-rate schedule; every batch is rotated by a
random global phase (the forward model is linear in the complex amplitude, so
a phase-rotated acquisition's fit is the phase-rotated fit).
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass

import torch

from ..recon.data import gaussian_coils, make_problem, poisson_disc_mask
from ..recon.model import GaussianSet
from ..recon.trainer import ReconConfig, Reconstructor
from .data import extract_patches, grid_gaussians, image_scale, plain_slot_targets, slot_loss, slot_targets
from .model import SLOT_FIELDS, SLOTS, InitMLP

TRAIN_SHAPE = (128, 128, 64)


def label_fit_config() -> ReconConfig:
    """Adam settings of the slot-structured label fits (rotation learning rate 0)."""
    return ReconConfig(lr_position=3e-3, lr_scale=1.5e-2, lr_rotation=0.0, lr_density=3e-2, gamma=0.998,
                       max_grad_norm=0.0, densify=False, hard_dc=True)


def training_problem(seed: int, device, shape=TRAIN_SHAPE, coils=None):
    mask = poisson_disc_mask((shape[1], shape[2]), 8, max(2, 24 * shape[1] // 256), seed)
    return make_problem(shape, 60, 16, mask, seed, noise_sigma=1e-3, device=device, phantom_on_device=True,
                        coils=coils)


def fit_slots(problem, voxelizer, steps: int, cfg: ReconConfig | None = None, eval_every: int | None = None):
    """Slot-structured plain fit; returns (GaussianSet, image scale, history)."""
    scale = image_scale(problem.zero_filled)
    g = GaussianSet(*grid_gaussians(problem.zero_filled, scale))
    rec = Reconstructor(problem, g, voxelizer, cfg or label_fit_config())
    hist = rec.run(steps, eval_every=eval_every or steps)
    return g, scale, hist


def fit_plain(problem, voxelizer, steps: int, n_points: int, cfg: ReconConfig | None = None, seed: int = 0,
              eval_every: int | None = None):
    """The plain method's fit: plain initialisation (n_points Gaussians at
    above-quantile voxels, identity rotations) refined by the loop with the
    label fits' Adam settings; returns (GaussianSet, image scale, history)."""
    scale = image_scale(problem.zero_filled)
    g = GaussianSet.from_image(problem.zero_filled, n_points, seed=seed)
    rec = Reconstructor(problem, g, voxelizer, cfg or label_fit_config())
    hist = rec.run(steps, eval_every=eval_every or steps)
    return rec.g, scale, hist


@dataclass
class SlotDataset:
    inputs: torch.Tensor     # bf16 [n, 1024]
    targets: torch.Tensor    # fp32 [n, 32, 9]
    phantom: torch.Tensor    # int32 [n], index of the source acquisition

    def __len__(self):
        return self.inputs.shape[0]


def build_dataset(seeds, device, fit_steps=500, shape=TRAIN_SHAPE, log=None, labels: str = "grid") -> SlotDataset:
    from ..tile_voxelizer import TileVoxelizer
    from .mfma import to_bf16

    vox = TileVoxelizer(shape, tile_size=8, device=str(device))
    coils = torch.from_numpy(gaussian_coils(shape, 16)).to(device)
    xs, ts, ps = [], [], []
    for i, seed in enumerate(seeds):
        t0 = time.perf_counter()
        p = training_problem(seed, device, shape, coils=coils)
        t1 = time.perf_counter()
        if labels == "plain":
            n_slots = (shape[0] // 8) * (shape[1] // 8) * (shape[2] // 8) * SLOTS
            g, scale, hist = fit_plain(p, vox, fit_steps, n_slots, seed=seed)
        elif labels == "grid":
            g, scale, hist = fit_slots(p, vox, fit_steps)
        else:
            raise ValueError(f"labels must be 'plain' or 'grid', got {labels!r}")
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        xs.append(to_bf16(extract_patches(p.zero_filled, scale)))
        ts.append(plain_slot_targets(g, shape, scale) if labels == "plain" else slot_targets(g, shape, scale))
        ps.append(torch.full((xs[-1].shape[0],), i, dtype=torch.int32, device=device))
        if log:
            log({"phantom": i, "seed": seed, "problem_s": round(t1 - t0, 3), "fit_s": round(t2 - t1, 3),
                 "nmse0": hist.rows[0]["nmse"], "nmse": hist.rows[-1]["nmse"], "psnr": hist.rows[-1]["psnr"],
                 "occupied": float(ts[-1][..., 8].mean())})
        del p, g
    return SlotDataset(torch.cat(xs), torch.cat(ts), torch.cat(ps))


def _rotate_phase(x: torch.Tensor, t: torch.Tensor, gen: torch.Generator):
    """Multiplies each sample's complex patch and target amplitudes by e^{i phi}."""
    n = x.shape[0]
    phi = torch.rand(n, 1, device=x.device, generator=gen) * (2 * math.pi)
    c, s = torch.cos(phi), torch.sin(phi)
    xf = x.float()
    half = xf.shape[1] // 2
    xr, xi = xf[:, :half], xf[:, half:]
    xo = torch.cat([c * xr - s * xi, s * xr + c * xi], dim=1)
    to = t.clone()
    ar, ai = t[..., 6], t[..., 7]
    to[..., 6] = c * ar - s * ai
    to[..., 7] = s * ar + c * ai
    return xo, to


def train_network(data: SlotDataset, epochs=30, batch=2048, lr=1e-3, val_phantoms=10, seed=0, log=None,
                  model: InitMLP | None = None):
    """Trains (or continues) the network; the last `val_phantoms` acquisitions are held out."""
    dev = data.inputs.device
    n_ph = int(data.phantom.max()) + 1
    train_idx = torch.nonzero(data.phantom < n_ph - val_phantoms, as_tuple=False).squeeze(-1)
    val_idx = torch.nonzero(data.phantom >= n_ph - val_phantoms, as_tuple=False).squeeze(-1)
    model = model or InitMLP(seed=seed).to(dev)
    opt = torch.optim.Adam(model.parameters(), lr=lr)
    steps_per_epoch = max(1, train_idx.numel() // batch)
    total = epochs * steps_per_epoch
    sched = torch.optim.lr_scheduler.LambdaLR(opt, lambda it: 0.5 * (1 + math.cos(math.pi * min(it, total) / total)))
    gen = torch.Generator(device=dev).manual_seed(seed)
    history = []
    t_start = time.perf_counter()
    for ep in range(epochs):
        perm = train_idx[torch.randperm(train_idx.numel(), device=dev, generator=gen)]
        tot = torch.zeros(3, device=dev)
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch]
            x, t = _rotate_phase(data.inputs[idx], data.targets[idx], gen)
            loss, bce, reg = slot_loss(model(x), t)
            opt.zero_grad(set_to_none=True)
            loss.backward()
            opt.step()
            sched.step()
            tot += torch.stack([loss.detach(), bce, reg])
        row = {"epoch": ep + 1, "train": (tot / steps_per_epoch).tolist(), "val": evaluate_network(model, data, val_idx),
               "elapsed_s": round(time.perf_counter() - t_start, 2)}
        history.append(row)
        if log:
            log(row)
    return model, history


@torch.no_grad()
def evaluate_network(model, data: SlotDataset, idx, batch=8192):
    if idx.numel() == 0:
        return None
    tot = torch.zeros(3, device=data.inputs.device)
    acc = torch.zeros((), device=data.inputs.device)
    for b in range(0, idx.numel(), batch):
        j = idx[b:b + batch]
        pred = model(data.inputs[j])
        loss, bce, reg = slot_loss(pred, data.targets[j])
        tot += torch.stack([loss, bce, reg]) * j.numel()
        acc += ((pred[..., 8] > 0) == (data.targets[j][..., 8] > 0.5)).float().sum()
    out = (tot / idx.numel()).tolist()
    return out + [float(acc) / (idx.numel() * SLOTS)]


def save_weights(model: InitMLP, path) -> None:
    """bf16 weights and biases (the precision the GEMMs consume) as safetensors."""
    from safetensors.torch import save_file

    tensors = {}
    for i, layer in enumerate(model.layers):
        tensors[f"layers.{i}.weight"] = layer.weight.detach().to(torch.bfloat16).cpu().contiguous()
        tensors[f"layers.{i}.bias"] = layer.bias.detach().float().cpu().contiguous()
    save_file(tensors, str(path), metadata={"format": "gstv-init-mlp", "slots": str(SLOTS),
                                            "slot_fields": str(SLOT_FIELDS)})


def load_weights(path, device) -> InitMLP:
    from safetensors.torch import load_file

    t = load_file(str(path))
    model = InitMLP()
    for i, layer in enumerate(model.layers):
        w, b = t[f"layers.{i}.weight"], t[f"layers.{i}.bias"]
        if tuple(w.shape) != tuple(layer.weight.shape):
            raise ValueError(f"layer {i}: weight shape {tuple(w.shape)} != {tuple(layer.weight.shape)}")
        layer.weight.data.copy_(w.float())
        layer.bias.data.copy_(b.float())
    return model.to(device)
