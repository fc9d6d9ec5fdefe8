"""Learned initialisation with the plain initialisation as fallback.

``LearnedInitialiser.propose`` runs the network over every 8^3 patch of the
zero-filled SENSE-adjoint image (BF16 MFMA GEMMs, batches of 16,384 patches)
and decodes the slots whose existence probability exceeds the threshold.

``initialise`` judges every learned initialisation by its data-consistency
residual ||M F S x - y|| / ||y|| (and reports its NMSE against the phantom
when the phantom is known).  A residual below `accept_residual` (the learned
set explains at least 1 - 0.3^2 = 91 % of the measured k-space energy) is
accepted outright; otherwise the learned set must beat the plain
initialisation's residual (gaussian/gaussian_model.py:308-355 of the
reference, here ``GaussianSet.from_image``), and when it does not, or the
network proposes no Gaussian, the plain set is used instead.
"""

from __future__ import annotations

import time
from pathlib import Path

import torch

from ..recon.data import fft3c
from ..recon.metrics import nmse, psnr
from ..recon.model import GaussianSet
from .data import decode, extract_patches, image_scale
from .mfma import to_bf16

DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "weights" / "init_mlp.safetensors"


class LearnedInitialiser:
    def __init__(self, model=None, weights=None, device="cuda:0", threshold=0.5, max_points=None):
        from .training import load_weights

        if model is None:
            path = Path(weights) if weights is not None else DEFAULT_WEIGHTS
            if not path.exists():
                raise FileNotFoundError(f"initialiser weights not found: {path}")
            model = load_weights(path, device)
        self.model = model.eval()
        self.threshold = threshold
        self.max_points = max_points

    @torch.no_grad()
    def propose(self, zero_filled: torch.Tensor, batch=16384):
        """Returns (positions, log_scales, rotations, density_real, density_imag)."""
        shape = tuple(zero_filled.shape)
        scale = image_scale(zero_filled)
        x = extract_patches(zero_filled, scale)
        preds = [self.model(to_bf16(x[b:b + batch])) for b in range(0, x.shape[0], batch)]
        pred = torch.cat(preds)
        return decode(pred, shape, scale, self.threshold, self.max_points)


@torch.no_grad()
def judge(problem, gaussians: GaussianSet, voxelizer) -> dict:
    """Data-consistency residual of the initial volume, and (phantom known) NMSE
    and PSNR of its hard-data-consistent image."""
    out = voxelizer(gaussians.positions, gaussians.scales(), gaussians.rotations, gaussians.density())
    vol = torch.complex(out[0], out[1]) if isinstance(out, tuple) else out
    k = fft3c(problem.coils * vol.unsqueeze(0))
    m = problem.mask.unsqueeze(0)
    res = torch.linalg.vector_norm(k * m - problem.kspace) / torch.linalg.vector_norm(problem.kspace)
    from ..recon.data import ifft3c

    dc = torch.sum(torch.conj(problem.coils) * ifft3c(k * (1.0 - m) + problem.kspace), dim=0)
    return {"dc_residual": float(res), "nmse": nmse(dc, problem.target), "psnr": psnr(dc, problem.target),
            "n": gaussians.num_points}


def initialise(problem, voxelizer, initialiser: LearnedInitialiser | None, plain_points=300_000, seed=0,
               plain_scale_vox=None, accept_residual=0.3):
    """Learned initialisation, judged, with the plain one as fallback; returns (GaussianSet, info)."""
    sync = torch.cuda.synchronize if problem.target.is_cuda else (lambda: None)
    sync()
    t0 = time.perf_counter()
    info = {"used": "plain"}
    chosen = None
    if initialiser is not None:
        params = initialiser.propose(problem.zero_filled)
        if params[0].shape[0] > 0:
            learned = GaussianSet(*params)
            info["learned"] = judge(problem, learned, voxelizer)
            if info["learned"]["dc_residual"] < accept_residual:
                chosen, info["used"] = learned, "learned"
    if chosen is None:
        plain = GaussianSet.from_image(problem.zero_filled, plain_points, mode="importance",
                                       scale_vox=plain_scale_vox, seed=seed)
        info["plain"] = judge(problem, plain, voxelizer)
        chosen = plain
        if "learned" in info and info["learned"]["dc_residual"] < info["plain"]["dc_residual"]:
            chosen, info["used"] = learned, "learned"
    sync()
    info["init_s"] = time.perf_counter() - t0
    return chosen, info
