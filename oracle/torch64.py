"""Differentiable FP64 restatement of the whole voxelizer call.  TEST INFRASTRUCTURE ONLY.

Float64 torch (CPU) version of ``TileVoxelizer.forward`` of the reference
(gaussian/tile_voxelizer.py:279-322 with the kernel math of
gaussian/cuda_ext/tile_voxelize_cuda_kernel.cu:118-141), so torch autograd gives
FP64 gradients with respect to positions, log-scales, quaternions and complex
amplitudes.  Tile lists come from the caller (the oracle's fp32 binning), so
the set of evaluated pairs is the reference's.

Pinned by tests/test_oracle.py against the NumPy oracle's forward and
analytic adjoint (which are in turn pinned by central finite differences),
and by FP64 central finite differences of its own gradients with respect to
positions, log-scales, quaternions and amplitudes.
"""

from __future__ import annotations

import math

import numpy as np
import torch


def quat_to_rot64(q: torch.Tensor) -> torch.Tensor:
    q = q / (q.norm(dim=-1, keepdim=True) + 1e-8)
    w, x, y, z = q.unbind(-1)
    r0 = torch.stack([1 - 2 * y * y - 2 * z * z, 2 * x * y - 2 * z * w, 2 * x * z + 2 * y * w], dim=-1)
    r1 = torch.stack([2 * x * y + 2 * z * w, 1 - 2 * x * x - 2 * z * z, 2 * y * z - 2 * x * w], dim=-1)
    r2 = torch.stack([2 * x * z - 2 * y * w, 2 * y * z + 2 * x * w, 1 - 2 * x * x - 2 * y * y], dim=-1)
    return torch.stack([r0, r1, r2], dim=-2)


def voxelize64(positions, log_scales, rotations, rho_r, rho_i, shape, tile, tile_ranges, pair_ids,
               return_active=False):
    """Returns (vol_r, vol_i) float64 flat [D*H*W]; all inputs float64 tensors.
    With ``return_active``, also the list of per-tile [voxels, pairs] masks of
    the evaluations with q <= 9 (for finite-difference checks that must not
    cross the cutoff)."""
    d, h, w = shape
    shape_t = torch.tensor([d, h, w], dtype=torch.float64)
    hs = shape_t * 0.5
    center = (positions + 1.0) * 0.5 * shape_t - 0.5
    scales = torch.exp(log_scales)
    rot = quat_to_rot64(rotations)
    t = rot * (1.0 / (scales + 1e-8)).unsqueeze(1)
    cov = t @ t.transpose(1, 2)
    nth, ntw = math.ceil(h / tile), math.ceil(w / tile)
    flats, vals_r, vals_i, acts = [], [], [], []
    ranges = np.asarray(tile_ranges)
    pids = np.asarray(pair_ids)
    for tid in range(ranges.shape[0]):
        s, e = int(ranges[tid, 0]), int(ranges[tid, 1])
        if s >= e:
            continue
        gids = torch.from_numpy(pids[s:e].astype(np.int64))
        td, rem = divmod(tid, nth * ntw)
        th, tw = divmod(rem, ntw)
        zz = torch.arange(td * tile, min(td * tile + tile, d), dtype=torch.float64)
        yy = torch.arange(th * tile, min(th * tile + tile, h), dtype=torch.float64)
        xx = torch.arange(tw * tile, min(tw * tile + tile, w), dtype=torch.float64)
        g0, g1, g2 = torch.meshgrid(zz, yy, xx, indexing="ij")
        vox = torch.stack([g0.reshape(-1), g1.reshape(-1), g2.reshape(-1)], dim=-1)
        dvec = (vox.unsqueeze(1) - center[gids].unsqueeze(0)) / hs
        q = torch.einsum("vgi,gij,vgj->vg", dvec, cov[gids], dvec)
        wgt = torch.exp(-0.5 * q) * (q <= 9.0).to(torch.float64)
        acts.append((q <= 9.0).detach())
        vals_r.append(wgt @ rho_r[gids])
        vals_i.append(wgt @ rho_i[gids])
        flats.append((vox[:, 0] * h + vox[:, 1]) * w + vox[:, 2])
    vol_r = torch.zeros(d * h * w, dtype=torch.float64)
    vol_i = torch.zeros(d * h * w, dtype=torch.float64)
    if flats:
        idx = torch.cat(flats).long()
        vol_r = vol_r.index_add(0, idx, torch.cat(vals_r))
        vol_i = vol_i.index_add(0, idx, torch.cat(vals_i))
    if return_active:
        return vol_r, vol_i, acts
    return vol_r, vol_i
