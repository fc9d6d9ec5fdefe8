"""PyTorch CPU restatement of the reference's per-tile voxelizer.  TEST / BASELINE INFRASTRUCTURE ONLY.

A from-scratch fp32 restatement of ``_tile_forward_pytorch``
(gaussian/tile_voxelizer.py:210-244): a Python loop over tiles, per tile an
``einsum`` quadratic form over [voxels, Gaussians, 3], ``exp * mask`` with the
q <= 9 cutoff, and a ``scatter_add_`` into the flat complex volume.  Backward
is torch autograd of that per-tile graph.

``bench.py`` times it on the host cores as the CPU baseline ("port"); the
tests check it against the FP64 oracle.  The product package never calls it.
"""

from __future__ import annotations

import math

import torch


def _tile_geometry(tile_id, shape, tile):
    d, h, w = shape
    nth, ntw = math.ceil(h / tile), math.ceil(w / tile)
    td, rem = divmod(tile_id, nth * ntw)
    th, tw = divmod(rem, ntw)
    zz = torch.arange(td * tile, min(td * tile + tile, d), dtype=torch.float32)
    yy = torch.arange(th * tile, min(th * tile + tile, h), dtype=torch.float32)
    xx = torch.arange(tw * tile, min(tw * tile + tile, w), dtype=torch.float32)
    g0, g1, g2 = torch.meshgrid(zz, yy, xx, indexing="ij")
    vox = torch.stack([g0, g1, g2], dim=-1).reshape(-1, 3)
    flat = g0.reshape(-1).long() * h * w + g1.reshape(-1).long() * w + g2.reshape(-1).long()
    return vox, flat


def _tile_contrib(tile_id, center_vox, cov_inv, density, half_shape, tile_ranges, pair_ids, shape, tile):
    s, e = int(tile_ranges[tile_id, 0]), int(tile_ranges[tile_id, 1])
    if s >= e:
        return None
    gids = pair_ids[s:e].long()
    vox, flat = _tile_geometry(tile_id, shape, tile)
    diff = (vox.unsqueeze(1) - center_vox[gids].unsqueeze(0)) / half_shape.view(1, 1, 3)
    q = torch.einsum("vgd,gde,vge->vg", diff, cov_inv[gids], diff)
    weights = torch.exp(-0.5 * q) * (q <= 9.0).float()
    return flat, (weights * density[gids].unsqueeze(0)).sum(dim=1)


def tile_forward(center_vox, cov_inv, density, half_shape, tile_ranges, pair_ids, shape, tile, tiles=None):
    """Complex64 volume [D, H, W] (fp32 math) over all tiles or the given subset."""
    d, h, w = shape
    vol = torch.zeros(d * h * w, dtype=torch.complex64)
    ids = range(tile_ranges.shape[0]) if tiles is None else tiles
    for t in ids:
        r = _tile_contrib(int(t), center_vox, cov_inv, density, half_shape, tile_ranges, pair_ids, shape, tile)
        if r is not None:
            vol.scatter_add_(0, r[0], r[1])
    return vol.view(d, h, w)


def tile_forward_backward(center_vox, cov_inv, density_r, density_i, half_shape, tile_ranges, pair_ids, grad,
                          shape, tile, tiles):
    """Per-tile forward immediately followed by its autograd backward for the
    loss <G, vol> (separable per tile, so memory stays O(tile)).  Returns the
    accumulated gradients (d_center, d_cov, d_rho_r, d_rho_i)."""
    c = center_vox.detach().requires_grad_()
    m = cov_inv.detach().requires_grad_()
    rr = density_r.detach().requires_grad_()
    ri = density_i.detach().requires_grad_()
    g_r = grad.real.reshape(-1)
    g_i = grad.imag.reshape(-1)
    for t in tiles:
        dens = torch.complex(rr, ri)
        r = _tile_contrib(int(t), c, m, dens, half_shape, tile_ranges, pair_ids, shape, tile)
        if r is None:
            continue
        flat, val = r
        loss = (val.real * g_r[flat]).sum() + (val.imag * g_i[flat]).sum()
        loss.backward()
    return c.grad, m.grad, rr.grad, ri.grad


class CpuTileVoxelizer:
    """The reference's CPU voxelizer call (gaussian/tile_voxelizer.py:279-322 with
    `_tile_forward_pytorch`), restated: torch prologue, NumPy binning of the
    oracle, per-tile fp32 forward with torch autograd.  Returns a complex64
    [D, H, W] volume like the reference's CPU path.  Test infrastructure: it lets
    the reconstruction harness run the reference's CPU path for parity runs."""

    def __init__(self, volume_shape, tile_size=8):
        self.volume_shape = tuple(volume_shape)
        self.tile_size = tile_size

    def __call__(self, positions, scales, rotations, density):
        from oracle import voxelizer_oracle as vo

        d, h, w = self.volume_shape
        shape_t = torch.tensor([d, h, w], dtype=torch.float32)
        half_shape = shape_t * 0.5
        center = (positions + 1.0) * 0.5 * shape_t - 0.5
        radius = (scales * half_shape).max(dim=-1)[0] * 3.0
        q = rotations / (rotations.norm(dim=-1, keepdim=True) + 1e-8)
        qw, qx, qy, qz = q.unbind(-1)
        rot = torch.stack([
            torch.stack([1 - 2 * qy * qy - 2 * qz * qz, 2 * qx * qy - 2 * qz * qw, 2 * qx * qz + 2 * qy * qw], -1),
            torch.stack([2 * qx * qy + 2 * qz * qw, 1 - 2 * qx * qx - 2 * qz * qz, 2 * qy * qz - 2 * qx * qw], -1),
            torch.stack([2 * qx * qz - 2 * qy * qw, 2 * qy * qz + 2 * qx * qw, 1 - 2 * qx * qx - 2 * qy * qy], -1),
        ], -2)
        t = rot * torch.reciprocal(scales + 1e-8).unsqueeze(1)
        cov = (t.unsqueeze(2) * t.unsqueeze(1)).sum(-1)
        ranges, pids = vo.bin_tiles(center.detach().numpy(), radius.detach().numpy(), self.volume_shape,
                                    self.tile_size)
        return tile_forward(center, cov, density, half_shape, torch.from_numpy(ranges), torch.from_numpy(pids),
                            self.volume_shape, self.tile_size)
