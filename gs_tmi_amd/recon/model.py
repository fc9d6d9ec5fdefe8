"""The Gaussian parameter set of 3DGSMR (gaussian/gaussian_model.py:22-355 of the
reference): positions in [-1, 1], log-scales, (w, x, y, z) quaternions and a
complex amplitude stored as two real parameters; plain initialisation from the
zero-filled image; split, clone and prune with a row map for the optimizer state."""

from __future__ import annotations

import numpy as np
import torch
import torch.nn as nn

from ..tile_voxelizer import _quaternion_to_rotation_matrix


class GaussianSet(nn.Module):
    def __init__(self, positions, log_scales, rotations, density_real, density_imag):
        super().__init__()
        self.positions = nn.Parameter(positions.float().contiguous())
        self.log_scales = nn.Parameter(log_scales.float().contiguous())
        self.rotations = nn.Parameter(rotations.float().contiguous())
        self.density_real = nn.Parameter(density_real.float().contiguous())
        self.density_imag = nn.Parameter(density_imag.float().contiguous())

    @property
    def num_points(self) -> int:
        return self.positions.shape[0]

    def scales(self) -> torch.Tensor:
        return torch.exp(self.log_scales)

    def density(self) -> torch.Tensor:
        return torch.complex(self.density_real, self.density_imag)

    def param_list(self):
        return [self.positions, self.log_scales, self.rotations, self.density_real, self.density_imag]

    # ------------------------------------------------------------ init --
    @classmethod
    def from_image(cls, image: torch.Tensor, num_points: int, mode: str = "importance", scale_vox: float | None = 1.0,
                   density_scale_k: float = 0.2, seed: int = 0, quantile: float = 0.90):
        """Plain initialisation (gaussian/gaussian_model.py:308-355): Gaussians at voxel
        centres chosen at random among voxels whose zero-filled magnitude exceeds the
        given quantile ("importance") or among all voxels ("random"); amplitude =
        density_scale_k x image value; identity rotation.  Isotropic voxel scale
        `scale_vox`, or, when None, the mean distance to the 3 nearest neighbours
        (:282-306)."""
        dev = image.device
        shape = tuple(image.shape)
        mag = image.abs().flatten()
        total = mag.numel()
        gen = torch.Generator(device="cpu").manual_seed(seed)
        if num_points >= total:
            idx = torch.arange(total)
        else:
            if mode == "importance":
                thr = torch.quantile(mag.float().cpu()[torch.randperm(total, generator=gen)[:1_000_000]], quantile)
                cand = torch.nonzero(mag.cpu() > thr, as_tuple=False).squeeze(-1)
                if cand.numel() < num_points:
                    cand = torch.arange(total)
            elif mode == "random":
                cand = torch.arange(total)
            else:
                raise ValueError(mode)
            idx = cand[torch.randperm(cand.numel(), generator=gen)[:num_points]]
        d, h, w = shape
        z, rem = idx // (h * w), idx % (h * w)
        y, x = rem // w, rem % w
        pos = torch.stack([(z.float() + 0.5) / d * 2 - 1, (y.float() + 0.5) / h * 2 - 1,
                           (x.float() + 0.5) / w * 2 - 1], dim=-1)
        hs = torch.tensor([d, h, w], dtype=torch.float32) / 2
        if scale_vox is not None:
            scales = (scale_vox / hs).expand(pos.shape[0], 3).clone()
        else:
            scales = _knn_scales(pos).expand(-1, 3).clone()
        vals = image.flatten()[idx.to(dev)].cpu() * density_scale_k
        rot = torch.zeros(pos.shape[0], 4)
        rot[:, 0] = 1.0
        return cls(pos.to(dev), torch.log(scales.clamp_min(1e-8)).to(dev), rot.to(dev), vals.real.to(dev),
                   vals.imag.to(dev))

    # ---------------------------------------------------- densify / prune --
    def _replace(self, keep: torch.Tensor, new_rows=None):
        """Keeps rows `keep` (index tensor) and appends `new_rows`; returns the row
        map (old row index per new row, -1 for appended rows)."""
        tensors = [p.data[keep] for p in self.param_list()]
        n_new = 0
        if new_rows is not None:
            n_new = new_rows[0].shape[0]
            tensors = [torch.cat([a, b.to(a.dtype)], dim=0) for a, b in zip(tensors, new_rows)]
        self.positions, self.log_scales, self.rotations, self.density_real, self.density_imag = \
            [nn.Parameter(t.contiguous()) for t in tensors]
        return torch.cat([keep, torch.full((n_new,), -1, dtype=keep.dtype, device=keep.device)])

    @torch.no_grad()
    def densify_and_split(self, grad_norm, grad_threshold, scale_threshold, offset_factor=1.0):
        """Long-axis split (gaussian/gaussian_model.py:85-138): each selected parent is
        replaced by two children with the longest axis halved and the others times
        0.85, shifted by +-(child long scale x offset_factor) along the rotated long
        axis, amplitude x 0.6.  Returns (count, row map) or (0, None)."""
        scales = self.scales()
        mask = (grad_norm > grad_threshold) & (scales.max(dim=-1)[0] > scale_threshold)
        cnt = int(mask.sum())
        if cnt == 0:
            return 0, None
        ps, pp, pr = scales[mask], self.positions[mask], self.rotations[mask]
        ar = torch.arange(cnt, device=ps.device)
        longest = ps.argmax(dim=-1)
        child = ps * 0.85
        child[ar, longest] = ps[ar, longest] * 0.5
        shift = torch.zeros_like(pp)
        shift[ar, longest] = child[ar, longest] * offset_factor
        shift = (_quaternion_to_rotation_matrix(pr) @ shift.unsqueeze(-1)).squeeze(-1)
        new = (torch.cat([pp + shift, pp - shift]), torch.log(torch.cat([child, child]) + 1e-8),
               torch.cat([pr, pr]), torch.cat([self.density_real[mask] * 0.6] * 2),
               torch.cat([self.density_imag[mask] * 0.6] * 2))
        keep = torch.nonzero(~mask, as_tuple=False).squeeze(-1)
        return cnt, self._replace(keep, new)

    @torch.no_grad()
    def densify_and_clone(self, grad_norm, grad_threshold, scale_threshold):
        """Clone (gaussian/gaussian_model.py:140-177): small high-gradient Gaussians are
        duplicated, both copies at half amplitude."""
        scales = self.scales()
        mask = (grad_norm > grad_threshold) & (scales.max(dim=-1)[0] <= scale_threshold)
        cnt = int(mask.sum())
        if cnt == 0:
            return 0, None
        self.density_real.data[mask] *= 0.5
        self.density_imag.data[mask] *= 0.5
        new = tuple(p.data[mask].clone() for p in self.param_list())
        keep = torch.arange(self.num_points, device=mask.device)
        return cnt, self._replace(keep, new)

    @torch.no_grad()
    def prune(self, keep_mask):
        keep = torch.nonzero(keep_mask, as_tuple=False).squeeze(-1)
        return int((~keep_mask).sum()), self._replace(keep)


def _knn_scales(pos: torch.Tensor) -> torch.Tensor:
    """Mean distance to the 3 nearest other Gaussians (normalised units)."""
    from scipy.spatial import cKDTree

    p = pos.cpu().numpy().astype(np.float32)
    if p.shape[0] < 2:
        return torch.full((p.shape[0], 1), 1e-3)
    k = min(4, p.shape[0])
    dist, _ = cKDTree(p).query(p, k=k, workers=-1)
    nd = np.asarray(dist[:, 1:], dtype=np.float32)
    return torch.from_numpy(np.maximum(nd.mean(axis=1, keepdims=True), 1e-6))
