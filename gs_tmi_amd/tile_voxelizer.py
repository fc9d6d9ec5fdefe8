"""Tile voxelizer for 3DGSMR on MI355X: the reference's Python interface over HIP.

Drop-in for ``gaussian/tile_voxelizer.py`` of the reference (3DGSMR):

* ``TileVoxelizer(volume_shape, tile_size=8, max_radius=20, use_cuda=True,
  strict_cuda=False, device="cuda:0")`` and ``.forward(positions, scales,
  rotations, density, chunk_size=4096)`` keep the constructor and call of
  gaussian/tile_voxelizer.py:247-322 and return ``(vol_r, vol_i)`` as fp32
  ``[D, H, W]`` tensors, as the reference's GPU path does (:309).
* ``_TileVoxelizeHIP`` keeps the 11-argument ``torch.autograd.Function``
  signature of ``_TileVoxelizeCUDA`` (gaussian/tile_voxelizer.py:165-207).
* ``build_tile_pairs(center_vox, radius_vox, volume_shape, tile_size)`` keeps
  the signature and outputs of gaussian/tile_voxelizer.py:87-162; the pair order
  inside a tile is ascending Gaussian id (a stable sort).

Everything below the prologue runs in hand-written HIP for gfx950
(``csrc/gstv.hip``) through the C ABI in ``include/gstv.h``.  There is no
PyTorch or CPU fallback: where the reference would fall back to its PyTorch
tile path (:264-274, :311-322), this package raises RuntimeError.
"""

from __future__ import annotations

import math
from typing import Tuple

import torch
import torch.nn as nn

from . import _native


def _load_hip_ext(strict: bool = False):
    """Loader with the reference's contract (gaussian/tile_voxelizer.py:23-74):
    memoised; strict raises RuntimeError, non-strict returns None on failure."""
    return _native.get_lib(strict=strict)


def _quaternion_to_rotation_matrix(q: torch.Tensor) -> torch.Tensor:
    """Rotation matrices of normalised (w, x, y, z) quaternions
    (gaussian/tile_voxelizer.py:77-84)."""
    q = q / (q.norm(dim=-1, keepdim=True) + 1e-8)
    w, x, y, z = q.unbind(-1)
    two = 2.0
    rows = (
        (1 - two * y * y - two * z * z, two * x * y - two * z * w, two * x * z + two * y * w),
        (two * x * y + two * z * w, 1 - two * x * x - two * z * z, two * y * z - two * x * w),
        (two * x * z - two * y * w, two * y * z + two * x * w, 1 - two * x * x - two * y * y),
    )
    return torch.stack([torch.stack(r, dim=-1) for r in rows], dim=-2)


def build_tile_pairs(center_vox: torch.Tensor, radius_vox: torch.Tensor, volume_shape, tile_size: int):
    """Bins Gaussians into tiles on the GPU (gaussian/tile_voxelizer.py:87-162).

    Returns ``(tile_ranges int32 [T, 2], pair_ids int32 [P])``.  Like the
    reference, it reads the pair total back to the host once (:118) to size
    ``pair_ids``.
    """
    offsets, total = _native.bin_count(center_vox, radius_vox, volume_shape, tile_size)
    total_pairs = int(total.item())
    return _native.bin_fill(center_vox, radius_vox, volume_shape, tile_size, offsets, total_pairs)


class _TileVoxelizeHIP(torch.autograd.Function):
    """Forward and hand-written adjoint of the tile rasterisation
    (gaussian/tile_voxelizer.py:165-207; kernels in csrc/gstv.hip)."""

    @staticmethod
    def forward(ctx, center_vox, cov_inv, density_r, density_i, tile_ranges, pair_ids, half_shape, tile_size,
                d_size, h_size, w_size):
        _load_hip_ext(strict=True)
        args = (
            center_vox.contiguous().float(),
            cov_inv.contiguous().float(),
            density_r.contiguous().float(),
            density_i.contiguous().float(),
            tile_ranges.contiguous().int(),
            pair_ids.contiguous().int(),
            half_shape.contiguous().float(),
        )
        vol_r, vol_i = _native.forward(*args, tile_size, d_size, h_size, w_size)
        ctx.save_for_backward(*args)
        ctx.tile_size = tile_size
        ctx.dims = (d_size, h_size, w_size)
        return vol_r, vol_i

    @staticmethod
    def backward(ctx, grad_vol_r, grad_vol_i):
        center_vox, cov_inv, density_r, density_i, tile_ranges, pair_ids, half_shape = ctx.saved_tensors
        d_size, h_size, w_size = ctx.dims
        nvox = d_size * h_size * w_size
        if grad_vol_r is None:
            grad_vol_r = torch.zeros(nvox, dtype=torch.float32, device=center_vox.device)
        if grad_vol_i is None:
            grad_vol_i = torch.zeros(nvox, dtype=torch.float32, device=center_vox.device)
        g_c, g_m, g_rr, g_ri = _native.backward(
            center_vox, cov_inv, density_r, density_i, tile_ranges, pair_ids, half_shape,
            grad_vol_r.contiguous().float(), grad_vol_i.contiguous().float(),
            ctx.tile_size, d_size, h_size, w_size)
        return g_c, g_m.view(-1, 3, 3), g_rr, g_ri, None, None, None, None, None, None, None


class _TileVoxelizeFusedHIP(torch.autograd.Function):
    """The whole TileVoxelizer call as one Function: fused HIP prologue
    (gaussian/tile_voxelizer.py:281-290), GPU binning (:292-293), forward tile
    kernel, and in backward the adjoint kernel plus the fused chain rule back to
    positions, linear scales and quaternions.  Same math as composing
    ``voxel_space_params`` with ``_TileVoxelizeHIP``; about thirty small torch
    kernels fewer per step."""

    @staticmethod
    def forward(ctx, positions, scales, rotations, density, volume_shape, tile_size, pairs=None):
        _load_hip_ext(strict=True)
        d_size, h_size, w_size = volume_shape
        scales_c = scales.detach().contiguous().float()
        rot_c = _native.aligned16(rotations.detach().contiguous().float())
        # The kernels read the complex64 density interleaved, in place: no
        # real / imag copies here, and one complex gradient in backward.
        dens = density.detach().reshape(-1).resolve_conj()
        if dens.dtype != torch.complex64:
            dens = dens.to(torch.complex64)
        dv = torch.view_as_real(dens.contiguous())
        # One pass makes the prologue's outputs and the Gaussian records,
        # zeroes the adjoint's accumulator, and counts the pairs whose tile
        # meets a Gaussian's extent box: binning keeps only the pairs the tile
        # kernels evaluate (the reference's lists hold about twice as many,
        # which the kernels would skip).
        center, radius, cov, half_shape, kept = _native.keep_prologue(positions, scales_c, rot_c, dv[:, 0],
                                                                      dv[:, 1], volume_shape, tile_size)
        n = center.shape[0]
        stream = torch.cuda.current_stream(center.device)
        total_host = torch.empty(1, dtype=torch.int64, pin_memory=True)
        with torch.cuda.stream(stream):
            total_host.copy_(kept.total, non_blocking=True)
        total_ready = torch.cuda.Event()
        total_ready.record(stream)
        cap = pairs.capacity(n) if pairs is not None else None

        def exact():
            total_ready.synchronize()
            total = int(total_host.numpy()[0])
            ranges, pids = _native.bin_fill(center, radius, volume_shape, tile_size, kept.offsets, total, kept=kept)
            return total, ranges, pids

        if cap is None:
            # No capacity known: the reference's order (:118), the host reads
            # the pair total and then sizes the pair arrays.
            total, tile_ranges, pair_ids = exact()
        else:
            # Binning into the known capacity, from the total in device memory:
            # the host does not wait before queueing the forward.
            tile_ranges, pair_ids = _native.bin_fill(center, radius, volume_shape, tile_size, kept.offsets, cap,
                                                     kept=kept, capped=True)
        # The forward keeps its records and per-pair column rows for the
        # adjoint of this call, which then rebuilds neither.
        vol_r, vol_i, ctx.kept = _native.forward(center, cov, dv[:, 0], dv[:, 1], tile_ranges, pair_ids,
                                                 half_shape, tile_size, d_size, h_size, w_size, keep=kept)
        if cap is not None:
            # The total is checked while the GPU runs the forward; above the
            # capacity the lists were empty, and the call is redone exactly.
            total_ready.synchronize()
            total = int(total_host.numpy()[0])
            if total > cap:
                total, tile_ranges, pair_ids = exact()
                vol_r, vol_i, ctx.kept = _native.forward(center, cov, dv[:, 0], dv[:, 1], tile_ranges, pair_ids,
                                                         half_shape, tile_size, d_size, h_size, w_size, keep=kept)
        if pairs is not None:
            pairs.update(n, total, cap)
        ctx.save_for_backward(scales_c, rot_c, center, cov, dv, tile_ranges, pair_ids, half_shape)
        ctx.density_dtype = density.dtype
        ctx.density_shape = density.shape
        ctx.tile_size = tile_size
        ctx.dims = (d_size, h_size, w_size)
        return vol_r.view(d_size, h_size, w_size), vol_i.view(d_size, h_size, w_size)

    @staticmethod
    def backward(ctx, grad_vol_r, grad_vol_i):
        scales_c, rot_c, center, cov, dv, tile_ranges, pair_ids, half_shape = ctx.saved_tensors
        d_size, h_size, w_size = ctx.dims
        shape = (d_size, h_size, w_size)
        if grad_vol_r is None:
            grad_vol_r = torch.zeros(shape, dtype=torch.float32, device=center.device)
        if grad_vol_i is None:
            grad_vol_i = torch.zeros(shape, dtype=torch.float32, device=center.device)
        g_dv = torch.empty_like(dv)
        # Adjoint kernel, then moments -> parameter gradients in one pass.
        g_p, g_s, g_q = _native.backward_params(
            center, cov, scales_c, rot_c, tile_ranges, pair_ids, half_shape,
            grad_vol_r.contiguous().float().view(-1), grad_vol_i.contiguous().float().view(-1),
            ctx.tile_size, d_size, h_size, w_size, ctx.kept, g_dv)
        if ctx.density_dtype.is_complex:
            g_d = torch.view_as_complex(g_dv).to(ctx.density_dtype)
        else:
            g_d = g_dv[:, 0].to(ctx.density_dtype)
        return g_p, g_s, g_q, g_d.reshape(ctx.density_shape), None, None, None


def voxel_space_params(positions, scales, rotations, volume_shape):
    """The prologue of gaussian/tile_voxelizer.py:281-290, op for op.

    Returns (center_vox, radius_vox, cov_inv, half_shape).  cov_inv = T T^T is
    formed as an explicit fp32 sum over k of T_ik T_jk (true fp32, no
    reduced-precision matmul path).
    """
    d_size, h_size, w_size = volume_shape
    shape_tensor = torch.tensor([d_size, h_size, w_size], device=positions.device, dtype=torch.float32)
    half_shape = shape_tensor * 0.5
    center_vox = (positions + 1.0) * 0.5 * shape_tensor - 0.5
    scale_vox = scales * half_shape
    radius_vox = scale_vox.max(dim=-1)[0] * 3.0
    rotation = _quaternion_to_rotation_matrix(rotations)
    s_inv = torch.reciprocal(scales + 1e-8)
    transform = rotation * s_inv.unsqueeze(1)
    cov_inv = (transform.unsqueeze(2) * transform.unsqueeze(1)).sum(dim=-1)
    return center_vox, radius_vox, cov_inv, half_shape


class TileVoxelizer(nn.Module):
    """Gaussians -> complex volume through 3D tiles (gaussian/tile_voxelizer.py:247-322)."""

    def __init__(self, volume_shape: Tuple[int, int, int], tile_size: int = 8, max_radius: int = 20,
                 use_cuda: bool = True, strict_cuda: bool = False, device: str = "cuda:0"):
        super().__init__()
        self.volume_shape = tuple(int(v) for v in volume_shape)
        self.tile_size = int(tile_size)
        self.max_radius = max_radius  # kept for signature parity; unused, as in the reference
        self.device = device
        self.strict_cuda = strict_cuda
        if self.tile_size < 1:
            raise ValueError(f"tile_size must be a positive integer, got {tile_size}")
        if not use_cuda:
            raise RuntimeError("TileVoxelizer(use_cuda=False) selects the reference's PyTorch CPU backend, "
                               "which this package does not provide; the HIP backend needs a ROCm GPU")
        if not torch.cuda.is_available():
            raise RuntimeError("the HIP tile voxelizer needs a ROCm GPU (torch.cuda.is_available() is False)")
        if _load_hip_ext(strict=False) is None:
            # Both modes raise: the reference's silent PyTorch fallback is not provided.
            _load_hip_ext(strict=True)
        self.use_cuda = True
        self.backend = "tile_hip"
        # Pair-array capacity carried from call to call (binning then never
        # waits for the host; see _native.PairCapacity).
        self.pairs = _native.PairCapacity()
        print(f"[TileVoxelizer] volume={self.volume_shape}, tile_size={self.tile_size}, "
              f"max_radius={max_radius}, backend={self.backend}")

    def forward(self, positions: torch.Tensor, scales: torch.Tensor, rotations: torch.Tensor,
                density: torch.Tensor, chunk_size: int = 4096):
        del chunk_size
        return _TileVoxelizeFusedHIP.apply(positions, scales, rotations, density, self.volume_shape,
                                           self.tile_size, self.pairs)

    def forward_unfused(self, positions, scales, rotations, density):
        """The same call composed as the reference composes it: torch prologue,
        binning, and the 11-argument Function (kept for Function-level parity)."""
        d_size, h_size, w_size = self.volume_shape
        center_vox, radius_vox, cov_inv, half_shape = voxel_space_params(positions, scales, rotations,
                                                                         self.volume_shape)
        with torch.no_grad():
            tile_ranges, pair_ids = build_tile_pairs(center_vox.detach(), radius_vox.detach(), self.volume_shape,
                                                     self.tile_size)
        vol_r, vol_i = _TileVoxelizeHIP.apply(center_vox, cov_inv, density.real.contiguous(),
                                              density.imag.contiguous(), tile_ranges, pair_ids, half_shape,
                                              self.tile_size, d_size, h_size, w_size)
        return vol_r.view(d_size, h_size, w_size), vol_i.view(d_size, h_size, w_size)


def tile_count(volume_shape, tile_size: int) -> int:
    return math.prod(math.ceil(s / tile_size) for s in volume_shape)
