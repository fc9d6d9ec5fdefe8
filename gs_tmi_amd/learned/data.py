"""Patches, slot grid, slot targets and decoding for the learned initialiser.

A volume is cut into 8^3 patches; patch p's input is its zero-filled complex
values (512 real then 512 imaginary parts) divided by the image scale (the
99.5th percentile of the zero-filled magnitude).  Each patch has 32 slots, one
per cell of a 4 x 4 x 2 grid of 2 x 2 x 4-voxel cells (D, H, W).  Slot s of
patch p is global slot p * 32 + s; patches are in row-major (pd, ph, pw)
order and cells in row-major (cd, ch, cw) order.

Slot fields: centre offset from the cell centre in half-cell units (3), log of
the voxel sigma per axis (3), complex amplitude over the image scale (2), and
an existence logit (1).  Rotations are the identity, as in plain
initialisation (gaussian/gaussian_model.py:54-56 of the reference).

Targets come from fits of the ordinary reconstruction loop, in one of two
forms:

* plain fits (``plain_slot_targets``): the plain initialisation
  (gaussian/gaussian_model.py:308-355 of the reference) fitted by the loop;
  the fitted Gaussians whose centres fall in a cell are merged into one
  axis-aligned Gaussian with the same complex mass, mass-weighted centre and
  the diagonal of the mass-weighted second moment (moment matching);
* slot-grid fits (``slot_targets``): one Gaussian placed at every cell centre
  (``grid_gaussians``) and fitted with rotations held fixed and no
  densification, so fitted Gaussian i is the fit of slot i.

A slot exists when its amplitude is at least `exist_threshold` of the image
scale.
"""

from __future__ import annotations

import math

import torch

from .model import PATCH, SLOT_FIELDS, SLOTS

CELL = (2, 2, 4)                     # voxels per cell along D, H, W
GRID = (PATCH // CELL[0], PATCH // CELL[1], PATCH // CELL[2])  # 4 x 4 x 2 cells per patch
OFFSET_LIMIT = 2.0                   # |offset| clamp, half-cells
LOG_SIGMA_RANGE = (math.log(0.25), math.log(8.0))
# Mass of a 3D Gaussian inside q <= 9 (chi-square, 3 dof): erf(3/sqrt2) - 3 sqrt(2/pi) e^-4.5.
INSIDE_CUTOFF = math.erf(3 / math.sqrt(2)) - 3 * math.sqrt(2 / math.pi) * math.exp(-4.5)


def image_scale(zero_filled: torch.Tensor) -> float:
    mag = zero_filled.abs().flatten().float()
    if mag.numel() > 4_000_000:
        g = torch.Generator(device="cpu").manual_seed(0)
        mag = mag[torch.randperm(mag.numel(), generator=g)[:4_000_000].to(mag.device)]
    return max(float(torch.quantile(mag, 0.995)), 1e-12)


def patch_grid(shape):
    if any(s % PATCH for s in shape):
        raise ValueError(f"volume shape {tuple(shape)} must be a multiple of {PATCH} on every axis")
    return tuple(s // PATCH for s in shape)


def extract_patches(zero_filled: torch.Tensor, scale: float) -> torch.Tensor:
    """[n_patches, 1024] fp32 network inputs in patch order."""
    d, h, w = zero_filled.shape
    pd, ph, pw = patch_grid((d, h, w))
    x = zero_filled / scale
    x = x.reshape(pd, PATCH, ph, PATCH, pw, PATCH).permute(0, 2, 4, 1, 3, 5).reshape(pd * ph * pw, PATCH ** 3)
    return torch.cat([x.real, x.imag], dim=1).float().contiguous()


def slot_cell_centres(shape, device) -> torch.Tensor:
    """[n_slots, 3] voxel coordinates of every cell centre, in global slot order."""
    pd, ph, pw = patch_grid(shape)
    p = torch.stack(torch.meshgrid(torch.arange(pd), torch.arange(ph), torch.arange(pw), indexing="ij"), -1)
    c = torch.stack(torch.meshgrid(torch.arange(GRID[0]), torch.arange(GRID[1]), torch.arange(GRID[2]),
                                   indexing="ij"), -1)
    cs = torch.tensor(CELL, dtype=torch.float32)
    origin = p.reshape(-1, 1, 3).float() * PATCH
    local = (c.reshape(1, -1, 3).float() + 0.5) * cs - 0.5
    return (origin + local).reshape(-1, 3).to(device)


def _to_normalised(centre_vox, shape):
    shp = torch.tensor(shape, dtype=torch.float32, device=centre_vox.device)
    return (centre_vox + 0.5) / shp * 2 - 1


def _to_voxel(positions, shape):
    shp = torch.tensor(shape, dtype=torch.float32, device=positions.device)
    return (positions + 1.0) * 0.5 * shp - 0.5


def slot_params_to_gaussians(fields: torch.Tensor, slot_index: torch.Tensor, shape, scale: float):
    """Slot fields [n, 8] (offset, log sigma, amplitude) of global slots `slot_index`
    -> (positions, log_scales, rotations, density_real, density_imag)."""
    dev = fields.device
    centre = slot_cell_centres(shape, dev)[slot_index]
    half = torch.tensor(CELL, dtype=torch.float32, device=dev) / 2
    vox = centre + fields[:, :3].clamp(-OFFSET_LIMIT, OFFSET_LIMIT) * half
    hs = torch.tensor(shape, dtype=torch.float32, device=dev) / 2
    log_sig = fields[:, 3:6].clamp(*LOG_SIGMA_RANGE)
    rot = torch.zeros(fields.shape[0], 4, device=dev)
    rot[:, 0] = 1.0
    return (_to_normalised(vox, shape), log_sig - torch.log(hs), rot, fields[:, 6] * scale, fields[:, 7] * scale)


@torch.no_grad()
def grid_gaussians(zero_filled: torch.Tensor, scale: float):
    """One Gaussian per slot at the cell centre, sigma = half the cell edge per
    axis, amplitude = cell mean of the zero-filled image times cell volume over
    the Gaussian's integral inside the q <= 9 cutoff (so the initial volume
    matches the cell means)."""
    shape = tuple(zero_filled.shape)
    pd, ph, pw = patch_grid(shape)
    cells = zero_filled.reshape(pd, GRID[0], CELL[0], ph, GRID[1], CELL[1], pw, GRID[2], CELL[2])
    mean = cells.mean(dim=(2, 5, 8)).permute(0, 2, 4, 1, 3, 5).reshape(-1)
    sig = torch.tensor(CELL, dtype=torch.float32) / 2
    gain = float(CELL[0] * CELL[1] * CELL[2] / ((2 * math.pi) ** 1.5 * float(sig.prod()) * INSIDE_CUTOFF))
    amp = mean * gain / scale
    n = amp.numel()
    fields = torch.zeros(n, 8, device=zero_filled.device)
    fields[:, 3:6] = torch.log(sig).to(zero_filled.device)
    fields[:, 6], fields[:, 7] = amp.real, amp.imag
    return slot_params_to_gaussians(fields, torch.arange(n, device=zero_filled.device), shape, scale)


@torch.no_grad()
def slot_targets(gaussians, shape, scale: float, exist_threshold: float = 0.03) -> torch.Tensor:
    """[n_patches, 32, 9] targets from a slot-structured fit (row i = slot i)."""
    dev = gaussians.positions.device
    n_slots = math.prod(patch_grid(shape)) * SLOTS
    if gaussians.num_points != n_slots:
        raise ValueError(f"fit has {gaussians.num_points} Gaussians, the slot grid {n_slots}")
    centre = slot_cell_centres(shape, dev)
    half = torch.tensor(CELL, dtype=torch.float32, device=dev) / 2
    off = ((_to_voxel(gaussians.positions.detach(), shape) - centre) / half).clamp(-OFFSET_LIMIT, OFFSET_LIMIT)
    hs = torch.tensor(shape, dtype=torch.float32, device=dev) / 2
    log_sig = (gaussians.log_scales.detach() + torch.log(hs)).clamp(*LOG_SIGMA_RANGE)
    ar = gaussians.density_real.detach() / scale
    ai = gaussians.density_imag.detach() / scale
    exist = (torch.hypot(ar, ai) >= exist_threshold).float()
    tgt = torch.cat([off, log_sig, ar[:, None], ai[:, None], exist[:, None]], dim=1)
    return tgt.view(-1, SLOTS, SLOT_FIELDS)


@torch.no_grad()
def plain_slot_targets(gaussians, shape, scale: float, exist_threshold: float = 0.03) -> torch.Tensor:
    """[n_patches, 32, 9] targets from an unstructured fit (any number of
    Gaussians, any rotations): every Gaussian goes to the cell holding its
    centre, and each cell's Gaussians are merged by moment matching with
    weights |m_i| (m_i = complex amplitude x Gaussian integral): centre
    sum w c / W, per-axis variance sum w (Sigma_i,kk + (c_k - mu_k)^2) / W,
    and the amplitude that gives the merged axis-aligned Gaussian the cell's
    total complex mass sum m_i.  Empty cells are absent slots."""
    from ..tile_voxelizer import _quaternion_to_rotation_matrix

    dev = gaussians.positions.device
    pd, ph, pw = patch_grid(shape)
    n_slots = pd * ph * pw * SLOTS
    c = _to_voxel(gaussians.positions.detach().float(), shape)
    hs = torch.tensor(shape, dtype=torch.float32, device=dev) / 2
    s = torch.exp(gaussians.log_scales.detach().float())
    rot = _quaternion_to_rotation_matrix(gaussians.rotations.detach().float())
    # Voxel-unit covariance diag(hs) R diag(s^2) R^T diag(hs): its diagonal and
    # determinant (the voxelizer's q is u^T cov_inv u with u in half-shape units).
    var = hs * hs * torch.einsum("nik,nk->ni", rot * rot, s * s)
    integral = (2 * math.pi) ** 1.5 * (s * hs).prod(dim=1)
    mr = gaussians.density_real.detach().float() * integral
    mi = gaussians.density_imag.detach().float() * integral
    w = torch.hypot(mr, mi)
    cs = torch.tensor(CELL, dtype=torch.float32, device=dev)
    cells = torch.tensor([pd * GRID[0], ph * GRID[1], pw * GRID[2]], device=dev)
    cell = torch.floor((c + 0.5) / cs).long().clamp(min=0)
    cell = torch.minimum(cell, cells - 1)
    g = torch.tensor(GRID, device=dev)
    patch, local = cell // g, cell % g
    slot = (((patch[:, 0] * ph + patch[:, 1]) * pw + patch[:, 2]) * SLOTS
            + (local[:, 0] * GRID[1] + local[:, 1]) * GRID[2] + local[:, 2])
    ok = torch.isfinite(w) & torch.isfinite(c).all(1) & torch.isfinite(var).all(1) & (w > 0)
    slot, c, var, w, mr, mi = slot[ok], c[ok], var[ok], w[ok], mr[ok], mi[ok]

    def acc(v):
        out = torch.zeros((n_slots,) + v.shape[1:], dtype=torch.float64, device=dev)
        return out.index_add_(0, slot, v.double())

    W = acc(w)
    occ = W > 0
    Wd = torch.where(occ, W, torch.ones_like(W))
    mu = acc(w[:, None] * c) / Wd[:, None]
    second = acc(w[:, None] * (var + c * c)) / Wd[:, None]
    s2 = (second - mu * mu).clamp(min=1e-6)
    sig_m = torch.sqrt(s2)
    amp_scale = 1.0 / ((2 * math.pi) ** 1.5 * sig_m.prod(dim=1) * scale)
    ar = acc(mr) * amp_scale
    ai = acc(mi) * amp_scale
    centre = slot_cell_centres(shape, dev).double()
    half = cs.double() / 2
    off = ((mu - centre) / half).clamp(-OFFSET_LIMIT, OFFSET_LIMIT)
    log_sig = torch.log(sig_m).clamp(*LOG_SIGMA_RANGE)
    exist = (occ & (torch.hypot(ar, ai) >= exist_threshold)).double()
    tgt = torch.cat([off, log_sig, ar[:, None], ai[:, None], exist[:, None]], dim=1).float()
    tgt[~occ, :8] = 0.0
    return tgt.view(-1, SLOTS, SLOT_FIELDS)


def slot_loss(pred: torch.Tensor, tgt: torch.Tensor, w_exist=1.0, beta=0.1):
    """Existence BCE over all slots plus smooth-L1 on the fields of occupied slots."""
    exist = tgt[..., 8]
    bce = torch.nn.functional.binary_cross_entropy_with_logits(pred[..., 8], exist)
    occ = exist > 0.5
    if occ.any():
        reg = torch.nn.functional.smooth_l1_loss(pred[..., :8][occ], tgt[..., :8][occ], beta=beta)
    else:
        reg = pred.sum() * 0.0
    return w_exist * bce + reg, bce.detach(), reg.detach()


@torch.no_grad()
def decode(pred: torch.Tensor, shape, scale: float, threshold=0.5, max_points=None):
    """Slots whose existence probability exceeds `threshold` -> Gaussian
    parameters (positions, log_scales, rotations, density_real, density_imag)."""
    prob = torch.sigmoid(pred[..., 8]).reshape(-1)
    idx = torch.nonzero(prob > threshold, as_tuple=False).squeeze(-1)
    if max_points is not None and idx.numel() > max_points:
        idx = idx[torch.topk(prob[idx], max_points).indices]
        idx = torch.sort(idx).values
    fields = pred.reshape(-1, SLOT_FIELDS)[idx, :8].float()
    return slot_params_to_gaussians(fields, idx, shape, scale)
