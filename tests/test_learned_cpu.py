"""Learned initialiser host logic on the CPU: patch order, slot grid, targets,
decoding, phase augmentation.  No GPU kernels run here."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from gs_tmi_amd.learned import data as ld
from gs_tmi_amd.learned.model import SLOTS
from gs_tmi_amd.recon.model import GaussianSet


def _image(shape, seed=0):
    g = torch.Generator().manual_seed(seed)
    return torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g))


def test_patch_order_and_values():
    img = _image((16, 24, 8))
    x = ld.extract_patches(img, 2.0)
    assert x.shape == (2 * 3 * 1, 1024)
    # patch (pd=1, ph=2, pw=0) is number 1*3 + 2 = 5; its voxel (i, j, k) is img[8+i, 16+j, k].
    p = x[5]
    for (i, j, k) in [(0, 0, 0), (7, 7, 7), (3, 1, 6)]:
        flat = (i * 8 + j) * 8 + k
        v = img[8 + i, 16 + j, k] / 2.0
        assert p[flat] == pytest.approx(float(v.real)) and p[512 + flat] == pytest.approx(float(v.imag))
    with pytest.raises(ValueError):
        ld.extract_patches(_image((12, 8, 8)), 1.0)


def test_slot_cell_centres_cover_every_cell_once():
    shape = (16, 8, 16)
    c = ld.slot_cell_centres(shape, "cpu")
    assert c.shape == (2 * 1 * 2 * SLOTS, 3)
    cell = torch.floor((c + 0.5) / torch.tensor(ld.CELL, dtype=torch.float32)).long()
    flat = (cell[:, 0] * (shape[1] // 2) + cell[:, 1]) * (shape[2] // 4) + cell[:, 2]
    assert torch.unique(flat).numel() == c.shape[0]
    # Slot 0 of patch 1 (pd=0, ph=0, pw=1) is the first cell of the volume's second W patch.
    assert torch.allclose(c[SLOTS], torch.tensor([0.5, 0.5, 8 + 1.5]))


def test_grid_init_targets_decode_round_trip():
    shape = (16, 16, 8)
    img = _image(shape, 1)
    scale = 1.7
    g = GaussianSet(*ld.grid_gaussians(img, scale))
    assert g.num_points == 2 * 2 * 1 * SLOTS
    t = ld.slot_targets(g, shape, scale, exist_threshold=0.0)
    assert torch.allclose(t[..., :3], torch.zeros(()), atol=1e-5)
    assert torch.allclose(t[..., 3:6], torch.log(torch.tensor(ld.CELL, dtype=torch.float32) / 2).expand(t.shape[0], SLOTS, 3),
                          atol=1e-5)
    gain = 16 / ((2 * math.pi) ** 1.5 * 2.0 * 0.970709)  # P(chi2_3 <= 9) = 0.970709
    cell_mean = img.reshape(2, 4, 2, 2, 4, 2, 1, 2, 4).mean(dim=(2, 5, 8)).permute(0, 2, 4, 1, 3, 5).reshape(-1)
    amp = torch.complex(t[..., 6], t[..., 7]).reshape(-1) * scale
    assert torch.allclose(amp, cell_mean * gain, atol=1e-5)
    # Decoding the targets (existence logits set) gives the Gaussian set back.
    pred = t.clone()
    pred[..., 8] = 10.0
    pred[0, 3, 8] = -10.0   # one slot absent
    params = ld.decode(pred, shape, scale)
    keep = torch.ones(g.num_points, dtype=torch.bool)
    keep[3] = False
    for a, b in zip(params, g.param_list()):
        assert torch.allclose(a, b.detach()[keep], atol=1e-5)


def test_targets_of_moved_gaussians():
    shape = (8, 8, 8)
    img = _image(shape, 2)
    g = GaussianSet(*ld.grid_gaussians(img, 1.0))
    with torch.no_grad():
        # Move slot 5 by +0.5 voxel along W (a quarter of its 4-voxel cell = 0.25 half-cells... 0.5/2)
        g.positions[5, 2] += 0.5 / 8 * 2
        g.log_scales[5] += math.log(1.5)
        g.density_real[5] = 0.01
        g.density_imag[5] = 0.0
    t = ld.slot_targets(g, shape, 1.0, exist_threshold=0.03).reshape(-1, 9)
    assert t[5, 2] == pytest.approx(0.25, abs=1e-5)
    assert t[5, 3:6].tolist() == pytest.approx((torch.log(torch.tensor([1.0, 1.0, 2.0]) * 1.5)).tolist(), abs=1e-5)
    assert t[5, 8] == 0.0
    with pytest.raises(ValueError):
        ld.slot_targets(GaussianSet(*[p[:10] for p in g.param_list()]), shape, 1.0)


def test_grid_init_reproduces_a_constant_image():
    """One Gaussian per 2x2x4 cell with sigma (1,1,2) and amplitude gain
    16 / ((2 pi)^1.5 * 2 * P(chi2_3 <= 9)) sums to ~ the constant away from the border: by
    Poisson summation the ripple of Gaussians spaced 2 sigma apart is
    2 exp(-pi^2 / 2) = 1.4% per axis, under 5% over three axes."""
    from oracle import voxelizer_oracle as vo
    from oracle.torch64 import voxelize64

    shape = (24, 24, 24)
    img = torch.full(shape, 0.8 + 0.3j, dtype=torch.complex64)
    g = GaussianSet(*ld.grid_gaussians(img, 1.0))
    pos = g.positions.detach().double()
    ls = g.log_scales.detach().double()
    rot = g.rotations.detach().double()
    center, radius = vo.prologue(pos.float().numpy(), torch.exp(g.log_scales.detach()).numpy(),
                                 rot.float().numpy(), shape)[:2]
    ranges, pids = vo.bin_tiles(center, radius, shape, 8)
    vr, vi = voxelize64(pos, ls, rot, g.density_real.detach().double(), g.density_imag.detach().double(), shape, 8,
                        ranges, pids)
    vol = torch.complex(vr, vi).reshape(shape)[8:16, 8:16, 8:16]
    err = (vol - torch.tensor(0.8 + 0.3j, dtype=torch.complex128)).abs() / abs(0.8 + 0.3j)
    assert 0.01 < float(err.max()) < 0.05
    assert float(err.mean()) < 0.025


def test_phase_rotation_is_consistent():
    from gs_tmi_amd.learned.training import _rotate_phase

    x = torch.randn(4, 1024)
    t = torch.randn(4, SLOTS, 9)
    gen = torch.Generator().manual_seed(0)
    xo, to = _rotate_phase(x, t, gen)
    gen = torch.Generator().manual_seed(0)
    phi = torch.rand(4, 1, generator=gen) * 2 * math.pi
    z = torch.complex(x[:, :512], x[:, 512:]) * torch.exp(1j * phi)
    assert torch.allclose(xo[:, :512], z.real, atol=1e-5) and torch.allclose(xo[:, 512:], z.imag, atol=1e-5)
    a = torch.complex(t[..., 6], t[..., 7]) * torch.exp(1j * phi)
    assert torch.allclose(to[..., 6], a.real, atol=1e-5) and torch.allclose(to[..., 7], a.imag, atol=1e-5)
    assert torch.equal(to[..., :6], t[..., :6]) and torch.equal(to[..., 8], t[..., 8])
    # Magnitudes are unchanged.
    assert np.allclose(z.abs().numpy(), torch.complex(x[:, :512], x[:, 512:]).abs().numpy(), atol=1e-5)


def _set(pos_vox, sig_vox, amp, shape, rot=None):
    """GaussianSet from voxel centres / sigmas (normalised parameters as the model keeps them)."""
    shp = torch.tensor(shape, dtype=torch.float32)
    pos = (torch.tensor(pos_vox, dtype=torch.float32) + 0.5) / shp * 2 - 1
    log_s = torch.log(torch.tensor(sig_vox, dtype=torch.float32) / (shp / 2))
    n = pos.shape[0]
    if rot is None:
        rot = torch.zeros(n, 4)
        rot[:, 0] = 1.0
    a = torch.tensor(amp, dtype=torch.complex64)
    return GaussianSet(pos, log_s, torch.as_tensor(rot, dtype=torch.float32), a.real.clone(), a.imag.clone())


def test_plain_targets_single_gaussian_is_itself():
    """An axis-aligned Gaussian alone in its cell becomes that cell's slot unchanged
    (offset, sigma, amplitude); every other slot is absent."""
    shape = (8, 8, 16)
    scale = 2.0
    g = _set([[2.3, 5.1, 9.7]], [[0.9, 1.4, 2.2]], [1.5 - 0.5j], shape)
    t = ld.plain_slot_targets(g, shape, scale)
    # Cell (1, 2, 2) of patch (0, 0, 1): slot 32 + (1 * 4 + 2) * 2 + 0 (cell W index 2 -> patch 1, local 0).
    s = SLOTS + (1 * 4 + 2) * 2 + 0
    flat = t.reshape(-1, 9)
    centre = ld.slot_cell_centres(shape, "cpu")[s]
    assert torch.allclose(flat[s, :3], (torch.tensor([2.3, 5.1, 9.7]) - centre) / (torch.tensor(ld.CELL) / 2),
                          atol=1e-5)
    assert torch.allclose(flat[s, 3:6], torch.log(torch.tensor([0.9, 1.4, 2.2])), atol=1e-5)
    assert torch.allclose(flat[s, 6:8], torch.tensor([0.75, -0.25]), atol=1e-5)
    assert flat[s, 8] == 1.0 and float(flat[:, 8].sum()) == 1.0


def test_plain_targets_merge_conserves_mass_and_moments():
    """Two Gaussians in one cell (one rotated, amplitudes of different phase):
    the merged slot carries their total complex mass, the |mass|-weighted
    centre and the diagonal of the weighted second moment, each Gaussian's
    own covariance from its rotation in the voxel frame of an anisotropic
    volume."""
    shape = (8, 16, 8)
    scale = 1.0
    q = torch.tensor([[1.0, 0.0, 0.0, 0.0], [0.9, 0.3, -0.2, 0.25]])
    pos = [[2.1, 2.4, 4.9], [2.8, 3.3, 5.6]]
    sig = [[0.6, 0.8, 1.0], [0.5, 0.7, 1.3]]
    amp = [1.0 + 0.5j, -0.3 + 0.8j]
    g = _set(pos, sig, amp, shape, rot=q)
    flat = ld.plain_slot_targets(g, shape, scale, exist_threshold=0.0).reshape(-1, 9)
    occ = torch.nonzero(flat[:, 8]).squeeze(-1)
    assert occ.numel() == 1
    s = int(occ)
    # Reference moments in float64.
    from gs_tmi_amd.tile_voxelizer import _quaternion_to_rotation_matrix
    hs = torch.tensor(shape, dtype=torch.float64) / 2
    sd = torch.tensor(sig, dtype=torch.float64) / hs  # normalised scales
    R = _quaternion_to_rotation_matrix(q.double())
    cov = torch.diag_embed(hs).unsqueeze(0) @ R @ torch.diag_embed(sd * sd) @ R.transpose(1, 2) @ torch.diag_embed(hs)
    mass = torch.tensor(amp, dtype=torch.complex128) * (2 * math.pi) ** 1.5 * torch.sqrt(torch.linalg.det(cov))
    w = mass.abs()
    c = torch.tensor(pos, dtype=torch.float64)
    mu = (w[:, None] * c).sum(0) / w.sum()
    var = (w[:, None] * (torch.diagonal(cov, dim1=1, dim2=2) + (c - mu) ** 2)).sum(0) / w.sum()
    centre = ld.slot_cell_centres(shape, "cpu")[s].double()
    assert torch.allclose(flat[s, :3].double(), (mu - centre) / (torch.tensor(ld.CELL) / 2), atol=1e-4)
    assert torch.allclose(flat[s, 3:6].double(), 0.5 * torch.log(var), atol=1e-4)
    merged_mass = torch.complex(flat[s, 6].double(), flat[s, 7].double()) * (2 * math.pi) ** 1.5 * torch.sqrt(var).prod()
    assert abs(complex(merged_mass) - complex(mass.sum())) < 1e-4 * float(w.sum())
