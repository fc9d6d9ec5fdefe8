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
