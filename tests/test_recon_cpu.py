"""Reconstruction harness on CPU: transforms, masks, coils, metrics, Gaussian set
updates, and a few Adam steps through the reference's CPU voxelizer path."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from gs_tmi_amd import recon
from oracle.cpu_tile_torch import CpuTileVoxelizer


def test_fft3c_is_orthonormal_and_centred():
    g = torch.Generator().manual_seed(0)
    x = torch.complex(torch.randn(2, 8, 6, 10, generator=g), torch.randn(2, 8, 6, 10, generator=g))
    k = recon.fft3c(x)
    assert torch.allclose(recon.ifft3c(k), x, atol=1e-5)
    assert math.isclose(float((k.abs() ** 2).sum()), float((x.abs() ** 2).sum()), rel_tol=1e-5)
    # A centred delta transforms to a constant; a constant to a centred delta.
    d = torch.zeros(8, 6, 10, dtype=torch.complex64)
    d[4, 3, 5] = 1.0
    kd = recon.fft3c(d)
    assert torch.allclose(kd.abs(), torch.full_like(kd.abs(), 1 / math.sqrt(480)), atol=1e-6)
    assert torch.allclose(kd, kd[0, 0, 0].expand_as(kd), atol=1e-6)


@pytest.mark.parametrize("shape,accel,acs", [((64, 64), 4, 16), ((32, 48), 8, 8)])
def test_variable_density_mask(shape, accel, acs):
    m = recon.variable_density_mask(shape, accel, acs, seed=1)
    assert m.sum() == round(shape[0] * shape[1] / accel)
    cy, cz = shape[0] // 2, shape[1] // 2
    assert m[cy - acs // 2:cy + acs // 2, cz - acs // 2:cz + acs // 2].all()
    # Density falls off from the centre.
    yy, zz = np.meshgrid(np.arange(shape[0]) - cy, np.arange(shape[1]) - cz, indexing="ij")
    r = np.sqrt(yy ** 2 + zz ** 2)
    assert m[r < 12].mean() > m[r > 20].mean()


def test_poisson_disc_mask():
    shape, accel, calib = (64, 48), 8, 12
    m = recon.poisson_disc_mask(shape, accel, calib, seed=3)
    assert m.sum() == int(shape[0] * shape[1] / accel)
    cy, cz = shape[0] // 2, shape[1] // 2
    assert m[cy - 6:cy + 6, cz - 6:cz + 6].all()
    assert not np.array_equal(m, recon.poisson_disc_mask(shape, accel, calib, seed=4))
    full = recon.expand_phase_mask(m, (5,) + shape)
    assert np.all(full.sum(axis=0) % 5 == 0)  # the readout axis is fully sampled


def test_coils_rss_and_metrics():
    s = recon.gaussian_coils((8, 8, 8), 6)
    assert np.allclose(np.sqrt((np.abs(s) ** 2).sum(0)), 1.0, atol=1e-5)
    t = torch.tensor([0.0, 1.0, 2.0, 3.0])
    p = t + torch.tensor([0.1, -0.1, 0.1, -0.1])
    assert recon.psnr(p, t) == pytest.approx(10 * math.log10(9 / 0.01))
    assert recon.nmse(p, t) == pytest.approx(0.04 / 14)


def test_gaussian_set_split_clone_prune_row_maps():
    img = torch.complex(torch.rand(8, 8, 8), torch.rand(8, 8, 8))
    g = recon.GaussianSet.from_image(img, 50, mode="random", scale_vox=1.5, seed=0)
    pos0 = g.positions.detach().clone()
    gn = torch.zeros(50)
    gn[[3, 7]] = 1.0
    n, rm = g.densify_and_split(gn, 0.5, 1e-4)
    # each split parent is replaced by two children
    assert n == 2 and g.num_points == 52 and (rm == -1).sum() == 4
    kept = rm[rm >= 0]
    assert torch.equal(g.positions.detach()[:48], pos0[kept])
    # children sit symmetrically about the parent, scales shrunk
    assert torch.allclose(g.positions.detach()[48:50] + g.positions.detach()[50:52], 2 * pos0[[3, 7]], atol=1e-6)
    n, rm = g.densify_and_clone(torch.cat([torch.ones(1), torch.zeros(51)]), 0.5, 10.0)
    assert n == 1 and g.num_points == 53 and rm[-1] == -1
    keep = torch.ones(53, dtype=torch.bool)
    keep[0] = False
    n, rm = g.prune(keep)
    assert n == 1 and g.num_points == 52 and rm[0] == 1


def test_adam_steps_through_cpu_voxelizer_reduce_loss():
    shape = (16, 16, 16)
    p = recon.make_problem(shape, 6, 4, recon.variable_density_mask((16, 16), 3, 6, 0), seed=2, noise_sigma=1e-3)
    g = recon.GaussianSet.from_image(p.zero_filled, 300, seed=0)
    r = recon.Reconstructor(p, g, CpuTileVoxelizer(shape), recon.ReconConfig.single_lr(1e-2))
    losses = [float(r.step()) for _ in range(8)]
    assert losses[-1] < 0.7 * losses[0]
    m = r.evaluate()
    assert m["psnr"] > 15 and m["nmse"] < 1.0


def test_config1_cpu_path_fixture_is_reproducible():
    """The committed configuration-1 run of the reference's CPU path
    (tests/golden/config1_cpu_path.json, against which the GPU test checks the
    HIP path's PSNR) is what its script computes: its first two steps,
    evaluated after each, are recomputed here."""
    import json
    from pathlib import Path

    from gs_tmi_amd.recon import Reconstructor
    from gs_tmi_amd.recon.configs import config1_gaussians, config1_problem, config1_recon_config
    from oracle.cpu_tile_torch import CpuTileVoxelizer

    golden = json.loads((Path(__file__).resolve().parent / "golden" / "config1_cpu_path.json").read_text())
    p = config1_problem("cpu")
    g = config1_gaussians(p)
    rows = Reconstructor(p, g, CpuTileVoxelizer(p.shape), config1_recon_config()).run(2, eval_every=1).rows
    assert golden["history"][0]["psnr"] == pytest.approx(rows[0]["psnr"], abs=1e-6)
    for a, b in zip(rows, golden["pin_two_steps"]):
        assert a["step"] == b["step"]
        assert a["psnr"] == pytest.approx(b["psnr"], abs=1e-3)
        assert a["nmse"] == pytest.approx(b["nmse"], rel=1e-3)
        if b["loss"] is not None:
            assert a["loss"] == pytest.approx(b["loss"], rel=1e-4)
