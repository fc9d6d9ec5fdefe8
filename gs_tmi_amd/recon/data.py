"""Synthetic MRI problems: phantoms, coil maps, sampling masks and the
centred FFT pair (data/transforms.py:11-55, data/dataset.py:232-250,
data/dataset.py:438-537 and data/mask_generator.py:22-400 of the reference)."""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch


def fft3c(image: torch.Tensor) -> torch.Tensor:
    """Centred orthonormal 3D FFT over the last three axes: ifftshift -> fftn -> fftshift
    (data/transforms.py:11-32)."""
    dims = (-3, -2, -1)
    return torch.fft.fftshift(torch.fft.fftn(torch.fft.ifftshift(image, dim=dims), dim=dims, norm="ortho"), dim=dims)


def ifft3c(kspace: torch.Tensor) -> torch.Tensor:
    """Centred orthonormal 3D inverse FFT (data/transforms.py:35-55)."""
    dims = (-3, -2, -1)
    return torch.fft.fftshift(torch.fft.ifftn(torch.fft.ifftshift(kspace, dim=dims), dim=dims, norm="ortho"), dim=dims)


def _grid(shape):
    axes = [np.linspace(-1.0, 1.0, s) for s in shape]
    return np.meshgrid(*axes, indexing="ij")


def ellipsoid_phantom(shape, n_ellipsoids, seed, mag_range=(0.2, 1.0)) -> np.ndarray:
    """Complex phantom: a sum of randomly rotated ellipsoids with magnitudes
    U[mag_range], times a smooth linear phase.  Returns complex64 [D, H, W]."""
    rng = np.random.default_rng(seed)
    z, y, x = _grid(shape)
    pts = np.stack([z, y, x], axis=-1).reshape(-1, 3)
    mag = np.zeros(pts.shape[0])
    for _ in range(n_ellipsoids):
        centre = rng.uniform(-0.55, 0.55, size=3)
        axes = rng.uniform(0.06, 0.35, size=3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, a, b, c = q
        rot = np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - c * w), 2 * (a * c + b * w)],
                        [2 * (a * b + c * w), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w)],
                        [2 * (a * c - b * w), 2 * (b * c + a * w), 1 - 2 * (a * a + b * b)]])
        local = (pts - centre) @ rot
        inside = np.sum((local / axes) ** 2, axis=1) <= 1.0
        mag[inside] += rng.uniform(*mag_range)
    mag = np.minimum(mag, 1.5).reshape(shape)
    g = rng.uniform(-math.pi / 2, math.pi / 2, size=3)
    phase = g[0] * z + g[1] * y + g[2] * x + rng.uniform(-math.pi, math.pi)
    return (mag * np.exp(1j * phase)).astype(np.complex64)


def ellipsoid_phantom_torch(shape, n_ellipsoids, seed, mag_range=(0.2, 1.0), device="cpu") -> torch.Tensor:
    """``ellipsoid_phantom`` evaluated with torch (float64) on `device` from the
    same random draws; the inside test can differ from the NumPy path only for
    points within round-off of an ellipsoid surface.  Returns complex64 [D, H, W]."""
    rng = np.random.default_rng(seed)
    axes1d = [torch.linspace(-1.0, 1.0, s, dtype=torch.float64, device=device) for s in shape]
    z, y, x = torch.meshgrid(*axes1d, indexing="ij")
    pts = torch.stack([z, y, x], dim=-1).reshape(-1, 3)
    mag = torch.zeros(pts.shape[0], dtype=torch.float64, device=device)
    for _ in range(n_ellipsoids):
        centre = rng.uniform(-0.55, 0.55, size=3)
        axes = rng.uniform(0.06, 0.35, size=3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, a, b, c = q
        rot = np.array([[1 - 2 * (b * b + c * c), 2 * (a * b - c * w), 2 * (a * c + b * w)],
                        [2 * (a * b + c * w), 1 - 2 * (a * a + c * c), 2 * (b * c - a * w)],
                        [2 * (a * c - b * w), 2 * (b * c + a * w), 1 - 2 * (a * a + b * b)]])
        local = (pts - torch.from_numpy(centre).to(device)) @ torch.from_numpy(rot).to(device)
        inside = torch.sum((local / torch.from_numpy(axes).to(device)) ** 2, dim=1) <= 1.0
        mag[inside] += rng.uniform(*mag_range)
    mag = torch.clamp(mag, max=1.5).reshape(shape)
    g = rng.uniform(-math.pi / 2, math.pi / 2, size=3)
    phase = g[0] * z + g[1] * y + g[2] * x + rng.uniform(-math.pi, math.pi)
    return torch.polar(mag, phase).to(torch.complex64)


def gaussian_coils(shape, n_coils, radius=1.3, width=0.8) -> np.ndarray:
    """Coil maps with Gaussian magnitude profiles centred on a sphere of the
    given radius (Fibonacci points), a smooth per-coil phase, normalised to
    root-sum-of-squares 1 (the convention of data/dataset.py:301-302).
    Returns complex64 [C, D, H, W]."""
    z, y, x = _grid(shape)
    maps = []
    golden = math.pi * (3.0 - math.sqrt(5.0))
    for c in range(n_coils):
        t = 1.0 - 2.0 * (c + 0.5) / n_coils
        r = math.sqrt(max(0.0, 1.0 - t * t))
        cz, cy, cx = radius * t, radius * r * math.cos(golden * c), radius * r * math.sin(golden * c)
        d2 = (z - cz) ** 2 + (y - cy) ** 2 + (x - cx) ** 2
        phase = 0.5 * (cz * z + cy * y + cx * x) + 2.0 * math.pi * c / n_coils
        maps.append(np.exp(-d2 / (2 * width ** 2)) * np.exp(1j * phase))
    maps = np.stack(maps)
    rss = np.sqrt(np.sum(np.abs(maps) ** 2, axis=0)) + 1e-12
    return (maps / rss).astype(np.complex64)


def _acs_box(phase_shape, acs):
    (ny, nz), h = phase_shape, acs // 2
    cy, cz = ny // 2, nz // 2
    return slice(max(0, cy - h), min(ny, cy + h)), slice(max(0, cz - h), min(nz, cz + h))


def variable_density_mask(phase_shape, accel, acs, seed) -> np.ndarray:
    """2D variable-density random mask over the phase-encode plane: a fully
    sampled acs x acs centre plus points drawn without replacement with
    weights exp(-r^2 / 2 sigma^2), sigma = min(shape) / 4, to a total of
    round(size / accel) (data/dataset.py:443-474).  Returns float32 [ny, nz]."""
    rng = np.random.default_rng(seed)
    ny, nz = phase_shape
    target = max(int(round(ny * nz / float(accel))), 1)
    a = np.arange(ny) - ny // 2
    b = np.arange(nz) - nz // 2
    ga, gb = np.meshgrid(a, b, indexing="ij")
    sigma = min(phase_shape) / 4.0
    weights = np.exp(-(ga ** 2 + gb ** 2) / (2 * sigma ** 2))
    mask = np.zeros(phase_shape, dtype=np.float32)
    sy, sz = _acs_box(phase_shape, acs)
    mask[sy, sz] = 1.0
    cand = np.flatnonzero(mask.ravel() == 0)
    remaining = min(max(target - int(mask.sum()), 0), cand.size)
    if remaining > 0:
        w = weights.ravel()[cand]
        chosen = rng.choice(cand, size=remaining, replace=False, p=w / w.sum())
        mask.ravel()[chosen] = 1.0
    return mask


def poisson_disc_mask(phase_shape, accel, calib, seed, p=2.0, damping=0.6) -> np.ndarray:
    """Variable-density Poisson-disc mask by dart throwing, as the reference's
    heuristic (data/mask_generator.py:22-63, 241-400): pdf = (1 - r)^p with a
    fully sampled calib x calib centre, scaled to the target count and clipped
    to [0, 1]; candidates in one random permutation are accepted when no
    accepted point lies within damping / sqrt(pdf) of them, until
    size / accel points are set.  Returns float32 [ny, nz]."""
    rng = np.random.default_rng(seed)
    ny, nz = phase_shape
    yy, zz = np.meshgrid(np.linspace(-1, 1, ny), np.linspace(-1, 1, nz), indexing="ij")
    r = np.sqrt(yy ** 2 + zz ** 2)
    pdf = np.power(np.maximum(0.0, 1.0 - r), p)
    sy, sz = _acs_box(phase_shape, calib)
    pdf[sy, sz] = 1.0
    target = int(ny * nz / accel)
    pdf = np.clip(pdf / pdf.sum() * target, 0.0, 1.0)
    r_grid = damping / np.sqrt(np.maximum(pdf, 1e-4))
    mask = np.zeros(phase_shape, dtype=np.int8)
    mask[sy, sz] = 1
    count = int(mask.sum())
    order = rng.permutation(ny * nz)
    for idx in order:
        if count >= target:
            break
        y, z = divmod(int(idx), nz)
        if mask[y, z]:
            continue
        rad = r_grid[y, z]
        ir = int(math.ceil(rad))
        y0, y1, z0, z1 = max(0, y - ir), min(ny, y + ir + 1), max(0, z - ir), min(nz, z + ir + 1)
        sub = mask[y0:y1, z0:z1]
        if sub.any():
            ey, ez = np.nonzero(sub)
            if np.any((ey + y0 - y) ** 2 + (ez + z0 - z) ** 2 < rad * rad):
                continue
        mask[y, z] = 1
        count += 1
    return mask.astype(np.float32)


def expand_phase_mask(mask2d: np.ndarray, shape, readout_axis: int = 0) -> np.ndarray:
    """Broadcast a phase-encode mask along the fully sampled readout axis."""
    phase_axes = [a for a in range(3) if a != readout_axis]
    view = [1, 1, 1]
    view[phase_axes[0]], view[phase_axes[1]] = mask2d.shape
    return np.broadcast_to(mask2d.reshape(view), shape).astype(np.float32)


@dataclass
class MRIProblem:
    """One simulated multi-coil acquisition (tensors on one device)."""
    target: torch.Tensor          # complex64 [D, H, W] ground truth
    coils: torch.Tensor           # complex64 [C, D, H, W]
    mask: torch.Tensor            # float32 [D, H, W]
    kspace: torch.Tensor          # complex64 [C, D, H, W], undersampled (zero where not sampled)
    zero_filled: torch.Tensor     # complex64 [D, H, W], sum_c conj(S_c) ifft3c(y_c)

    @property
    def shape(self):
        return tuple(self.target.shape)


def make_problem(shape, n_ellipsoids, n_coils, mask2d, seed, noise_sigma=0.0, device="cpu",
                 phantom_on_device=False, coils=None) -> MRIProblem:
    """Simulates y_c = M * (fft3c(S_c x) + n_c) and the zero-filled SENSE adjoint
    (data/dataset.py:232-250, 527-537), with the known coil maps used for
    reconstruction.  `phantom_on_device` evaluates the phantom with torch on
    `device`; `coils` reuses precomputed maps of this shape."""
    if phantom_on_device:
        target = ellipsoid_phantom_torch(shape, n_ellipsoids, seed, device=device)
    else:
        target = torch.from_numpy(ellipsoid_phantom(shape, n_ellipsoids, seed)).to(device)
    if coils is None:
        coils = torch.from_numpy(gaussian_coils(shape, n_coils)).to(device)
    mask = torch.from_numpy(expand_phase_mask(mask2d, shape)).to(device)
    k_full = fft3c(coils * target.unsqueeze(0))
    if noise_sigma > 0:
        gen = torch.Generator(device="cpu").manual_seed(seed + 7919)
        noise = torch.complex(torch.randn(k_full.shape, generator=gen), torch.randn(k_full.shape, generator=gen))
        k_full = k_full + noise.to(device) * noise_sigma
    kspace = k_full * mask.unsqueeze(0)
    zero_filled = torch.sum(torch.conj(coils) * ifft3c(kspace), dim=0)
    return MRIProblem(target=target, coils=coils, mask=mask, kspace=kspace, zero_filled=zero_filled)
