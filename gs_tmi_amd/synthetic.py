"""Seeded synthetic Gaussian sets (workload generation for tests and benchmarks).

Parameters follow the reference's parameterisation (gaussian/gaussian_model.py:38-61,
gaussian/tile_voxelizer.py:279): positions in [-1, 1] normalised units (p = -1
at voxel -0.5, p = 1 at voxel size - 0.5), linear scales in units of the half
shape, unit quaternions in (w, x, y, z) order, and a complex amplitude as two
fp32 planes.
"""

from __future__ import annotations

import math

import numpy as np

F32 = np.float32


def make_gaussians(n, shape, seed, sigma_median=1.5, sigma_log_sd=0.5, sigma_mode="lognormal", means="uniform"):
    """Seeded synthetic Gaussians in the reference's parameterisation.

    Returns (positions f32 [N,3] in [-1,1], scales f32 [N,3] linear, rotations
    f32 [N,4] unit (w,x,y,z), rho_r f32 [N], rho_i f32 [N]).
    sigma_mode "lognormal": per-axis voxel sigma ~ LogNormal(ln sigma_median, sigma_log_sd);
    "lognormal0": voxel log-sigma ~ N(0, sigma_log_sd) (the adjoint-check reading).
    Drawn in float64 with numpy default_rng(seed), cast to fp32 once.
    """
    rng = np.random.default_rng(seed)
    shape_f = np.asarray(shape, dtype=np.float64)
    hs = shape_f / 2.0
    if means == "uniform":
        pos = rng.uniform(-1.0, 1.0, size=(n, 3))
    else:  # voxel coordinates uniform in [0, size]
        c = rng.uniform(0.0, 1.0, size=(n, 3)) * shape_f
        pos = (c + 0.5) / hs - 1.0
    if sigma_mode == "lognormal":
        s_vox = np.exp(rng.normal(math.log(sigma_median), sigma_log_sd, size=(n, 3)))
    elif sigma_mode == "lognormal0":
        s_vox = np.exp(rng.normal(0.0, sigma_log_sd, size=(n, 3)))
    else:
        raise ValueError(sigma_mode)
    scales = s_vox / hs
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    rho_r = rng.normal(size=n)
    rho_i = rng.normal(size=n)
    return (pos.astype(F32), scales.astype(F32), q.astype(F32), rho_r.astype(F32), rho_i.astype(F32))
