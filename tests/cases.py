"""Seeded input cases shared by the parity tests (numpy, fp32)."""

from __future__ import annotations

import math

import numpy as np

from gs_tmi_amd.synthetic import make_gaussians
from oracle import voxelizer_oracle as vo


def adjoint_check_case(seed=0, reading="voxel"):
    """32^3 volume, 512 Gaussians, means U[0,32]^3 voxels, random unit
    quaternions, complex amplitudes N(0,1)+iN(0,1).

    "log-scales N(0,0.3)" is read two ways: "voxel" (primary) takes the voxel
    sigma as exp(N(0, 0.3)); "literal" takes the reference-unit scale as
    exp(N(0, 0.3)), i.e. a voxel sigma of about 16, so every Gaussian covers
    every tile.
    """
    shape = (32, 32, 32)
    n = 512
    if reading == "voxel":
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_mode="lognormal0", sigma_log_sd=0.3,
                                                     means="voxel")
    else:
        rng = np.random.default_rng(seed)
        c = rng.uniform(0.0, 32.0, size=(n, 3))
        pos = ((c + 0.5) / 16.0 - 1.0).astype(np.float32)
        scales = np.exp(rng.normal(0.0, 0.3, size=(n, 3))).astype(np.float32)
        q = rng.normal(size=(n, 4))
        rot = (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)
        rr = rng.normal(size=n).astype(np.float32)
        ri = rng.normal(size=n).astype(np.float32)
    return dict(shape=shape, tile=8, pos=pos, scales=scales, rot=rot, rr=rr, ri=ri)


def microbench_case(n=1_000_000, shape=(256, 256, 128), seed=0):
    """Voxelizer microbenchmark: means uniform in the volume, log-normal voxel
    sigma (median 1.5, sigma 0.5), random rotations, N(0,1)+iN(0,1) amplitudes."""
    pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=1.5, sigma_log_sd=0.5)
    return dict(shape=shape, tile=8, pos=pos, scales=scales, rot=rot, rr=rr, ri=ri)


def edge_case(kind, seed=0):
    rng = np.random.default_rng(1000 + seed)
    if kind == "partial_tiles":
        shape, n, tile = (37, 29, 50), 300, 8
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=1.5, sigma_log_sd=0.5)
    elif kind == "outside":
        shape, n, tile = (24, 40, 16), 300, 8
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=2.0, sigma_log_sd=0.5)
        pos = (pos * 1.6).astype(np.float32)  # a third of the centres fall outside [-1, 1]
    elif kind.startswith("tile"):
        tile = int(kind[4:])
        shape, n = (23, 18, 31), 200
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=1.2, sigma_log_sd=0.6)
    elif kind == "needles":
        shape, n, tile = (32, 32, 32), 256, 8
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=1.5, sigma_log_sd=0.3)
        hs = np.asarray(shape, np.float64) / 2
        s_vox = np.stack([rng.uniform(0.3, 0.6, n), rng.uniform(0.3, 0.6, n), rng.uniform(4.0, 9.0, n)], axis=1)
        scales = (s_vox / hs).astype(np.float32)
    elif kind == "large_sigma":
        shape, n, tile = (32, 32, 32), 64, 8
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=12.0, sigma_log_sd=0.3)
    elif kind == "tiny_sigma":
        shape, n, tile = (16, 16, 16), 400, 8
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=0.2, sigma_log_sd=0.3)
    elif kind == "anisotropic_volume":
        shape, n, tile = (64, 32, 16), 800, 8
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=1.5, sigma_log_sd=0.6)
    elif kind == "extreme_aniso":
        # Needles and pancakes with axis ratios up to several hundred: the
        # quadratic form's determinant cancels strongly, which the column
        # bound of the tile kernels must survive.
        shape, n, tile = (32, 32, 32), 200, 8
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=1.5, sigma_log_sd=0.3)
        hs = np.asarray(shape, np.float64) / 2
        s_vox = np.minimum(np.exp(rng.uniform(np.log(0.05), np.log(30.0), size=(n, 3))), 30.0)
        scales = (s_vox / hs).astype(np.float32)
    elif kind == "crowded":
        # About 1,400 Gaussians per tile: the tile kernels stage more than one
        # batch per tile and two pairs per thread.
        shape, n, tile = (16, 16, 16), 1500, 8
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=3.0, sigma_log_sd=0.3)
    elif kind == "single":
        shape, n, tile = (16, 16, 16), 1, 8
        pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=2.0, sigma_log_sd=0.2)
    else:
        raise ValueError(kind)
    return dict(shape=shape, tile=tile, pos=pos, scales=scales, rot=rot, rr=rr, ri=ri)


def random_case(seed):
    """Randomised small case: volume sides 1-40 voxels, tile 1-16, 1-300
    Gaussians with log-normal voxel sigma (median 0.3-4, log-sd 0.2-0.8), a
    fifth of the centres pushed outside the volume."""
    rng = np.random.default_rng(5000 + seed)
    shape = tuple(int(v) for v in rng.integers(1, 41, size=3))
    tile = int(rng.integers(1, 17))
    n = int(rng.integers(1, 301))
    pos, scales, rot, rr, ri = make_gaussians(n, shape, seed, sigma_median=float(rng.uniform(0.3, 4.0)),
                                              sigma_log_sd=float(rng.uniform(0.2, 0.8)))
    out = rng.random(n) < 0.2
    pos[out] = (pos[out] * 1.5).astype(np.float32)
    return dict(shape=shape, tile=tile, pos=pos, scales=scales, rot=rot, rr=rr, ri=ri)


EDGE_KINDS = ["partial_tiles", "outside", "tile1", "tile3", "tile5", "tile12", "tile16", "tile17", "tile24", "tile32",
              "needles", "extreme_aniso", "crowded",
              "large_sigma", "tiny_sigma", "anisotropic_volume", "single"]


def rel_l2(a, b, mask=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is not None:
        a = a[mask]
        b = b[mask]
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def tile_lists(ranges, pair_ids):
    return [tuple(pair_ids[s:e]) for s, e in np.asarray(ranges)]


def num_tiles(shape, tile):
    return math.prod(math.ceil(s / tile) for s in shape)


def case_by_name(name):
    """The named parity case: adjoint_voxel / adjoint_literal (seeds 0 / 1),
    an edge kind, or random<i>."""
    if name == "adjoint_voxel":
        return adjoint_check_case(0, "voxel")
    if name == "adjoint_literal":
        return adjoint_check_case(1, "literal")
    if name.startswith("random"):
        return random_case(int(name[6:]))
    return edge_case(name)


# Cases of the Function-level tests, and of the full-call (fused) tests.
FUNCTION_CASES = ["adjoint_voxel", "adjoint_literal"] + EDGE_KINDS + [f"random{i}" for i in range(12)]
CALL_CASES = ["adjoint_voxel"] + EDGE_KINDS + [f"random{i}" for i in range(4)]


def function_upstream(shape):
    """Upstream gradient of the Function-level tests (flat planes)."""
    n = int(np.prod(shape))
    rng = np.random.default_rng(7)
    return rng.normal(size=n).astype(np.float32), rng.normal(size=n).astype(np.float32)


def call_upstream(shape, seed=4):
    """Upstream gradient of the full-call tests ([D, H, W] planes)."""
    rng = np.random.default_rng(seed)
    return rng.normal(size=shape).astype(np.float32), rng.normal(size=shape).astype(np.float32)
