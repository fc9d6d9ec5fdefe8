"""gstv_backward_noalloc (the adjoint without a workspace) against gstv_backward
(GPU tests): two calls in a row through the C ABI, the second with fewer
Gaussians, each equal to gstv_backward's result up to the order of fp32 atomic
additions; sizes above the limit are refused with a clear error."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from tests import cases

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _inputs(c):
    from gs_tmi_amd import build_tile_pairs, voxel_space_params

    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    center, radius, cov, hs = voxel_space_params(t(c["pos"]), t(c["scales"]), t(c["rot"]), c["shape"])
    ranges, pids = build_tile_pairs(center, radius, c["shape"], c["tile"])
    return center, cov.contiguous(), t(c["rr"]), t(c["ri"]), ranges, pids, hs


def test_backward_noalloc_matches_backward_on_two_calls():
    from gs_tmi_amd import _native

    runs = [("64^3, 20000 Gaussians", cases.microbench_case(n=20_000, shape=(64, 64, 64), seed=4)),
            ("partial tiles, 300 Gaussians", cases.edge_case("partial_tiles"))]
    for label, c in runs:
        d, h, w = c["shape"]
        print(f"[case] backward_noalloc: {label}, shape={c['shape']} tile={c['tile']}", flush=True)
        center, cov, rr, ri, ranges, pids, hs = _inputs(c)
        rng = np.random.default_rng(31)
        g_r = torch.from_numpy(rng.normal(size=d * h * w).astype(np.float32)).to(DEV)
        g_i = torch.from_numpy(rng.normal(size=d * h * w).astype(np.float32)).to(DEV)
        want = _native.backward(center, cov, rr, ri, ranges, pids, hs, g_r, g_i, c["tile"], d, h, w)
        got = _native.backward_noalloc(center, cov, rr, ri, ranges, pids, hs, g_r, g_i, c["tile"], d, h, w)
        torch.cuda.synchronize()
        for nm, a, b in zip(["center", "cov_inv", "rho_r", "rho_i"], got, want):
            assert a.shape == b.shape
            e = cases.rel_l2(a.cpu().numpy(), b.cpu().numpy())
            print(f"  d/d{nm}: no-allocation vs workspace rel L2 {e:.2e}", flush=True)
            assert np.isfinite(a.cpu().numpy()).all()
            assert e <= 1e-6, (label, nm, e)


def test_backward_noalloc_refuses_more_gaussians_than_its_array_holds():
    from gs_tmi_amd import _native

    n = _native.NOALLOC_MAX_GAUSSIANS + 1
    print(f"[case] backward_noalloc with {n} Gaussians (refused before any launch)", flush=True)
    z = torch.zeros(n, 3, device=DEV)
    cov = torch.zeros(n, 9, device=DEV)
    rho = torch.zeros(n, device=DEV)
    ranges = torch.zeros(8, 2, dtype=torch.int32, device=DEV)
    pids = torch.zeros(1, dtype=torch.int32, device=DEV)
    hs = torch.full((3,), 8.0, device=DEV)
    g = torch.zeros(16 * 16 * 16, device=DEV)
    with pytest.raises(RuntimeError, match="too many Gaussians"):
        _native.backward_noalloc(z, cov, rho, rho, ranges, pids, hs, g, g, 8, 16, 16, 16)
