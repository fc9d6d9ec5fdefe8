"""End-to-end TileVoxelizer autograd and full-size properties (GPU tests)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import voxelizer_oracle as vo
from tests import cases
from tests import reference_fp32 as rf

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _announce(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("reading", ["voxel", "literal"])
def test_tile_voxelizer_param_gradients_vs_fp64(reading):
    """Gradients w.r.t. positions, log-scales, quaternions and complex amplitude
    through the full call, against FP64 torch autograd of the same call, within
    1e-5 relative L2 (or 1.5x the reference's own fp32 error on the case, where
    that exceeds 1e-5; tests/golden/reference_fp32_errors.json)."""
    from gs_tmi_amd import TileVoxelizer

    c = cases.adjoint_check_case(0 if reading == "voxel" else 1, reading)
    shape, tile = c["shape"], c["tile"]
    _announce(f"end-to-end autograd, adjoint-check case ({reading})")
    table = rf.load_table()
    vox = TileVoxelizer(shape, tile_size=tile, device=DEV)
    pos = torch.from_numpy(c["pos"]).to(DEV).requires_grad_()
    log_s = torch.from_numpy(np.log(c["scales"])).to(DEV).requires_grad_()
    rot = torch.from_numpy(c["rot"]).to(DEV).requires_grad_()
    dr = torch.from_numpy(c["rr"]).to(DEV).requires_grad_()
    di = torch.from_numpy(c["ri"]).to(DEV).requires_grad_()
    vol_r, vol_i = vox(pos, torch.exp(log_s), rot, torch.complex(dr, di))
    g_r, g_i = cases.call_upstream(shape, seed=3)
    loss = (vol_r * torch.from_numpy(g_r).to(DEV)).sum() + (vol_i * torch.from_numpy(g_i).to(DEV)).sum()
    loss.backward()
    torch.cuda.synchronize()

    got = dict(pos=pos.grad, log_s=log_s.grad, rot=rot.grad, dr=dr.grad, di=di.grad)
    ref, ok, vol64 = rf.fp64_call_reference(c, torch.exp(log_s).detach().cpu().numpy(), g_r, g_i)
    for k in got:
        e = cases.rel_l2(got[k].cpu().numpy()[ok], ref[k][ok])
        tol = rf.tolerance(table, "param", reading, k)
        print(f"  d/d{k}: rel L2 {e:.3e} (reference fp32 {table['param'][reading][k]:.2e}, bound {tol:.2e})")
        assert e <= tol, (k, e, tol)
    # Volume parity through the full call.
    e_vol = cases.rel_l2(np.concatenate([vol_r.detach().cpu().numpy().ravel(), vol_i.detach().cpu().numpy().ravel()]),
                         vol64)
    tol = rf.tolerance(table, "param", reading, "vol")
    print(f"  volume rel L2 {e_vol:.3e} (bound {tol:.2e})")
    assert e_vol <= tol


def _microbench_inputs(n, shape, seed=0):
    from gs_tmi_amd import build_tile_pairs, voxel_space_params

    c = cases.microbench_case(n=n, shape=shape, seed=seed)
    pos = torch.from_numpy(c["pos"]).to(DEV)
    sc = torch.from_numpy(c["scales"]).to(DEV)
    rot = torch.from_numpy(c["rot"]).to(DEV)
    center, radius, cov, hs = voxel_space_params(pos, sc, rot, shape)
    ranges, pids = build_tile_pairs(center, radius, shape, c["tile"])
    rr = torch.from_numpy(c["rr"]).to(DEV)
    ri = torch.from_numpy(c["ri"]).to(DEV)
    return c, center, radius, cov, hs, ranges, pids, rr, ri


def test_full_size_properties():
    """Voxelizer microbenchmark size (256x256x128, 1M Gaussians): binning equals
    the oracle's; forward is deterministic, exactly linear in density, matches
    the FP64 oracle on sampled tiles; backward passes the adjoint dot-product test."""
    from gs_tmi_amd import _native

    shape = (256, 256, 128)
    d, h, w = shape
    _announce("full-size properties 256x256x128, 1M Gaussians")
    c, center, radius, cov, hs, ranges, pids, rr, ri = _microbench_inputs(1_000_000, shape)
    torch.cuda.synchronize()
    c32, r32, _, _ = vo.prologue(c["pos"], c["scales"], c["rot"], shape)
    r_exp, p_exp = vo.bin_tiles(c32, r32, shape, 8)
    print(f"  pairs {p_exp.size}")
    np.testing.assert_array_equal(ranges.cpu().numpy(), r_exp)
    np.testing.assert_array_equal(pids.cpu().numpy(), p_exp)
    del r_exp, p_exp

    v1 = _native.forward(center, cov, rr, ri, ranges, pids, hs, 8, d, h, w)
    v2 = _native.forward(center, cov, rr, ri, ranges, pids, hs, 8, d, h, w)
    v3 = _native.forward(center, cov, 2 * rr, 2 * ri, ranges, pids, hs, 8, d, h, w)
    assert torch.equal(v1[0], v2[0]) and torch.equal(v1[1], v2[1])
    assert torch.equal(2 * v1[0], v3[0]) and torch.equal(2 * v1[1], v3[1])

    # Sampled tiles against the FP64 oracle.
    rng = np.random.default_rng(11)
    tiles = rng.choice(ranges.shape[0], size=24, replace=False)
    rn = ranges.cpu().numpy()
    pn = pids.cpu().numpy()
    sub = np.zeros_like(rn)
    sub[tiles] = rn[tiles]
    er, ei, amb = vo.forward(center.cpu().numpy(), cov.cpu().numpy(), c["rr"], c["ri"], sub, pn, hs.cpu().numpy(),
                             shape, 8, band=1e-4)
    vr = v1[0].cpu().numpy()
    vi = v1[1].cpu().numpy()
    mask = np.zeros(d * h * w, dtype=bool)
    for t in tiles:
        vox_flat = vo._tile_voxels(int(t), shape, 8)[1]
        mask[vox_flat] = True
    mask &= ~amb
    e = cases.rel_l2(np.concatenate([vr[mask], vi[mask]]), np.concatenate([er[mask], ei[mask]]))
    print(f"  sampled-tile forward rel L2 {e:.3e}")
    assert e <= 1e-5

    # Adjoint dot product: the forward is linear in (rho_r, rho_i).  The
    # difference is measured against |vol| |G| (Cauchy-Schwarz), which does not
    # blow up when <vol, G> happens to be small.
    gen = torch.Generator(device="cpu").manual_seed(21)
    g_r = torch.randn(d * h * w, generator=gen).to(DEV)
    g_i = torch.randn(d * h * w, generator=gen).to(DEV)
    b1 = _native.backward(center, cov, rr, ri, ranges, pids, hs, g_r, g_i, 8, d, h, w)
    _, _, grr, gri = b1
    lhs = (v1[0].double() * g_r.double()).sum() + (v1[1].double() * g_i.double()).sum()
    rhs = (rr.double() * grr.double()).sum() + (ri.double() * gri.double()).sum()
    scale = float(torch.sqrt((v1[0].double() ** 2).sum() + (v1[1].double() ** 2).sum()) *
                  torch.sqrt((g_r.double() ** 2).sum() + (g_i.double() ** 2).sum()))
    rel = abs(float(lhs - rhs)) / scale
    print(f"  <J rho, G> = {float(lhs):.6e}, <rho, J^T G> = {float(rhs):.6e}, diff / (|vol||G|) {rel:.2e}")
    assert rel < 1e-7
    # Run to run, the backward differs only by the order of fp32 atomic adds.
    b2 = _native.backward(center, cov, rr, ri, ranges, pids, hs, g_r, g_i, 8, d, h, w)
    for name, x, y in zip(["center", "cov_inv", "rho_r", "rho_i"], b1, b2):
        e = cases.rel_l2(x.cpu().numpy(), y.cpu().numpy())
        print(f"  backward run-to-run {name}: rel L2 {e:.2e}")
        assert e < 1e-6, (name, e)


# The two prologues (HIP, and torch on the GPU) give cov_inv within fp32
# rounding of each other (not bit for bit), which the quadratic form's
# conditioning amplifies for needle-shaped and strongly anisotropic Gaussians
# (axis ratios up to ~600): the exact function itself moves by up to ~1e-4 in
# the volume and ~1e-3 in the gradients between the two fp32 inputs.  The two
# calls must agree within 2e-6 (volume) and 2e-5 (gradients); where the exact
# volumes on the two inputs differ by more than that, the volumes within that
# exact difference (plus 1e-6) and each within 1e-5 of the exact function of
# its own inputs (the gradients' mutual difference is then not a defined
# target: the Function-level tests hold the composition's gradients to FP64 on
# its own inputs, and the fused call is held to FP64 of the parameters below).


def _intrinsic_volume_difference(c, s_gpu):
    """Relative L2 difference (real, imaginary) of the FP64 volumes evaluated
    exactly on the two prologues' fp32 cov_inv (HIP and torch on the GPU), same
    centres and lists; and those two volumes (real then imaginary, flat)."""
    from gs_tmi_amd import _native, voxel_space_params

    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    shape, tile = c["shape"], c["tile"]
    cen, rad, cov_t, hs = voxel_space_params(t(c["pos"]), t(s_gpu), t(c["rot"]), shape)
    _, _, cov_h, _ = _native.prologue(t(c["pos"]), t(s_gpu), t(c["rot"]), shape)
    c32, h32 = cen.cpu().numpy(), hs.cpu().numpy()
    ranges, pids = vo.bin_tiles(c32, rad.cpu().numpy(), shape, tile)
    vols = []
    for cv in (cov_h, cov_t):
        er, ei, _ = vo.forward(c32, cv.cpu().numpy(), c["rr"], c["ri"], ranges, pids, h32, shape, tile)
        vols.append(np.concatenate([er, ei]))
    n = vols[0].size // 2
    return [cases.rel_l2(vols[0][k * n:(k + 1) * n], vols[1][k * n:(k + 1) * n]) for k in (0, 1)], vols


@pytest.mark.parametrize("name", cases.CALL_CASES)
def test_fused_call_matches_unfused_composition(name):
    """TileVoxelizer.forward (fused HIP prologue and epilogue) against the
    reference's composition (torch prologue + binning + 11-argument Function),
    and against FP64 autograd of the call: volume and every parameter gradient
    within 1e-5 relative L2, or 1.5x the reference's own fp32 error on the case
    where that exceeds 1e-5 (tests/golden/reference_fp32_errors.json)."""
    from gs_tmi_amd import TileVoxelizer

    c = cases.case_by_name(name)
    shape, tile = c["shape"], c["tile"]
    _announce(f"fused vs unfused {name} shape={shape} tile={tile}")
    table = rf.load_table()
    vox = TileVoxelizer(shape, tile_size=tile, device=DEV)
    g_r_np, g_i_np = cases.call_upstream(shape)
    g_r, g_i = torch.from_numpy(g_r_np).to(DEV), torch.from_numpy(g_i_np).to(DEV)
    out = {}
    for mode in ("fused", "unfused"):
        leaves = [torch.from_numpy(np.ascontiguousarray(a)).to(DEV).requires_grad_()
                  for a in (c["pos"], np.log(c["scales"]), c["rot"], c["rr"], c["ri"])]
        pos, log_s, rot, dr, di = leaves
        fn = vox.forward if mode == "fused" else vox.forward_unfused
        vr, vi = fn(pos, torch.exp(log_s), rot, torch.complex(dr, di))
        ((vr * g_r).sum() + (vi * g_i).sum()).backward()
        out[mode] = [vr.detach(), vi.detach()] + [t.grad for t in leaves]
        if mode == "fused":
            s_gpu = torch.exp(log_s).detach().cpu().numpy()
    names = ["vol_r", "vol_i", "pos", "log_s", "rot", "dr", "di"]
    intrinsic = None
    for k, (nm, a, b) in enumerate(zip(names[:2], out["fused"], out["unfused"])):
        e = cases.rel_l2(a.cpu().numpy(), b.cpu().numpy())
        tol = 2e-6
        if e > tol:
            intrinsic = intrinsic or _intrinsic_volume_difference(c, s_gpu)
            tol = intrinsic[0][k] + 1e-6
        print(f"  {nm}: fused vs unfused rel L2 {e:.2e} (bound {tol:.2e})")
        assert e <= tol, (nm, e, tol)
    if intrinsic is not None:
        # Each call's volume is the exact function of its own fp32 inputs.
        for mode, exact in zip(("fused", "unfused"), intrinsic[1]):
            got = np.concatenate([out[mode][0].cpu().numpy().ravel(), out[mode][1].cpu().numpy().ravel()])
            e = cases.rel_l2(got, exact)
            print(f"  volume: {mode} vs FP64 on its own prologue's cov_inv rel L2 {e:.2e}")
            assert e <= 1e-5, (mode, e)
    else:
        for nm, a, b in zip(names[2:], out["fused"][2:], out["unfused"][2:]):
            e = cases.rel_l2(a.cpu().numpy(), b.cpu().numpy())
            print(f"  {nm}: fused vs unfused rel L2 {e:.2e}")
            assert e < 2e-5, (nm, e)
    # The fused call against FP64 autograd of the reference's call (its own
    # parity, not only its agreement with the composition).
    ref, ok, vol64 = rf.fp64_call_reference(c, s_gpu, g_r_np, g_i_np)
    for mode in ("fused",):
        got = out[mode]
        e = cases.rel_l2(np.concatenate([got[0].cpu().numpy().ravel(), got[1].cpu().numpy().ravel()]), vol64)
        tol = rf.tolerance(table, "call", name, "vol")
        print(f"  volume: {mode} vs FP64 rel L2 {e:.2e} (reference fp32 {table['call'][name]['vol']:.2e}, "
              f"bound {tol:.2e})")
        assert e <= tol, (mode, "volume", e, tol)
        for nm, a in zip(names[2:], got[2:]):
            e = cases.rel_l2(a.cpu().numpy()[ok], ref[nm][ok])
            tol = rf.tolerance(table, "call", name, nm)
            print(f"  d/d{nm}: {mode} vs FP64 rel L2 {e:.2e} (reference fp32 {table['call'][name][nm]:.2e}, "
                  f"bound {tol:.2e})")
            assert e <= tol, (mode, nm, e, tol)


def test_fused_call_density_layouts():
    """The fused call reads the complex64 density in place; other complex
    dtypes and shapes are converted, and their gradients come back in the
    caller's dtype and shape."""
    from gs_tmi_amd import TileVoxelizer

    c = cases.edge_case("partial_tiles")
    shape, tile = c["shape"], c["tile"]
    _announce(f"fused call density layouts shape={shape} tile={tile}")
    vox = TileVoxelizer(shape, tile_size=tile, device=DEV)
    rng = np.random.default_rng(5)
    g_r = torch.from_numpy(rng.normal(size=shape).astype(np.float32)).to(DEV)
    g_i = torch.from_numpy(rng.normal(size=shape).astype(np.float32)).to(DEV)
    pos = torch.from_numpy(c["pos"]).to(DEV)
    scales = torch.from_numpy(c["scales"]).to(DEV)
    rot = torch.from_numpy(c["rot"]).to(DEV)
    z64 = torch.complex(torch.from_numpy(c["rr"]), torch.from_numpy(c["ri"])).to(DEV)
    out = {}
    for name, dens in (("c64", z64), ("c128", z64.to(torch.complex128)), ("column", z64.view(-1, 1))):
        dens = dens.clone().requires_grad_()
        vr, vi = vox(pos, scales, rot, dens)
        ((vr * g_r).sum() + (vi * g_i).sum()).backward()
        assert dens.grad.dtype == dens.dtype and dens.grad.shape == dens.shape, name
        out[name] = (vr.detach(), vi.detach(), dens.grad.reshape(-1).to(torch.complex64))
    for name in ("c128", "column"):
        assert torch.equal(out[name][0], out["c64"][0]) and torch.equal(out[name][1], out["c64"][1])
        e = cases.rel_l2(torch.view_as_real(out[name][2]).cpu().numpy(),
                         torch.view_as_real(out["c64"][2]).cpu().numpy())
        print(f"  density grad {name} vs complex64: rel L2 {e:.2e}")
        assert e < 5e-6, (name, e)


def test_fused_call_backward_twice():
    """backward(retain_graph=True) twice through the fused call: the second
    adjoint rebuilds its zeroed accumulator and gives the same gradients."""
    from gs_tmi_amd import TileVoxelizer

    c = cases.edge_case("partial_tiles")
    shape, tile = c["shape"], c["tile"]
    _announce(f"fused call backward twice shape={shape} tile={tile}")
    vox = TileVoxelizer(shape, tile_size=tile, device=DEV)
    rng = np.random.default_rng(6)
    g_r = torch.from_numpy(rng.normal(size=shape).astype(np.float32)).to(DEV)
    g_i = torch.from_numpy(rng.normal(size=shape).astype(np.float32)).to(DEV)
    leaves = [torch.from_numpy(np.ascontiguousarray(a)).to(DEV).requires_grad_()
              for a in (c["pos"], np.log(c["scales"]), c["rot"], c["rr"], c["ri"])]
    pos, log_s, rot, dr, di = leaves
    vr, vi = vox(pos, torch.exp(log_s), rot, torch.complex(dr, di))
    loss = (vr * g_r).sum() + (vi * g_i).sum()
    first = torch.autograd.grad(loss, leaves, retain_graph=True)
    second = torch.autograd.grad(loss, leaves)
    for nm, a, b in zip(["pos", "log_s", "rot", "dr", "di"], first, second):
        e = cases.rel_l2(b.cpu().numpy(), a.cpu().numpy())
        print(f"  second backward {nm}: rel L2 {e:.2e}")
        assert e < 5e-6, (nm, e)


def test_fused_call_pair_capacity_and_overflow():
    """The fused call bins into the pair capacity its voxelizer carries from
    earlier calls (the total stays on the device).  The first call (no
    capacity yet) reads the total as the reference does; later calls give the
    same volume bit for bit; a call whose total exceeds the capacity is redone
    with the exact size (same volume, same gradients up to the order of fp32
    atomics) and the capacity grows."""
    from gs_tmi_amd import TileVoxelizer

    shape = (64, 64, 64)
    _announce(f"fused call pair capacity and overflow shape={shape}")
    c = cases.microbench_case(n=20_000, shape=shape, seed=8)
    t = lambda a: torch.from_numpy(a).to(DEV)  # noqa: E731
    vox = TileVoxelizer(shape, tile_size=8, device=DEV)
    gen = torch.Generator(device="cpu").manual_seed(9)
    g_r = torch.randn(shape, generator=gen).to(DEV)
    g_i = torch.randn(shape, generator=gen).to(DEV)

    def call():
        leaves = [t(a).requires_grad_() for a in (c["pos"], np.log(c["scales"]), c["rot"], c["rr"], c["ri"])]
        pos, log_s, rot, dr, di = leaves
        vr, vi = vox(pos, torch.exp(log_s), rot, torch.complex(dr, di))
        grads = torch.autograd.grad((vr * g_r).sum() + (vi * g_i).sum(), leaves)
        return vr.detach(), vi.detach(), [g.detach() for g in grads]

    assert vox.pairs.capacity(20_000) is None
    r0 = call()
    cap = vox.pairs.capacity(20_000)
    assert cap is not None and vox.pairs.overflows == 0
    r1 = call()  # capped
    assert torch.equal(r0[0], r1[0]) and torch.equal(r0[1], r1[1])
    vox.pairs.cap = 100  # far below the total: this call overflows and is redone
    r2 = call()
    assert vox.pairs.overflows == 1 and vox.pairs.capacity(20_000) >= cap - vox.pairs.slack
    assert torch.equal(r0[0], r2[0]) and torch.equal(r0[1], r2[1])
    for nm, a, b in zip(["pos", "log_s", "rot", "dr", "di"], r1[2], r2[2]):
        e = cases.rel_l2(b.cpu().numpy(), a.cpu().numpy())
        print(f"  overflow redo {nm}: rel L2 {e:.2e}")
        assert e < 5e-6, (nm, e)


def test_full_size_fused_call_matches_reference_lists():
    """Voxelizer microbenchmark size: the fused call (extent-box tile lists)
    gives the volume of the forward kernel on the reference's tile lists with
    the same prologue up to fp32 summation order (the lists split into
    512-pair batches differently), and the same density gradient up to the
    order of fp32 atomic additions."""
    from gs_tmi_amd import TileVoxelizer, _native, build_tile_pairs

    shape = (256, 256, 128)
    d, h, w = shape
    _announce("full-size fused call vs reference lists 256x256x128, 1M Gaussians")
    c = cases.microbench_case(n=1_000_000, shape=shape, seed=3)
    t = lambda a: torch.from_numpy(a).to(DEV)  # noqa: E731
    pos, sc, rot, rr, ri = t(c["pos"]), t(c["scales"]), t(c["rot"]), t(c["rr"]), t(c["ri"])
    center, radius, cov, hs = _native.prologue(pos, sc, rot, shape)
    ranges, pids = build_tile_pairs(center, radius, shape, 8)
    vr, vi = _native.forward(center, cov, rr, ri, ranges, pids, hs, 8, d, h, w)
    gen = torch.Generator(device="cpu").manual_seed(5)
    g_r = torch.randn(d * h * w, generator=gen).to(DEV)
    g_i = torch.randn(d * h * w, generator=gen).to(DEV)
    _, _, grr, gri = _native.backward(center, cov, rr, ri, ranges, pids, hs, g_r, g_i, 8, d, h, w)
    dens = torch.complex(rr, ri).requires_grad_()
    fr, fi = TileVoxelizer(shape, tile_size=8, device=DEV)(pos, sc, rot, dens)
    ((fr.reshape(-1) * g_r).sum() + (fi.reshape(-1) * g_i).sum()).backward()
    torch.cuda.synchronize()
    fr, fi = fr.detach().reshape(-1), fi.detach().reshape(-1)
    ev = cases.rel_l2(torch.stack([fr, fi]).cpu().numpy(), torch.stack([vr, vi]).cpu().numpy())
    dmax = float(torch.maximum((fr - vr).abs().max(), (fi - vi).abs().max()))
    print(f"  volume rel L2 {ev:.2e}, max abs difference {dmax:.2e}")
    assert ev < 1e-6
    gd = dens.grad
    e = cases.rel_l2(torch.stack([gd.real, gd.imag]).cpu().numpy(), torch.stack([grr, gri]).cpu().numpy())
    print(f"  pairs: reference {pids.numel()}; density gradient rel L2 {e:.2e}")
    assert e < 1e-6


def test_fused_call_without_gaussians_or_pairs():
    """The fused call with no Gaussians, and with Gaussians whose extent boxes
    miss every tile (no pairs): zero volumes and zero (or empty) gradients."""
    from gs_tmi_amd import TileVoxelizer

    shape = (16, 8, 12)
    _announce(f"fused call without Gaussians / pairs shape={shape}")
    vox = TileVoxelizer(shape, tile_size=8, device=DEV)
    for n, far in ((0, False), (5, True)):
        pos = torch.full((n, 3), 9.0 if far else 0.0, device=DEV, requires_grad=True)
        log_s = torch.full((n, 3), -4.0, device=DEV, requires_grad=True)
        rot = torch.tensor([[1.0, 0.0, 0.0, 0.0]] * n, device=DEV).reshape(n, 4).requires_grad_()
        dens = torch.ones(n, dtype=torch.complex64, device=DEV, requires_grad=True)
        vr, vi = vox(pos, torch.exp(log_s), rot, dens)
        assert vr.shape == shape and float(vr.detach().abs().sum() + vi.detach().abs().sum()) == 0.0
        (vr.sum() + vi.sum()).backward()
        for t in (pos, log_s, rot, dens):
            assert t.grad is not None and t.grad.shape == t.shape
            assert float(t.grad.abs().sum()) == 0.0

