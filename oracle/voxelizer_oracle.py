"""CPU oracle for the tile voxelizer.  TEST INFRASTRUCTURE ONLY.

Only the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker: the product
path (``gs_tmi_amd``) never calls it and has no CPU fallback.

A NumPy restatement of the reference's tile voxelizer (3DGSMR,
``3DGS/3dgsVC/gaussian/``):

* ``prologue``      gaussian/tile_voxelizer.py:281-290 and :77-84.  Centre and
                    radius are rounded to fp32 after every operation, exactly as
                    torch evaluates them, because binning must be bit-exact.
* ``tile_boxes``,
  ``bin_tiles``     gaussian/tile_voxelizer.py:87-162.  The within-tile order of
                    the reference (``torch.argsort``, :147) is unspecified; the
                    oracle uses the stable order (ascending Gaussian id).
* ``forward``       gaussian/cuda_ext/tile_voxelize_cuda_kernel.cu:118-141 in FP64.
* ``backward``      gaussian/cuda_ext/tile_voxelize_cuda_kernel.cu:220-252 in FP64,
                    i.e. the reference's own adjoint formulas.
  Both add only the evaluations with q <= 9 (the kernel's branch, :130 and
  :233), so a non-finite density or upstream gradient reaches exactly the
  quantities the reference's kernel lets it reach.

Pinning (each function is checked by tests/test_oracle_*.py against something
other than itself):
* prologue:  bit-exact against torch fp32 ops for centre/radius; cov_inv against
             the closed form R diag(1/s^2) R^T; quaternion sign and scale invariance.
* binning:   brute force over every voxel of each clamped box (tile of a voxel is
             v // tile), on tiny volumes including partial tiles and out-of-volume
             centres; sortedness and range consistency.
* forward:   closed form of an axis-aligned Gaussian at a voxel centre; linearity
             in density; permutation invariance; integer translation; empty tiles 0.
* backward:  FP64 central finite differences of ``forward`` (h = 1e-6) over all
             parameters whose active set does not change at theta +- h; adjoint
             dot-product identity.
No reference-held golden vector exists for this path (the reference ships no
tests or fixtures), so these checks are the pins.
"""

from __future__ import annotations

import math

import numpy as np

F32 = np.float32


# ---------------------------------------------------------------- prologue --

def quaternion_to_rotation(q: np.ndarray) -> np.ndarray:
    """R(q / (|q| + 1e-8)) for q in (w, x, y, z) order, float64.

    Follows gaussian/tile_voxelizer.py:77-84 (and gaussian/gaussian_model.py:12-19).
    """
    q = np.asarray(q, dtype=np.float64)
    q = q / (np.linalg.norm(q, axis=-1, keepdims=True) + 1e-8)
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    r = np.empty(q.shape[:-1] + (3, 3), dtype=np.float64)
    r[..., 0, 0] = 1 - 2 * y * y - 2 * z * z
    r[..., 0, 1] = 2 * x * y - 2 * z * w
    r[..., 0, 2] = 2 * x * z + 2 * y * w
    r[..., 1, 0] = 2 * x * y + 2 * z * w
    r[..., 1, 1] = 1 - 2 * x * x - 2 * z * z
    r[..., 1, 2] = 2 * y * z - 2 * x * w
    r[..., 2, 0] = 2 * x * z - 2 * y * w
    r[..., 2, 1] = 2 * y * z + 2 * x * w
    r[..., 2, 2] = 1 - 2 * x * x - 2 * y * y
    return r


def prologue(positions, scales, rotations, shape):
    """Voxel-space parameters; gaussian/tile_voxelizer.py:281-290.

    Returns (center_vox f32 [N,3], radius_vox f32 [N], cov_inv f64 [N,3,3],
    half_shape f32 [3]).  ``scales`` are linear (the caller passes exp(log-scale)).
    centre and radius follow torch's fp32 evaluation order op by op:
    c = ((p + 1) * 0.5) * shape - 0.5 ;  r = max_a(s_a * hs_a) * 3.
    """
    p = np.asarray(positions, dtype=F32)
    s = np.asarray(scales, dtype=F32)
    shape32 = np.asarray(shape, dtype=F32)
    hs = shape32 * F32(0.5)
    c = ((p + F32(1.0)) * F32(0.5)) * shape32 - F32(0.5)
    r = (s * hs).max(axis=-1) * F32(3.0)
    rot = quaternion_to_rotation(rotations)
    s_inv = 1.0 / (s.astype(np.float64) + 1e-8)
    t = rot * s_inv[:, None, :]
    cov = t @ np.swapaxes(t, 1, 2)
    return c.astype(F32), r.astype(F32), cov, hs.astype(F32)


# ----------------------------------------------------------------- binning --

def tile_grid(shape, tile):
    d, h, w = shape
    return (math.ceil(d / tile), math.ceil(h / tile), math.ceil(w / tile))


def _sat_f2i(x: np.ndarray) -> np.ndarray:
    """fp32 -> int32 with truncation, saturation and NaN -> 0 (device cast)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros(x.shape, dtype=np.int64)
    fin = np.isfinite(x)
    out[fin] = np.trunc(x[fin]).astype(np.int64)
    out[x == np.inf] = 2**31 - 1
    out[x == -np.inf] = -(2**31)
    return np.clip(out, -(2**31), 2**31 - 1)


def tile_boxes(center_vox, radius_vox, shape, tile):
    """Per-Gaussian tile box (tmin [N,3], tcount [N,3]); gaussian/tile_voxelizer.py:97-116.

    bbox = floor(c - r) .. ceil(c + r) in fp32, clamped to the volume, then
    divided by the tile size and clamped to the tile grid.  An inverted box
    (negative radius) gets a count of 0 on that axis.
    """
    c = np.asarray(center_vox, dtype=F32)
    r = np.asarray(radius_vox, dtype=F32)[:, None]
    size = np.asarray(shape, dtype=np.int64)
    nt = np.asarray(tile_grid(shape, tile), dtype=np.int64)
    lo = _sat_f2i(np.floor(c - r))
    hi = _sat_f2i(np.ceil(c + r))
    lo = np.clip(lo, 0, size - 1)
    hi = np.clip(hi, 0, size - 1)
    tl = np.clip(lo // tile, 0, nt - 1)
    th = np.clip(hi // tile, 0, nt - 1)
    cnt = np.maximum(th - tl + 1, 0)
    return tl, cnt


def expand_pairs(center_vox, radius_vox, shape, tile):
    """(tile id, Gaussian id) pairs in Gaussian-major, row-major-local order
    (gaussian/tile_voxelizer.py:126-145)."""
    nt = tile_grid(shape, tile)
    tl, cnt = tile_boxes(center_vox, radius_vox, shape, tile)
    per = cnt.prod(axis=1)
    n = per.shape[0]
    total = int(per.sum())
    gids = np.repeat(np.arange(n, dtype=np.int64), per)
    starts = (np.cumsum(per) - per).astype(np.int64)
    local = np.arange(total, dtype=np.int64) - np.repeat(starts, per)
    rh = np.repeat(cnt[:, 1], per)
    rw = np.repeat(cnt[:, 2], per)
    ld = local // (rh * rw)
    lr = local % (rh * rw)
    lh = lr // rw
    lw = lr % rw
    td = np.repeat(tl[:, 0], per) + ld
    th = np.repeat(tl[:, 1], per) + lh
    tw = np.repeat(tl[:, 2], per) + lw
    keys = td * nt[1] * nt[2] + th * nt[2] + tw
    return keys, gids


def bin_tiles(center_vox, radius_vox, shape, tile):
    """tile_ranges int32 [T,2] and pair_ids int32 [P] (canonical stable order)."""
    nt = tile_grid(shape, tile)
    num_tiles = nt[0] * nt[1] * nt[2]
    keys, gids = expand_pairs(center_vox, radius_vox, shape, tile)
    ranges = np.zeros((num_tiles, 2), dtype=np.int32)
    if keys.size == 0:
        return ranges, np.zeros(0, dtype=np.int32)
    order = np.argsort(keys, kind="stable")
    ks = keys[order]
    pair_ids = gids[order].astype(np.int32)
    change = np.ones(ks.size, dtype=bool)
    change[1:] = ks[1:] != ks[:-1]
    starts = np.nonzero(change)[0]
    ends = np.concatenate([starts[1:], [ks.size]])
    ranges[ks[starts], 0] = starts
    ranges[ks[starts], 1] = ends
    return ranges, pair_ids


# ------------------------------------------------------- forward / backward --

def _tile_voxels(tile_id, shape, tile):
    d, h, w = shape
    nt = tile_grid(shape, tile)
    td, rem = divmod(tile_id, nt[1] * nt[2])
    th, tw = divmod(rem, nt[2])
    zz = np.arange(td * tile, min(td * tile + tile, d))
    yy = np.arange(th * tile, min(th * tile + tile, h))
    xx = np.arange(tw * tile, min(tw * tile + tile, w))
    g0, g1, g2 = np.meshgrid(zz, yy, xx, indexing="ij")
    vox = np.stack([g0.ravel(), g1.ravel(), g2.ravel()], axis=-1)
    flat = (vox[:, 0] * h + vox[:, 1]) * w + vox[:, 2]
    return vox.astype(np.float64), flat


def _tile_terms(tile_id, center, cov, ranges, pair_ids, half_shape, shape, tile):
    s, e = int(ranges[tile_id, 0]), int(ranges[tile_id, 1])
    if s >= e:
        return None
    gids = np.asarray(pair_ids[s:e], dtype=np.int64)
    vox, flat = _tile_voxels(tile_id, shape, tile)
    hs = np.asarray(half_shape, dtype=np.float64)
    dvec = (vox[:, None, :] - center[gids][None, :, :]) / hs  # [V, G, 3]
    m = cov[gids]  # [G, 3, 3]
    md = np.einsum("gij,vgj->vgi", m, dvec)
    q = np.einsum("vgi,vgi->vg", dvec, md)
    return gids, flat, dvec, md, q


def forward(center_vox, cov_inv, rho_r, rho_i, tile_ranges, pair_ids, half_shape, shape, tile, band=0.0):
    """FP64 forward; returns (vol_r, vol_i, ambiguous) over the flat volume.

    ``ambiguous`` flags voxels with a contributing pair whose q lies within
    ``band`` (relative) of the 9.0 cutoff, where an fp32 evaluation may decide
    the other way.
    """
    center = np.asarray(center_vox, dtype=np.float64)
    cov = np.asarray(cov_inv, dtype=np.float64).reshape(-1, 3, 3)
    rr = np.asarray(rho_r, dtype=np.float64)
    ri = np.asarray(rho_i, dtype=np.float64)
    nvox = int(np.prod(shape))
    vol_r = np.zeros(nvox)
    vol_i = np.zeros(nvox)
    amb = np.zeros(nvox, dtype=bool)
    for t in range(tile_ranges.shape[0]):
        terms = _tile_terms(t, center, cov, tile_ranges, pair_ids, half_shape, shape, tile)
        if terms is None:
            continue
        gids, flat, _, _, q = terms
        act = q <= 9.0
        with np.errstate(over="ignore", invalid="ignore"):
            w = np.where(act, np.exp(-0.5 * q), 0.0)
            # Only evaluations with q <= 9 add anything (the reference's
            # `if (mahal <= 9.0f)`, .cu:130): a non-finite density reaches only
            # those voxels, where a weighted sum would give 0 * inf = NaN.
            vol_r[flat] = np.where(act, w * rr[gids][None, :], 0.0).sum(axis=1)
            vol_i[flat] = np.where(act, w * ri[gids][None, :], 0.0).sum(axis=1)
        if band > 0:
            amb[flat] |= (np.abs(q - 9.0) < 9.0 * band).any(axis=1)
    return vol_r, vol_i, amb


def backward(center_vox, cov_inv, rho_r, rho_i, tile_ranges, pair_ids, half_shape, grad_r, grad_i, shape, tile,
             band=0.0):
    """FP64 adjoint with the reference's formulas (tile_voxelize_cuda_kernel.cu:233-252).

    Returns (d_center [N,3], d_cov [N,3,3], d_rho_r [N], d_rho_i [N], ambiguous [N]),
    ``ambiguous`` flagging Gaussians with a pair within ``band`` of the cutoff.
    """
    center = np.asarray(center_vox, dtype=np.float64)
    cov = np.asarray(cov_inv, dtype=np.float64).reshape(-1, 3, 3)
    rr = np.asarray(rho_r, dtype=np.float64)
    ri = np.asarray(rho_i, dtype=np.float64)
    gr = np.asarray(grad_r, dtype=np.float64).ravel()
    gi = np.asarray(grad_i, dtype=np.float64).ravel()
    hs = np.asarray(half_shape, dtype=np.float64)
    n = center.shape[0]
    d_c = np.zeros((n, 3))
    d_m = np.zeros((n, 3, 3))
    d_rr = np.zeros(n)
    d_ri = np.zeros(n)
    amb = np.zeros(n, dtype=bool)
    for t in range(tile_ranges.shape[0]):
        terms = _tile_terms(t, center, cov, tile_ranges, pair_ids, half_shape, shape, tile)
        if terms is None:
            continue
        gids, flat, dvec, md, q = terms
        act = q <= 9.0  # [V, G]; only these evaluations add anything (.cu:233)
        with np.errstate(over="ignore", invalid="ignore"):
            w = np.where(act, np.exp(-0.5 * q), 0.0)
            vgr = gr[flat][:, None]
            vgi = gi[flat][:, None]
            s = np.where(act, (rr[gids][None, :] * vgr + ri[gids][None, :] * vgi) * (-0.5) * w, 0.0)
            a3 = act[:, :, None]
            # gids are unique inside one tile, so fancy-index accumulation is exact.
            d_rr[gids] += np.where(act, w * vgr, 0.0).sum(axis=0)
            d_ri[gids] += np.where(act, w * vgi, 0.0).sum(axis=0)
            d_c[gids] += np.where(a3, s[:, :, None] * (-2.0 * md / hs), 0.0).sum(axis=0)
            d_m[gids] += np.where(a3[..., None], s[:, :, None, None] * dvec[:, :, :, None] * dvec[:, :, None, :],
                                  0.0).sum(axis=0)
        if band > 0:
            amb[gids] |= (np.abs(q - 9.0) < 9.0 * band).any(axis=0)
    return d_c, d_m, d_rr, d_ri, amb


# ------------------------------------------------------------------ helpers --

def fp32_cov_inv(scales, rotations):
    """cov_inv as the build's prologue forms it in fp32 (R * (1/(s+1e-8)) then
    sum_k T_ik T_jk), used to hand identical fp32 inputs to kernel and oracle."""
    rot = quaternion_to_rotation(np.asarray(rotations, dtype=F32)).astype(F32)
    s_inv = (F32(1.0) / (np.asarray(scales, dtype=F32) + F32(1e-8))).astype(F32)
    t = (rot * s_inv[:, None, :]).astype(F32)
    return np.einsum("gik,gjk->gij", t, t).astype(F32)
