"""CPU check of the column bound the tile kernels cull with.

For a column of voxels (H in [y0, y1], W in [x0, x1], all D rows) the kernels
keep only the D rows inside the intersection of the D extents of the
ellipsoid's projections on (D, H) and (D, W), each cut by the column's strip:

    u_0 in al t -/+ ga sqrt(rho_i^2 - t^2),   t = clamp(+/-ta, lo, hi),
    al = cof_0i / cof_ii,  ga = sqrt(A_jj det A) / cof_ii,  ta = al rho_i^2 / rho_0,

with A the voxel-unit quadratic form sym(M) / (hs hs^T) of the Gaussian,
cofactors and determinant in fp64 (built once per Gaussian), rho_a = e_a - 0.01
the widened extents.  This restates that arithmetic in NumPy, and the kernels'
fp32 evaluation of q' = -log2(e)/2 q as squares of linear forms (factors
rounded to fp32 from an fp64 factorisation), and checks, on Gaussians from
near-spherical to needles and pancakes with axis ratios of several hundred,
that no voxel whose fp32 q passes the cutoff lies outside the kept rows -- and
that the bound does cull.
"""

from __future__ import annotations

import numpy as np

from oracle import voxelizer_oracle as vo

F = np.float32
KQ = -0.72134752044448170368  # -log2(e) / 2
THR = F(9.0 * KQ)


def _fma(a, b, c):
    """fp32 fused multiply-add (exact product and sum in fp64, one rounding)."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(F)


def _form(cov, hs):
    m = cov.astype(np.float32).astype(np.float64)
    s = 0.5 * (m + np.transpose(m, (0, 2, 1)))
    h = hs.astype(np.float64)
    return s / (h[None, :, None] * h[None, None, :])


def _records(center, cov, hs):
    """The kernels' factors of q' (fp32) and the extents e (fp32)."""
    a = _form(cov, hs) * -KQ
    with np.errstate(all="ignore"):
        y1h = np.sqrt(np.maximum(a[:, 1, 1], 0))
        y1d = np.where(y1h > 0, a[:, 1, 0] / y1h, 0)
        y1w = np.where(y1h > 0, a[:, 1, 2] / y1h, 0)
        q00 = a[:, 0, 0] - y1d * y1d
        q02 = a[:, 0, 2] - y1d * y1w
        q22 = a[:, 2, 2] - y1w * y1w
        y2d = np.sqrt(np.maximum(q00, 0))
        y2w = np.where(y2d > 0, q02 / y2d, 0)
        y3w = q22 - y2w * y2w
    s = 0.5 * (cov.astype(np.float64) + np.transpose(cov.astype(np.float64), (0, 2, 1)))
    si = np.linalg.inv(s)
    e = (3.0 * 1.02 * hs.astype(np.float64)[None, :] * np.sqrt(np.einsum("nii->ni", si)) + 0.01).astype(np.float32)
    return tuple(x.astype(F) for x in (y1h, y1d, y1w, y2d, y2w, y3w)), e


def _coefs(cov, hs, e, d=np.float64):
    """Column-bound coefficients, cofactors formed in precision d."""
    a = _form(cov, hs).astype(d)
    a00, a11, a22, a01, a02, a12 = a[:, 0, 0], a[:, 1, 1], a[:, 2, 2], a[:, 0, 1], a[:, 0, 2], a[:, 1, 2]
    cof00 = a11 * a22 - a12 * a12
    cof11 = a00 * a22 - a02 * a02
    cof22 = a00 * a11 - a01 * a01
    cof01 = a02 * a12 - a01 * a22
    cof02 = a01 * a12 - a02 * a11
    det = a00 * cof00 + a01 * cof01 + a02 * cof02
    rho = (e - F(0.01)).astype(np.float64)
    with np.errstate(all="ignore"):
        h = (cof01 / cof11, np.sqrt(a22 * det) / cof11)
        w = (cof02 / cof22, np.sqrt(a11 * det) / cof22)
        h = tuple(x.astype(F) for x in (h[0], h[1], h[0] * (rho[:, 1] * rho[:, 1] / rho[:, 0]), rho[:, 1]))
        w = tuple(x.astype(F) for x in (w[0], w[1], w[0] * (rho[:, 2] * rho[:, 2] / rho[:, 0]), rho[:, 2]))
    ok = (cof11 > 0) & (cof22 > 0) & (det > 0) & np.isfinite(rho).all(1)
    for arr in h + w:
        ok &= np.isfinite(arr)
    return h, w, ok


def _span(al, ga, ta, rho, lo, hi):
    tp = np.clip(ta, lo, hi)
    tm = np.clip(-ta, lo, hi)
    sp = np.sqrt(np.maximum(rho * rho - tp * tp, F(0)))
    sm = np.sqrt(np.maximum(rho * rho - tm * tm, F(0)))
    return al * tm - ga * sm, al * tp + ga * sp


def _check(pos, scales, rot, shape, cof_dtype=np.float64):
    """(misses, kept rows, box rows) of the column bound over every column of
    every Gaussian's box; a miss is a live voxel outside the kept rows."""
    center, _, cov, hs64 = vo.prologue(pos, scales, rot, shape)
    (y1h, y1d, y1w, y2d, y2w, y3w), e = _records(center, cov, hs64)
    hcoef, wcoef, ok = _coefs(cov, hs64, e, cof_dtype)
    misses = kept = total = 0
    zi, yi, xi = (np.arange(v) for v in shape)
    z, y, x = np.meshgrid(zi, yi, xi, indexing="ij")
    for g in range(len(pos)):
        c = center[g]
        lo = np.maximum(np.ceil(c - e[g]), 0)
        hi = np.minimum(np.floor(c + e[g]), np.asarray(shape) - 1)
        if np.any(lo > hi):
            continue
        u0, u1, u2 = z.astype(F) - c[0], y.astype(F) - c[1], x.astype(F) - c[2]
        # The kernels' order: per D row t1 = y1d u0 + y1w u2, y2, r; per voxel y1, q'.
        t1 = _fma(y1d[g], u0, (y1w[g] * u2).astype(F))
        y2 = _fma(y2d[g], u0, (y2w[g] * u2).astype(F))
        r = _fma(-y2, y2, -((y3w[g] * u2).astype(F) * u2).astype(F))
        yy1 = _fma(y1h[g], u1, t1)
        live = _fma(-yy1, yy1, r) >= THR
        for y0 in range(int(lo[1]), int(hi[1]) + 1, 4):
            y1 = min(y0 + 3, int(hi[1]))
            for x0 in range(int(lo[2]), int(hi[2]) + 1, 4):
                x1 = min(x0 + 3, int(hi[2]))
                if ok[g]:
                    hmin, hmax = _span(*(v[g] for v in hcoef), F(y0) - c[1], F(y1) - c[1])
                    wmin, wmax = _span(*(v[g] for v in wcoef), F(x0) - c[2], F(x1) - c[2])
                    zl = max(np.ceil(c[0] + max(hmin, wmin) - F(0.01)), lo[0])
                    zh = min(np.floor(c[0] + min(hmax, wmax) + F(0.01)), hi[0])
                else:
                    zl, zh = lo[0], hi[0]
                col = live[int(lo[0]):int(hi[0]) + 1, y0:y1 + 1, x0:x1 + 1]
                rows = np.nonzero(col.any(axis=(1, 2)))[0] + lo[0]
                if rows.size and not (zl <= rows.min() and rows.max() <= zh):
                    misses += 1
                kept += max(zh - zl + 1, 0)
                total += hi[0] - lo[0] + 1
    return misses, kept, total


def test_column_bound_keeps_every_active_voxel():
    rng = np.random.default_rng(11)
    shape = (24, 24, 24)
    hs = np.asarray(shape, np.float32) * 0.5
    n = 160
    sig = np.concatenate([np.exp(rng.normal(np.log(1.5), 0.5, (n // 2, 3))),
                          np.minimum(np.exp(rng.uniform(np.log(0.05), np.log(20.0), (n - n // 2, 3))), 20.0)])
    cen = rng.uniform(0, 24, (n, 3))
    pos = ((cen + 0.5) / hs - 1).astype(np.float32)
    scales = (sig / hs).astype(np.float32)
    rot = rng.normal(size=(n, 4)).astype(np.float32)
    misses, kept, total = _check(pos, scales, rot, shape)
    assert misses == 0
    assert kept < 0.8 * total  # the bound removes a real share of the box rows


def test_needle_needs_fp64_cofactors():
    """A needle (voxel sigma 12.5 x 6.8 x 0.056, rotated) whose form's
    determinant cancels by ~1e4: with fp32 cofactors the bound drops live
    voxels, with fp64 cofactors (as the kernels' per-Gaussian table forms them)
    it keeps them all."""
    shape = (40, 40, 40)
    pos = np.array([[-0.4479787051677704, -0.09544815868139267, -0.3014869689941406]], np.float32)
    scales = np.array([[0.6251696944236755, 0.3388972878456116, 0.00277737551368773]], np.float32)
    rot = np.array([[1.7393678426742554, -0.49591073393821716, 0.3289696276187897, -0.2585725486278534]], np.float32)
    assert _check(pos, scales, rot, shape)[0] == 0
    assert _check(pos, scales, rot, shape, np.float32)[0] > 0


def _hw_coefs(cov, hs, e):
    """The (H, W) projection of the binning's tile test: the H range along
    u_2 = t, al = cof_12 / cof_22, ga = sqrt(A_00 det A) / cof_22,
    ta = al rho_2^2 / rho_1 (fp64, rounded to fp32)."""
    a = _form(cov, hs)
    a00, a11, a22, a01, a02, a12 = a[:, 0, 0], a[:, 1, 1], a[:, 2, 2], a[:, 0, 1], a[:, 0, 2], a[:, 1, 2]
    cof00 = a11 * a22 - a12 * a12
    cof22 = a00 * a11 - a01 * a01
    cof01 = a02 * a12 - a01 * a22
    cof02 = a01 * a12 - a02 * a11
    cof12 = a01 * a02 - a00 * a12
    det = a00 * cof00 + a01 * cof01 + a02 * cof02
    rho = (e - F(0.01)).astype(np.float64)
    with np.errstate(all="ignore"):
        al = cof12 / cof22
        return tuple(x.astype(F) for x in (al, np.sqrt(a00 * det) / cof22, al * (rho[:, 2] * rho[:, 2] / rho[:, 1]),
                                           rho[:, 2]))


def _check_tiles(pos, scales, rot, shape, tile, margin=F(0.05)):
    """(wrongly dead tiles, dead tiles, box tiles) of the binning's tile test
    (live_tiles): a tile of a Gaussian's extent box is dead when the D rows
    that the (D, H) and (D, W) projections leave over the tile's whole H and
    W strips, or the H rows that the (H, W) projection leaves over its W strip,
    are empty (bounds widened by `margin`).  A dead tile is wrong if a voxel of
    it passes the kernels' fp32 cutoff (its pair would then add something)."""
    center, _, cov, hs64 = vo.prologue(pos, scales, rot, shape)
    (y1h, y1d, y1w, y2d, y2w, y3w), e = _records(center, cov, hs64)
    hcoef, wcoef, ok = _coefs(cov, hs64, e)
    hw = _hw_coefs(cov, hs64, e)
    ok &= np.isfinite(np.stack(hw, 1)).all(1) & (hw[3] > 0)
    wrong = dead = total = 0
    zi, yi, xi = (np.arange(v) for v in shape)
    z, y, x = np.meshgrid(zi, yi, xi, indexing="ij")
    size = np.asarray(shape)
    for g in range(len(pos)):
        c = center[g]
        blo, bhi = np.ceil(c - e[g]), np.floor(c + e[g])
        lo, hi = np.maximum(blo, 0), np.minimum(bhi, size - 1)
        if np.any(lo > hi) or not ok[g]:
            continue
        u0, u1, u2 = z.astype(F) - c[0], y.astype(F) - c[1], x.astype(F) - c[2]
        t1 = _fma(y1d[g], u0, (y1w[g] * u2).astype(F))
        y2 = _fma(y2d[g], u0, (y2w[g] * u2).astype(F))
        r = _fma(-y2, y2, -((y3w[g] * u2).astype(F) * u2).astype(F))
        yy1 = _fma(y1h[g], u1, t1)
        live = _fma(-yy1, yy1, r) >= THR
        tl, th = (lo // tile).astype(int), (hi // tile).astype(int)
        for t0 in range(tl[0], th[0] + 1):
            for t1_ in range(tl[1], th[1] + 1):
                for t2 in range(tl[2], th[2] + 1):
                    t = np.array([t0, t1_, t2])
                    vlo = np.maximum(blo, t * tile).astype(F)
                    vhi = np.minimum(bhi, np.minimum(t * tile + tile, size) - 1).astype(F)
                    total += 1
                    hmn, hmx = _span(*(v[g] for v in hcoef), vlo[1] - c[1], vhi[1] - c[1])
                    wmn, wmx = _span(*(v[g] for v in wcoef), vlo[2] - c[2], vhi[2] - c[2])
                    ymn, ymx = _span(*(v[g] for v in hw), vlo[2] - c[2], vhi[2] - c[2])
                    zl = max(np.ceil(c[0] + max(hmn, wmn) - margin), vlo[0])
                    zh = min(np.floor(c[0] + min(hmx, wmx) + margin), vhi[0])
                    yl = max(np.ceil(c[1] + ymn - margin), vlo[1])
                    yh = min(np.floor(c[1] + ymx + margin), vhi[1])
                    if not (zl > zh or yl > yh):
                        continue
                    dead += 1
                    a, b = vlo.astype(int), vhi.astype(int)
                    wrong += bool(live[a[0]:b[0] + 1, a[1]:b[1] + 1, a[2]:b[2] + 1].any())
    return wrong, dead, total


def test_binning_tile_test_drops_only_tiles_without_evaluations():
    """The fused call's binning drops a (tile, Gaussian) pair when the tile
    test finds no live voxel; a dropped tile must hold no voxel whose fp32 q
    passes the cutoff, on Gaussians from spheres to needles and pancakes, over
    several tile sizes -- and the test must drop a real share of the box
    tiles."""
    rng = np.random.default_rng(12)
    shape = (24, 24, 24)
    hs = np.asarray(shape, np.float32) * 0.5
    n = 120
    sig = np.concatenate([np.exp(rng.normal(np.log(1.5), 0.5, (n // 2, 3))),
                          np.minimum(np.exp(rng.uniform(np.log(0.05), np.log(20.0), (n - n // 2, 3))), 20.0)])
    cen = rng.uniform(0, 24, (n, 3))
    pos = ((cen + 0.5) / hs - 1).astype(np.float32)
    scales = (sig / hs).astype(np.float32)
    rot = rng.normal(size=(n, 4)).astype(np.float32)
    for tile in (3, 4, 8, 12):
        wrong, dead, total = _check_tiles(pos, scales, rot, shape, tile)
        assert wrong == 0, tile
        assert dead > 0.1 * total, (tile, dead, total)
    needle = (np.array([[-0.4479787051677704, -0.09544815868139267, -0.3014869689941406]], np.float32),
              np.array([[0.6251696944236755, 0.3388972878456116, 0.00277737551368773]], np.float32),
              np.array([[1.7393678426742554, -0.49591073393821716, 0.3289696276187897, -0.2585725486278534]],
                       np.float32))
    assert _check_tiles(*needle, (40, 40, 40), 8)[0] == 0
