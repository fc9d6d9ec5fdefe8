"""CPU check of the column bound the tile kernels cull with.

For a column of voxels (H in [y0, y1], W in [x0, x1], all D rows) the kernels
keep only the D rows inside the intersection of the D extents of the
ellipsoid's projections on (D, H) and (D, W), each cut by the column's strip:

    u_0 in al t -/+ ga sqrt(rho_i^2 - t^2),   t = clamp(+/-ta, lo, hi),
    al = cof_0i / cof_ii,  ga = sqrt(A'_jj det A') / cof_ii,  ta = al rho_i^2 / rho_0,

with A' the fp32 scaled quadratic form of the Gaussian record, cofactors and
determinant in fp64, rho_a = e_a - 0.01 the widened extents.  This restates
that arithmetic in NumPy (fp32 where the kernels use fp32) and checks, on
Gaussians from near-spherical to needles and pancakes with axis ratios of
several hundred, that no voxel whose fp32 quadratic form passes the cutoff
lies outside the kept rows -- and that the bound does cull.
"""

from __future__ import annotations

import numpy as np

from oracle import voxelizer_oracle as vo

F = np.float32
KQ = F(-0.72134752044448170368)  # -log2(e) / 2
THR = F(9.0 * -0.72134752044448170368)


def _records(center, cov, hs):
    m = cov.astype(np.float32)
    h = hs.astype(np.float32)
    a00 = KQ * (m[:, 0, 0] / (h[0] * h[0]))
    a11 = KQ * (m[:, 1, 1] / (h[1] * h[1]))
    a22 = KQ * (m[:, 2, 2] / (h[2] * h[2]))
    b01 = KQ * ((m[:, 0, 1] + m[:, 1, 0]) / (h[0] * h[1]))
    b02 = KQ * ((m[:, 0, 2] + m[:, 2, 0]) / (h[0] * h[2]))
    b12 = KQ * ((m[:, 1, 2] + m[:, 2, 1]) / (h[1] * h[2]))
    s = 0.5 * (cov.astype(np.float64) + np.transpose(cov.astype(np.float64), (0, 2, 1)))
    si = np.linalg.inv(s)
    e = (3.0 * 1.02 * hs.astype(np.float64)[None, :] * np.sqrt(np.einsum("nii->ni", si)) + 0.01).astype(np.float32)
    return a00, a11, a22, b01, b02, b12, e


def _coefs(a00, a11, a22, b01, b02, b12, e, d=np.float64):
    s00, s11, s22 = a00.astype(d), a11.astype(d), a22.astype(d)
    s01, s02, s12 = 0.5 * b01.astype(d), 0.5 * b02.astype(d), 0.5 * b12.astype(d)
    cof00 = s11 * s22 - s12 * s12
    cof11 = s00 * s22 - s02 * s02
    cof22 = s00 * s11 - s01 * s01
    cof01 = s02 * s12 - s01 * s22
    cof02 = s01 * s12 - s02 * s11
    det = s00 * cof00 + s01 * cof01 + s02 * cof02
    c11, c22, c01, c02, dt = (x.astype(np.float32) for x in (cof11, cof22, cof01, cof02, det))
    rho = e - F(0.01)
    with np.errstate(all="ignore"):
        h = (c01 / c11, np.sqrt(a22 * dt) / c11)
        w = (c02 / c22, np.sqrt(a11 * dt) / c22)
        h = (h[0], h[1], h[0] * (rho[:, 1] * rho[:, 1] / rho[:, 0]), rho[:, 1])
        w = (w[0], w[1], w[0] * (rho[:, 2] * rho[:, 2] / rho[:, 0]), rho[:, 2])
    ok = (cof11 > 0) & (cof22 > 0) & (det < 0) & np.isfinite(rho).all(1)
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
    a00, a11, a22, b01, b02, b12, e = _records(center, cov, hs64)
    hcoef, wcoef, ok = _coefs(a00, a11, a22, b01, b02, b12, e, cof_dtype)
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
        # The forward kernel's Horner order: (a00 u0 + b02 u2) u0 + a22 u2^2, then H.
        a_part = (a00[g] * u0 + b02[g] * u2) * u0 + a22[g] * u2 * u2
        b_part = b01[g] * u0 + b12[g] * u2
        live = ((a11[g] * u1 + b_part) * u1 + a_part) >= THR
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
    """A needle (voxel sigma 0.067 x 16 x 4.4, rotated) whose form's determinant
    cancels by ~1e4: with fp32 cofactors the bound drops live voxels, with fp64
    cofactors it keeps them all."""
    shape = (40, 40, 40)
    pos = np.array([[0.45545065, -0.53229886, -0.22736628]], np.float32)
    scales = np.array([[0.003345329547300935, 0.7984070181846619, 0.22116735577583313]], np.float32)
    rot = np.array([[-0.14371681, 0.44219768, 0.7676122, 0.35766584]], np.float32)
    assert _check(pos, scales, rot, shape)[0] == 0
    assert _check(pos, scales, rot, shape, np.float32)[0] > 0
