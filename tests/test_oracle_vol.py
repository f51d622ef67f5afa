"""Pins the oracle's volume path (GlobalRepetition3D_ with its four metrics,
as intended where the reference does not run); no GPU needed.

* the ellipsoid footprint: Gauss's sphere lattice counts (radius 1, 2, 3, 4:
  7, 33, 123, 257 points) and an independent numpy evaluation of the float
  predicate for ellipsoids, in z-y-x order;
* the 3D statistics and the noise variance (float mean of std^2 over the
  inner voxels) against FP64 numpy;
* the maps of all four metrics: float form against the FP64 form, and the
  FP64 form against an independent vectorised numpy restatement (block
  matrix, correlation or distance matrices, weights, flat and threshold
  rules, finalize, inversion, mask, normalisation);
* symmetries that the full-ellipsoid reference gather makes exact: flipping
  the volume along z (or transposing x and y with a square block) flips
  (transposes) the FP64 map; the reference's one-slice-short gather would
  break them;
* flat reference blocks score 0.
"""
import numpy as np
import pytest

import oracle as O


def _ellipsoid_numpy(bw, bh, bd):
    brw, brh, brd = bw // 2, bh // 2, bd // 2
    out = []
    for z in range(bd):
        for y in range(bh):
            for x in range(bw):
                fx, fy, fz = np.float32(x - brw), np.float32(y - brh), np.float32(z - brd)
                v = (fx * fx) / np.float32(brw * brw) + (fy * fy) / np.float32(brh * brh)
                v = np.float32(v) + (fz * fz) / np.float32(brd * brd)
                if np.float32(v) <= np.float32(1.0):
                    out.append((x - brw, y - brh, z - brd))
    return out


def test_footprint3d():
    for r, k in [(1, 7), (2, 33), (3, 123), (4, 257)]:
        assert O.footprint3d(2 * r + 1, 2 * r + 1, 2 * r + 1)[0] == k
    for b in [(5, 3, 3), (3, 5, 7), (9, 5, 3), (7, 7, 5)]:
        k, dx, dy, dz = O.footprint3d(*b)
        assert list(zip(dx, dy, dz)) == _ellipsoid_numpy(*b)


def _blobs(d, h, w, seed, n=12):
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.mgrid[0:d, 0:h, 0:w]
    v = np.full((d, h, w), 0.1)
    for _ in range(n):
        cz, cy, cx = rng.uniform(0, d), rng.uniform(0, h), rng.uniform(0, w)
        v += rng.uniform(0.5, 1.0) * np.exp(-((zz - cz) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2) / (2 * 1.8 ** 2))
    v += 0.03 * rng.standard_normal(v.shape)
    return ((v - v.min()) / (v.max() - v.min())).astype(np.float32)


def _stats64(vol, bw, bh, bd):
    k, dx, dy, dz = O.footprint3d(bw, bh, bd)
    d, h, w = vol.shape
    brw, brh, brd = bw // 2, bh // 2, bd // 2
    v = vol.astype(np.float64)
    blocks = np.stack([v[brd + z:d - brd + z, brh + y:h - brh + y, brw + x:w - brw + x]
                       for x, y, z in zip(dx, dy, dz)], -1)  # inner voxels x K
    return blocks, blocks.mean(-1), blocks.std(-1, ddof=1)


def test_stats3d_and_noise_var():
    vol = _blobs(14, 18, 20, 1)
    m, s = O.local_stats3d_f32(vol, 5, 3, 3)
    _, m64, s64 = _stats64(vol, 5, 3, 3)
    np.testing.assert_allclose(m[1:-1, 1:-1, 2:-2], m64, atol=1e-6)
    np.testing.assert_allclose(s[1:-1, 1:-1, 2:-2], s64, atol=1e-6)
    assert not m[0].any() and not s[:, :, :2].any()
    nv = O.vol_noise_var(s, 5, 3, 3)
    assert nv == pytest.approx((s64 ** 2).mean(), rel=1e-5)


def _global3d_numpy(vol, metric, bw, bh, bd, c):
    """GlobalRepetition3D_ per metric in FP64 numpy, from the kernels' rules:
    rows need var >= threshold and std != 0, columns var >= threshold (and std
    != 0 for AbsDiffStds); Pearson/SSIM score 0 against flat columns; NRMSE
    adds sqrt(mse) * weight; finalize, inversion, mask and normalisation."""
    d, h, w = vol.shape
    brw, brh, brd = bw // 2, bh // 2, bd // 2
    _, s32 = O.local_stats3d_f32(vol, bw, bh, bd)
    inner = s32[brd:d - brd, brh:h - brh, brw:w - brw].ravel()
    nv = np.float32(O.vol_noise_var(s32, bw, bh, bd))
    thr = np.float32(nv * np.float32(c))
    blocks, mu, sd = _stats64(vol, bw, bh, bd)
    B = blocks.reshape(-1, blocks.shape[-1]) - mu.reshape(-1, 1)
    mu, sd = mu.ravel(), sd.ravel()
    K = B.shape[1]
    v32 = (inner * inner).astype(np.float32)
    valid = v32.astype(np.float64) >= float(thr)
    rows = valid & (v32 != 0)
    cols = rows if metric == O.ABSDIFF else valid
    eps = float(np.float32(1e-7))
    fp = float(nv)
    c1, c2 = float(np.float32(0.0001)), float(np.float32(0.0009))
    sc, mc, Bc = sd[cols], mu[cols], B[cols]
    flat_c = (inner[cols] == 0)[None, :]
    S = np.zeros(len(sd))
    W = np.zeros(len(sd))
    for r0 in range(0, len(sd), 64):  # row chunks keep the pair tensors small
        r1 = min(r0 + 64, len(sd))
        dd = sd[r0:r1, None] - sc[None, :]
        wt = np.exp(-dd ** 2 / (fp + eps))
        if metric == O.ABSDIFF:
            T = np.abs(dd) * wt
        elif metric == O.NRMSE:
            mse = ((B[r0:r1, None, :] - Bc[None, :, :]) ** 2).sum(-1) / K
            T = np.sqrt(mse) * wt
        else:
            cov = B[r0:r1] @ Bc.T / (K - 1)
            if metric == O.PEARSON:
                T = np.maximum(cov / (sd[r0:r1, None] * sc[None, :] + eps), 0.0) * wt
            else:
                m = mu[r0:r1, None]
                T = (2 * m * mc[None, :] + c1) * (2 * cov + c2) / ((m * m) * (mc[None, :] ** 2) + c1) / \
                    (sd[r0:r1, None] ** 2 + sc[None, :] ** 2 + c2) * wt
                T = np.maximum(T, 0.0)
            T = np.where(flat_c, 0.0, T)
        S[r0:r1] = np.where(rows[r0:r1], T.sum(1), 0.0)
        W[r0:r1] = np.where(rows[r0:r1], wt.sum(1), 0.0)
    mask = (v32 > thr).astype(np.float64)
    npx = float(mask.sum())
    out = S / (W * npx + eps)
    if metric in (O.NRMSE, O.ABSDIFF):
        out = np.where(mask > 0, np.where(out == 0, 1.0, 1.0 / np.where(out == 0, 1.0, out)), out)
    if c > 0:
        out = out * mask
    sel = mask > 0
    mn, mx = out[sel].min(), out[sel].max()
    out = np.where(sel, (out - mn) / (mx - mn + eps), out)
    full = np.zeros((d, h, w))
    full[brd:d - brd, brh:h - brh, brw:w - brw] = out.reshape(d - 2 * brd, h - 2 * brh, w - 2 * brw)
    return full


@pytest.mark.parametrize("b,c", [((3, 3, 3), 0.0), ((5, 5, 3), 0.0), ((3, 3, 3), 0.5), ((3, 5, 5), 1.0)])
def test_pearson3d_against_numpy(b, c):
    vol = _blobs(12, 16, 18, 2)
    r64 = O.global_pearson3d(vol, *b, c=c, fp64=True)
    np.testing.assert_allclose(r64.out, _global3d_numpy(vol, O.PEARSON, *b, c), atol=1e-9)
    r32 = O.global_pearson3d(vol, *b, c=c)
    assert np.abs(r32.out - r64.out).max() < 1e-4


@pytest.mark.parametrize("metric", [O.SSIM, O.NRMSE, O.ABSDIFF])
@pytest.mark.parametrize("b,c", [((3, 3, 3), 0.0), ((5, 3, 3), 0.5), ((3, 5, 5), 1.0)])
def test_metrics3d_against_numpy(metric, b, c):
    vol = _blobs(10, 14, 16, 5)
    vol[2:8, 3:9, 4:10] = 0.5  # flat region: flat rows and (for AbsDiffStds) skipped flat columns
    r64 = O.global_repetition3d(vol, metric, *b, c=c, fp64=True)
    np.testing.assert_allclose(r64.out, _global3d_numpy(vol, metric, *b, c), atol=1e-8)
    r32 = O.global_repetition3d(vol, metric, *b, c=c)
    assert np.abs(r32.out - r64.out).max() < 2e-4
    with pytest.raises(ValueError):
        O.global_repetition3d(vol, 4, *b)


def test_pearson3d_symmetries_and_flat_rows():
    vol = _blobs(12, 16, 16, 3)
    a = O.global_pearson3d(vol, 3, 3, 3, fp64=True).out
    # (the noise variance and validity come from the float statistics, whose
    # summation order the flip changes: agreement to 1e-7, where a
    # one-slice-short reference gather differs by O(0.1))
    np.testing.assert_allclose(O.global_pearson3d(vol[::-1].copy(), 3, 3, 3, fp64=True).out, a[::-1], atol=1e-7)
    np.testing.assert_allclose(O.global_pearson3d(vol.transpose(0, 2, 1).copy(), 3, 3, 3, fp64=True).out,
                               a.transpose(0, 2, 1), atol=1e-7)
    flat = vol.copy()
    flat[3:9, 4:10, 4:10] = 0.5  # flat region: its inner reference blocks score 0
    r = O.global_pearson3d(flat, 3, 3, 3)
    assert not r.S[5:7, 6:8, 6:8].any() and not r.out[5:7, 6:8, 6:8].any()
