"""Pins the oracle's 3D block repetition (BlockRepetition3D_ ->
CLUtils.calculateBlockRepetitionMap3D, as intended); no GPU needed.

* Utils.getReferenceBlock3D: the ellipsoid's pixels in z, y, x order, the
  two normalisations and the mean/std against FP64 numpy; a constant block
  gives zeros with mean and std 0; even sides and a depth below 3 refused;
* the raw kernel outputs of all four metrics against an independent
  vectorised FP64 numpy restatement (block matrix, min/max remaps, local
  mean subtraction, dot products, distances, the flat special cases);
* the maps: inversion (NRMSE), 1 - d (AbsDiffStds), the 3D mask and the
  normalisation against numpy on the oracle's raw output;
* the decisions: with the block radius (not the block depth, as
  getBlockPearson3D passes it) the Pearson map is non-zero inside; with
  the x radius (not the block width, as getBlockSsim3D passes it) the SSIM
  block has the reference block's K voxels.
"""
import numpy as np
import pytest

import oracle as O

EPS = float(np.float32(1e-7))


def _vol(d, h, w, seed, n=10):
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.mgrid[0:d, 0:h, 0:w]
    v = np.full((d, h, w), 0.1)
    for _ in range(n):
        cz, cy, cx = rng.uniform(0, d), rng.uniform(0, h), rng.uniform(0, w)
        v += rng.uniform(0.5, 1.0) * np.exp(-((zz - cz) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2) / (2 * 1.6 ** 2))
    v += 0.03 * rng.standard_normal(v.shape)
    return ((v - v.min()) / (v.max() - v.min())).astype(np.float32)


def _ellipsoid_pixels(block):
    bd, bh, bw = block.shape
    k, dx, dy, dz = O.footprint3d(bw, bh, bd)
    return np.array([block[bd // 2 + z, bh // 2 + y, bw // 2 + x] for x, y, z in zip(dx, dy, dz)], np.float64)


def _norm64(a):
    return (a - a.min()) / (a.max() - a.min() + EPS)


def test_reference_block3d():
    rng = np.random.default_rng(3)
    for shape in [(3, 3, 3), (5, 5, 5), (3, 7, 5), (7, 5, 9)]:
        blk = rng.random(shape).astype(np.float32)
        px, m, s = O.reference_block3d(blk)
        p = _norm64(_ellipsoid_pixels(blk))
        p = _norm64(p - p.mean())
        assert len(px) == O.footprint3d(shape[2], shape[1], shape[0])[0]
        np.testing.assert_allclose(px, p, atol=1e-6)
        assert px.min() == 0.0 and m == pytest.approx(p.mean(), abs=1e-6)
        assert s == pytest.approx(p.std(ddof=1), abs=1e-6)
    px, m, s = O.reference_block3d(np.full((3, 5, 5), 0.7, np.float32))
    assert not px.any() and m == 0.0 and s == 0.0
    assert O.reference_block3d(np.ones((3, 4, 5), np.float32)) is None
    assert O.reference_block3d(np.ones((1, 5, 5), np.float32)) is None


def _blocks(vol, bw, bh, bd):
    k, dx, dy, dz = O.footprint3d(bw, bh, bd)
    d, h, w = vol.shape
    brw, brh, brd = bw // 2, bh // 2, bd // 2
    v = vol.astype(np.float64)
    return np.stack([v[brd + z:d - brd + z, brh + y:h - brh + y, brw + x:w - brw + x] for x, y, z in zip(dx, dy, dz)],
                    -1)


def _raw_numpy(metric, vol, tmpl, bw, bh, bd, tm, ts):
    """The four kernels' formulas in FP64 over the inner voxels."""
    d, h, w = vol.shape
    brw, brh, brd = bw // 2, bh // 2, bd // 2
    B = _blocks(vol, bw, bh, bd)
    m32, s32 = O.local_stats3d_f32(vol, bw, bh, bd)
    cm = m32[brd:d - brd, brh:h - brh, brw:w - brw].astype(np.float64)[..., None]
    cs = s32[brd:d - brd, brh:h - brh, brw:w - brw].astype(np.float64)
    t = tmpl.astype(np.float64)
    K = B.shape[-1]
    if metric == O.ABSDIFF:
        r = np.abs(ts - cs)
    elif metric == O.NRMSE:
        r = np.sqrt(((t - (B - cm)) ** 2).mean(-1))
    elif metric == O.PEARSON:
        t1 = (B - B.min(-1, keepdims=True)) / (np.ptp(B, -1)[..., None] + EPS) - cm
        t3 = (t1 - t1.min(-1, keepdims=True)) / (np.ptp(t1, -1)[..., None] + EPS)
        cov = (t3 * t).sum(-1) / (K - 1)
        r = np.maximum(cov / (ts * cs + EPS), 0.0)
        r = np.where(cs == 0, 1.0 if ts == 0 else 0.0, r)
    else:
        cov = ((B - cm) * t).sum(-1) / (K - 1)
        c1, c2 = float(np.float32(0.0001)), float(np.float32(0.0009))
        mu = cm[..., 0]
        r = (2 * tm * mu + c1) * (2 * cov + c2) / ((tm * tm + mu * mu + c1) * (ts * ts + cs * cs + c2))
        r = np.where(cs == 0, 1.0 if ts == 0 else 0.0, np.maximum(r, 0.0))
    full = np.zeros((d, h, w))
    full[brd:d - brd, brh:h - brh, brw:w - brw] = r
    return full


def _template(vol, z, y, x, bw, bh, bd, noise, seed):
    blk = vol[z - bd // 2:z + bd // 2 + 1, y - bh // 2:y + bh // 2 + 1, x - bw // 2:x + bw // 2 + 1].copy()
    blk += noise * np.random.default_rng(seed).standard_normal(blk.shape).astype(np.float32)
    return O.reference_block3d(blk)


@pytest.mark.parametrize("metric", [O.PEARSON, O.SSIM, O.NRMSE, O.ABSDIFF])
@pytest.mark.parametrize("b", [(3, 3, 3), (5, 5, 3), (3, 5, 7)])
def test_raw_against_numpy(metric, b):
    vol = _vol(12, 18, 20, 1)
    vol[6:10, 8:14, 8:14] = 0.5  # flat region: test std 0
    px, tm, ts = _template(vol, 5, 9, 10, *b, 0.02, 4)
    m32, s32 = O.local_stats3d_f32(vol, *b)
    got = O.block_raw3d(metric, vol, m32, s32, px, *b, tm, ts)
    want = _raw_numpy(metric, vol, px, *b, tm, ts)
    scale = max(np.abs(want).max(), 1.0)
    assert np.abs(got - want).max() <= 2e-5 * scale
    bw, bh, bd = b
    inner = got[bd // 2:-(bd // 2), bh // 2:-(bh // 2), bw // 2:-(bw // 2)]
    assert inner.any()  # with the block depth as the z radius every voxel would be border


def test_flat_template_special_cases():
    vol = _vol(10, 14, 14, 2)
    vol[3:7, 4:10, 4:10] = 0.25
    px, tm, ts = O.reference_block3d(np.full((3, 3, 3), 0.4, np.float32))
    m32, s32 = O.local_stats3d_f32(vol, 3, 3, 3)
    for metric in (O.PEARSON, O.SSIM):
        got = O.block_raw3d(metric, vol, m32, s32, px, 3, 3, 3, tm, ts)
        inner = np.zeros_like(vol, bool)
        inner[1:-1, 1:-1, 1:-1] = True
        np.testing.assert_array_equal(got[inner], np.where(s32[inner] == 0, 1.0, 0.0))


def _map_numpy(metric, raw, stds, bw, bh, bd, c):
    d, h, w = raw.shape
    brw, brh, brd = bw // 2, bh // 2, bd // 2
    inner = np.zeros(raw.shape, bool)
    inner[brd:d - brd, brh:h - brh, brw:w - brw] = True
    out = raw.astype(np.float64)
    if metric == O.NRMSE:
        out = np.where(inner, np.where(out == 0, 1.0, 1.0 / np.where(out == 0, 1.0, out)), out)
    elif metric == O.ABSDIFF:
        out = np.where(inner, 1.0 - out, out)
    sel = inner
    if c > 0:
        nv = float((stds[inner].astype(np.float64) ** 2).mean())
        mask = inner & ((stds.astype(np.float64) ** 2) > nv * c)
        out = out * mask
        sel = mask
    mn, mx = out[sel].min(), out[sel].max()
    return np.where(sel, (out - mn) / (mx - mn + EPS), out)


@pytest.mark.parametrize("metric", [O.PEARSON, O.SSIM, O.NRMSE, O.ABSDIFF])
@pytest.mark.parametrize("c", [0.0, 0.7])
def test_maps_against_numpy(metric, c):
    vol = _vol(12, 16, 18, 5)
    b = (5, 3, 3)
    px, tm, ts = _template(vol, 6, 8, 9, *b, 0.01, 7)
    m32, s32 = O.local_stats3d_f32(vol, *b)
    raw = O.block_raw3d(metric, vol, m32, s32, px, *b, tm, ts)
    got = O.block_map3d(metric, vol, px, *b, tm, ts, c)
    want = _map_numpy(metric, raw, s32, *b, c)
    # the mask decision sits on float sums: compare where it is not borderline
    assert np.abs(got - want).max() <= 1e-5
    with pytest.raises(ValueError):
        O.block_map3d(5, vol, px, *b, tm, ts)
