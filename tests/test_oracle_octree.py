"""Pins the oracle's octree GAT estimate of a volume (OcTree_.java:136-486,
the "Quad/Octree" branch of Utils.getInputImage3D); no GPU needed.  The
F quantile and SimpleRegression it shares with the quadtree are pinned in
test_oracle_quadtree.py.

* the node criterion: a constant volume gives block variance 0 and noise
  variance ~0 inside (equal residuals); an FP64 numpy restatement (double
  block variance, the 7v - neighbours pseudo-residual over sqrt(56),
  volume-border neighbours dropped, n - 1 denominator) agrees to float
  rounding;
* the robust mean and the LTS variance of a box equal the quadtree's on the
  box's voxels laid out as a (nh*nd) x nw image: the same float sequence in
  z, y, x order;
* the tree: every leaf either cannot divide by size or fails the Fisher test
  with the reference's width*height degrees of freedom, every internal node
  passes it (thresholds from scipy); the leaves tile the volume in
  pre-order (children x fastest, then y, then z);
* the whole estimate statistically: gain-3 Poisson counts on tiles of
  constant intensity give a slope of about 3c, c the LTS factor; g0 and eDC
  are float values;
* applyGATtree with float parameters against its closed form, including the
  zero branch v < -ref/g0 (also for a negative g0, where it differs from
  the quadtree's g0 v + ref < 0);
* the reference's errors: one voxel per slice in a visited node, fewer than
  3 leaves.
"""
import numpy as np
import pytest

import oracle as O

st = pytest.importorskip("scipy.stats")

P01 = float(np.float32(0.01))


def _tiles(d, h, w, t, gain, seed):
    rng = np.random.default_rng(seed)
    lam = np.kron(rng.uniform(20, 400, (d // t, h // t, w // t)), np.ones((t, t, t)))
    return (gain * rng.poisson(lam)).astype(np.float32)


def _criterion64(vol, x0, y0, z0, nw, nh, nd):
    v = vol.astype(np.float64)
    box = v[z0:z0 + nd, y0:y0 + nh, x0:x0 + nw]
    bvar = box.var()
    p = np.pad(v, 1)
    lap = 7 * v
    d, h, w = v.shape
    for dz, dy, dx in [(0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]:
        lap = lap - p[1 + dz:1 + dz + d, 1 + dy:1 + dy + h, 1 + dx:1 + dx + w]
    r = lap[z0:z0 + nd, y0:y0 + nh, x0:x0 + nw] / np.sqrt(56.0)
    return bvar, r.var(ddof=1)


def test_criterion_against_numpy():
    rng = np.random.default_rng(1)
    vol = (rng.poisson(50, (12, 20, 18)) * 2.0 + rng.normal(0, 3, (12, 20, 18))).astype(np.float32)
    for box in [(0, 0, 0, 18, 20, 12), (3, 5, 2, 8, 7, 5), (10, 12, 8, 8, 8, 4), (0, 13, 0, 9, 7, 6)]:
        bv, nv = O.oc_criterion(vol, *box)
        bv64, nv64 = _criterion64(vol, *box)
        assert bv == pytest.approx(bv64, rel=2e-6) and nv == pytest.approx(nv64, rel=2e-5)
    const = np.full((8, 8, 8), 7.0, np.float32)
    bv, nv = O.oc_criterion(const, 2, 2, 2, 4, 4, 4)  # equal residuals 7/sqrt(56) inside: variance ~0
    assert bv == 0.0 and nv == pytest.approx(0.0, abs=1e-10)


def test_leaf_statistics_match_the_quadtree_on_the_flattened_box():
    rng = np.random.default_rng(2)
    vol = (rng.poisson(80, (10, 16, 14)) * 1.5).astype(np.float32)
    vol[3, 4, 5] = 4000.0  # an outlier inside the box
    x0, y0, z0, nw, nh, nd = 2, 3, 1, 8, 6, 5
    flat = vol[z0:z0 + nd, y0:y0 + nh, x0:x0 + nw].reshape(nd * nh, nw).copy()
    h, w = vol.shape[1], vol.shape[2]
    m3 = O.lib().or_oc_robust_mean(vol.reshape(-1), w, h, x0, y0, z0, nw, nh, nd, 50)
    m2 = O.qt_robust_mean(flat, 0, 0, nw, nd * nh)
    assert np.float32(m3) == np.float32(m2)
    v3 = O.lib().or_oc_lts_variance(vol.reshape(-1), w, h, x0, y0, z0, nw, nh, nd, float(m2), 0.75)
    v2 = O.qt_lts_variance(flat, 0, 0, nw, nd * nh, float(m2), 0.75)
    assert np.float32(v3) == np.float32(v2)


def _path_codes(boxes, w, h, d):
    """Octal child path of each leaf box, root-relative (pre-order sorts them)."""
    codes = []
    for x, y, z, bw, bh, bd in boxes:
        cw, ch, cd, cx, cy, cz, path = w, h, d, 0, 0, 0, []
        while (cw, ch, cd) != (bw, bh, bd):
            hw, hh, hd = cw // 2, ch // 2, cd // 2
            bx, by, bz = int(x >= cx + hw), int(y >= cy + hh), int(z >= cz + hd)
            path.append(bx + 2 * by + 4 * bz)
            cx, cy, cz, cw, ch, cd = cx + bx * hw, cy + by * hh, cz + bz * hd, hw, hh, hd
        assert (cx, cy, cz) == (x, y, z)
        codes.append(path)
    return codes


@pytest.mark.parametrize("shape,min_size,alpha", [((32, 48, 40), 4, 0.01), ((24, 33, 29), 3, 0.05),
                                                   ((16, 64, 64), 4, 0.001)])
def test_tree_decisions_and_preorder(shape, min_size, alpha):
    rng = np.random.default_rng(sum(shape))
    d, h, w = shape
    vol = _tiles(d - d % 8, h - h % 8, w - w % 8, 8, 2.0, 3)
    vol = np.pad(vol, ((0, d % 8), (0, h % 8), (0, w % 8)), mode="edge")
    vol = vol + rng.normal(0, 2.0, vol.shape).astype(np.float32)
    r = O.octree_gat(vol, min_size, alpha)
    a = float(np.float32(alpha))
    for x, y, z, bw, bh, bd in r.boxes:
        bv, nv = O.oc_criterion(vol, x, y, z, bw, bh, bd)
        ratio = np.float32(min(bv, nv)) / np.float32(max(bv, nv))
        thr = np.float32(st.f.ppf(a, bw * bh - 1, bw * bh - 1))
        assert bw <= min_size or bh <= min_size or bd <= min_size or not ratio < thr
    codes = _path_codes(r.boxes, w, h, d)
    assert codes == sorted(codes)
    # internal nodes: every proper prefix of a leaf's path passed the test
    seen = set()
    for (x, y, z, bw, bh, bd), path in zip(r.boxes, codes):
        cw, ch, cd, cx, cy, cz = w, h, d, 0, 0, 0
        for c in path:
            if (cx, cy, cz, cw, ch, cd) not in seen:
                seen.add((cx, cy, cz, cw, ch, cd))
                bv, nv = O.oc_criterion(vol, cx, cy, cz, cw, ch, cd)
                ratio = np.float32(min(bv, nv)) / np.float32(max(bv, nv))
                assert ratio < np.float32(st.f.ppf(a, cw * ch - 1, cw * ch - 1))
            hw, hh, hd = cw // 2, ch // 2, cd // 2
            cx, cy, cz = cx + (c & 1) * hw, cy + ((c >> 1) & 1) * hh, cz + (c >> 2) * hd
            cw, ch, cd = hw, hh, hd
    if all(s % 2 == 0 for s in shape):
        covered = np.zeros(vol.shape, np.int32)
        for x, y, z, bw, bh, bd in r.boxes:
            covered[z:z + bd, y:y + bh, x:x + bw] += 1
        assert covered.max() == 1


def test_gain_recovery_and_float_parameters():
    c = st.chi2.cdf(st.chi2.ppf(0.75, 1), 3) / 0.75
    r = O.octree_gat(_tiles(64, 64, 64, 16, 3.0, 7))
    assert r.g0 == pytest.approx(3.0 * c, rel=0.05)
    assert r.g0 == float(np.float32(r.g0)) and r.edc == float(np.float32(r.edc))
    assert len(r.means) == len(r.boxes) >= 3


def test_gat_tree3d_closed_form():
    rng = np.random.default_rng(4)
    v = rng.uniform(-50, 300, (4, 9, 11)).astype(np.float32)
    for g0, edc in [(np.float32(1.3), np.float32(-12.5)), (np.float32(0.6), np.float32(40.0)),
                    (np.float32(-0.8), np.float32(25.0))]:
        g = float(g0)
        ref = 3.0 / 8.0 * g * g + float(edc)
        x = v.astype(np.float64)
        with np.errstate(invalid="ignore"):
            want = np.where(x < -ref / g, 0.0, (2.0 / g) * np.sqrt(g * x + ref)).astype(np.float32)
        np.testing.assert_array_equal(O.gat_tree3d(v, g0, edc), want)


def test_errors():
    with pytest.raises(ValueError):  # a 1 x 1 x d node reaches the F test
        O.octree_gat(np.random.default_rng(0).normal(100, 10, (8, 1, 1)).astype(np.float32), 1)
    # zero-mean noise (no border bias in the pseudo-residuals) stays one leaf: the regression needs 3
    with pytest.raises(ValueError):
        O.octree_gat(np.random.default_rng(0).normal(0, 10, (16, 16, 16)).astype(np.float32))
