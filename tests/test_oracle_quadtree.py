"""Pins the oracle's quadtree GAT estimate (QuadTree_.java:135-427, called by
Utils.getInputImage2D, Utils.java:686-703); no GPU needed.

* the third-party pieces (commons-math3 3.6.1, not in the reference tree):
  the regularised incomplete beta against scipy.special.betainc, the F
  quantile against scipy.stats.f.ppf and closed forms (F(2,2): x/(1+x);
  the median of F(d,d) is 1), SimpleRegression against numpy's least
  squares and an exact line;
* the node criterion: a constant image gives block variance 0 and noise
  variance 0 inside, and an FP64 numpy restatement (the 5z - neighbours
  pseudo-residual with the sqrt(30) scale, image-border neighbours dropped,
  n - 1 denominator) agrees to float rounding;
* the robust mean: a constant leaf returns the constant; outliers far from a
  bulk leave the bulk's mean; an FP64 numpy restatement of the iteration;
* the LTS variance: a closed form on a two-valued leaf;
* the tree: every leaf either cannot divide by size or fails the Fisher
  test, every internal node passes it (thresholds from scipy), leaves tile
  the covered region in pre-order;
* the whole estimate statistically: Poisson counts with gain g on tiles of
  constant intensity give a slope of about c g, c = E[Z^2 | Z^2 <= q] with q
  the 75% point of chi^2_1 (the LTS trimming), and the transform of
  applyGATtree against its closed form.
"""
import numpy as np
import pytest

import oracle as O

st = pytest.importorskip("scipy.stats")
sp = pytest.importorskip("scipy.special")

P01 = float(np.float32(0.01))  # the reference passes 0.01f, widened to double


def test_beta_and_f_quantile_against_scipy():
    for a, b, x in [(1, 1, 0.3), (2.5, 3.5, 0.4), (10, 10, 0.45), (500, 500, 0.49), (3, 1.5, 0.9), (0.5, 0.5, 0.2)]:
        assert O.beta_reg(a, b, x) == pytest.approx(sp.betainc(a, b, x), rel=1e-12)
    for d in [1, 2, 3, 15, 63, 255, 1023, 16383, 262143]:
        assert O.f_quantile(d, d, P01) == pytest.approx(st.f.ppf(P01, d, d), rel=1e-11)
    for d1, d2, p in [(3, 7, 0.1), (20, 4, 0.9), (7, 12, 0.2)]:
        assert O.f_quantile(d1, d2, p) == pytest.approx(st.f.ppf(p, d1, d2), rel=1e-11)
    for p in (0.05, 0.3, 0.8):
        assert O.f_quantile(2, 2, p) == pytest.approx(p / (1 - p), rel=1e-14)
    assert O.f_quantile(40, 40, 0.5) == pytest.approx(1.0, abs=1e-14)


def test_simple_regression():
    assert O.simple_regression([1.0, 2.0, 4.0], [5.0, 7.0, 11.0]) == (2.0, 3.0)
    rng = np.random.default_rng(1)
    x = rng.random(500) * 100
    y = 0.37 * x + 12.5 + rng.normal(0, 2, 500)
    s, i = O.simple_regression(x, y)
    ps, pi = np.polyfit(x, y, 1)
    assert s == pytest.approx(ps, rel=1e-12) and i == pytest.approx(pi, rel=1e-12)
    assert np.isnan(O.simple_regression([3.0, 3.0, 3.0], [1.0, 2.0, 4.0])[0])  # sumXX = 0


def _criterion64(img, x0, y0, nw, nh):
    """computeCriterion in FP64 numpy (QuadTree_.java:170-232)."""
    d = img.astype(np.float64)
    h, w = d.shape
    blk = d[y0:y0 + nh, x0:x0 + nw]
    bvar = (blk * blk).mean() - blk.mean() ** 2
    pad = np.pad(d, 1)
    present = np.pad(np.ones_like(d), 1)
    lap = 5.0 * d
    for dy, dx in [(0, -1), (0, 1), (-1, 0), (1, 0)]:
        lap -= (pad * present)[1 + dy:1 + dy + h, 1 + dx:1 + dx + w]
    r = (lap / np.sqrt(30.0))[y0:y0 + nh, x0:x0 + nw]
    return bvar, ((r - r.mean()) ** 2).sum() / (r.size - 1)


def test_criterion_closed_form_and_fp64():
    img = np.ones((40, 40), np.float32)
    bv, nv = O.qt_criterion(img, 8, 8, 16, 16)  # interior: every residual is (5-4)/sqrt(30)
    assert bv == 0.0 and nv < 1e-12
    rng = np.random.default_rng(2)
    img = (20 + 5 * rng.standard_normal((37, 53)) + np.linspace(0, 30, 53)[None, :]).astype(np.float32)
    for x0, y0, nw, nh in [(0, 0, 53, 37), (0, 0, 26, 18), (26, 18, 26, 18), (13, 9, 13, 9), (40, 30, 4, 4)]:
        bv, nv = O.qt_criterion(img, x0, y0, nw, nh)
        b64, n64 = _criterion64(img, x0, y0, nw, nh)
        assert bv == pytest.approx(b64, rel=2e-4) and nv == pytest.approx(n64, rel=1e-4)


def _robust_mean64(v, max_iter=50):
    mean = 0.0
    var = 0.0
    for i, x in enumerate(v):
        mean += x
        var += (x - mean) ** 2
    mean /= len(v)
    var = max(var / (len(v) - 1), 1e-7)
    for _ in range(max_iter):
        prev = mean
        w = np.exp(-((v - mean) ** 2) / (2 * var))
        mean = (w * v).sum() / max(w.sum(), 1e-7)
        if abs(mean - prev) <= 1e-5:
            break
    return mean


def test_robust_mean():
    img = np.full((16, 16), 7.25, np.float32)
    assert O.qt_robust_mean(img, 0, 0, 16, 16) == 7.25
    rng = np.random.default_rng(3)
    img = (50 + rng.standard_normal((32, 32))).astype(np.float32)
    img[::7, ::5] = 400.0  # far outliers
    # reference quirk (QuadTree_.java:277-284): the one-pass variance adds
    # (v - running sum)^2, so it is of the order of the squared sum, every
    # Leclerc weight is ~1 and the "robust" mean is the plain mean
    m = O.qt_robust_mean(img, 0, 0, 32, 32)
    assert m == pytest.approx(img.astype(np.float64).mean(), abs=1e-3)
    assert m == pytest.approx(_robust_mean64(img.astype(np.float64).ravel()), abs=1e-3)
    # values alternating +-0.1 keep the running sum small and a last outlier
    # adds nothing to the one-pass variance: here the weights matter
    img = np.tile(np.array([0.1, -0.1], np.float32), 8).reshape(4, 4)
    img[3, 3] = 3.0
    v64 = img.astype(np.float32).astype(np.float64).ravel()
    m = O.qt_robust_mean(img, 0, 0, 4, 4)
    assert m == pytest.approx(_robust_mean64(v64), abs=1e-4)
    assert abs(m - v64.mean()) > 0.05


def test_lts_variance_closed_form():
    # 16 pixels: 8 at m + 1, 8 at m + 3: residuals 1 and 9; h = floor(0.75 * 16) = 12
    img = np.zeros((4, 4), np.float32)
    img.ravel()[:8] = 11.0
    img.ravel()[8:] = 13.0
    v = O.qt_lts_variance(img, 0, 0, 4, 4, 10.0, 0.75)
    assert v == np.float32(np.float32(8 * 1 + 4 * 9) / np.float32(np.float32(12) + np.float32(1e-7)))
    assert O.qt_lts_variance(img, 0, 0, 4, 4, 10.0, 0.0) == 0.0


def _tile_image(tile, nt, gain, seed, offset=0.0):
    rng = np.random.default_rng(seed)
    lam = rng.uniform(20, 400, (nt, nt))
    lam = np.kron(lam, np.ones((tile, tile)))
    return (gain * rng.poisson(lam) + offset).astype(np.float32)


def _leaf_checks(img, res, min_size=4, alpha=P01):
    h, w = img.shape
    covered = np.zeros((h, w), np.int32)
    thr = {}

    def threshold(n):
        if n not in thr:
            thr[n] = np.float32(st.f.ppf(alpha, n - 1, n - 1))
        return thr[n]

    def passes(x, y, nw, nh):
        bv, nv = O.qt_criterion(img, x, y, nw, nh)
        lo, hi = min(bv, nv), max(bv, nv)
        return np.float32(lo) / np.float32(hi) < threshold(nw * nh)

    for x, y, nw, nh in res.rects:
        covered[y:y + nh, x:x + nw] += 1
        if nw > min_size and nh > min_size:
            assert not passes(x, y, nw, nh)
    assert covered.max() == 1
    # pre-order: a leaf of level k is reached by one child choice per level
    # (a child's origin is its parent's plus the level's size W_l when it is
    # on the right, H_l below); the path code (2*ybit + xbit per level, top-
    # left first) extended to the deepest level must increase along the list
    sizes = [(w, h)]
    while sizes[-1][0] > min_size and sizes[-1][1] > min_size:
        sizes.append((sizes[-1][0] // 2, sizes[-1][1] // 2))
    depth = len(sizes) - 1

    def path(x, y, nw, nh):
        k = sizes.index((nw, nh))
        code, rx, ry = 0, x, y
        for l in range(1, k + 1):
            bx, by = int(rx >= sizes[l][0]), int(ry >= sizes[l][1])
            rx, ry = rx - bx * sizes[l][0], ry - by * sizes[l][1]
            code = code * 4 + 2 * by + bx
        assert (rx, ry) == (0, 0)
        return code * 4 ** (depth - k)

    codes = [path(*r) for r in res.rects.tolist()]
    assert codes == sorted(codes) and len(set(codes)) == len(codes)
    return covered


def test_tree_decisions_against_scipy_thresholds():
    img = _tile_image(32, 4, 2.0, seed=5)  # 128 x 128, 4 x 4 tiles
    res = O.quadtree_gat(img)
    covered = _leaf_checks(img, res)
    assert covered.all()  # 128 is a power of two: nothing dropped
    # internal nodes: every ancestor of a leaf divided (its criterion passes)
    seen = set()
    for x, y, nw, nh in res.rects:
        s, cx, cy = 1, 0, 0
        while (128 >> (s - 1)) > nw:
            size = 128 >> (s - 1)
            cx, cy = (x // size) * size, (y // size) * size
            if (cx, cy, size) not in seen:
                seen.add((cx, cy, size))
                bv, nv = O.qt_criterion(img, cx, cy, size, size)
                t = np.float32(st.f.ppf(P01, size * size - 1, size * size - 1))
                assert np.float32(min(bv, nv)) / np.float32(max(bv, nv)) < t
            s += 1
    # odd sizes drop the last column/row of an odd node
    img = _tile_image(10, 5, 1.5, seed=6)[:47, :45]
    res = O.quadtree_gat(img)
    covered = _leaf_checks(img, res)
    assert covered[:, -1].sum() == 0 and covered[-1, :].sum() == 0


def test_estimate_recovers_gain_statistically():
    # Poisson counts times gain on 64 x 64 tiles: var = gain * mean per tile;
    # the LTS variance keeps the smallest 75% squared residuals, scaling the
    # variance by c = P(chi2_3 <= q) / 0.75 with q the 0.75 point of chi2_1
    q = st.chi2.ppf(0.75, 1)
    c = st.chi2.cdf(q, 3) / 0.75
    for gain in (1.0, 3.0):
        img = _tile_image(64, 8, gain, seed=7)  # 512 x 512
        res = O.quadtree_gat(img)
        assert res.g0 == pytest.approx(c * gain, rel=0.08)
        assert len(res.means) >= 64


def test_gat_tree_closed_form_and_errors():
    x = np.array([-100.0, 0.0, 3.5, 250.0], np.float32)
    g0, edc = 0.8, -4.0
    ref = 3.0 / 8.0 * g0 * g0 + edc
    v = x.astype(np.float64)
    want = np.where(g0 * v + ref < 0, 0.0, 2.0 / g0 * np.sqrt(np.maximum(g0 * v + ref, 0))).astype(np.float32)
    np.testing.assert_array_equal(O.gat_tree(x, g0, edc), want)
    bad = np.ones((16, 16), np.float32)
    bad[3, 4] = np.nan
    with pytest.raises(ValueError) as e:
        O.quadtree_gat(bad)
    assert e.value.args[0] == -5
    with pytest.raises(ValueError) as e:  # a 4-wide image is one leaf: SimpleRegression.regress needs 3
        O.quadtree_gat(np.arange(64, dtype=np.float32).reshape(16, 4))
    assert e.value.args[0] == -1
    with pytest.raises(ValueError):
        O.quadtree_gat(np.ones((1, 1), np.float32))
