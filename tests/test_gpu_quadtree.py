"""GPU quadtree GAT estimate (the "Quad/Octree" branch of
Utils.getInputImage2D, QuadTree_.java:135-427) against the oracle, through
the C ABI: the leaves (pre-order rectangles), their robust means and LTS
variances bit for bit, g0 and eDC identical; applyGATtree bit-exact; the
stabilised input image identical; the reference's error cases."""
import ctypes

import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from sred_amd import synthetic, utils

pytestmark = pytest.mark.gpu

P01 = float(np.float32(0.01))


def structured(h, w, seed, gain=2.0, sigma=3.0, offset=10.0):
    """Poisson counts of rings and blobs plus Gaussian read noise."""
    s = max(h, w)
    base = synthetic.rings_and_blobs(s, seed=seed)[:h, :w].astype(np.float64)
    rng = np.random.default_rng(seed)
    lam = 5.0 + 60.0 * base / max(base.max(), 1e-6)
    return (gain * rng.poisson(lam) + rng.normal(offset, sigma, (h, w))).astype(np.float32)


def run_device(ctx, img, min_size=4, alpha=0.01, max_iter=50, frac=0.75):
    h, w = img.shape
    cap = max(16, w * h // 2)
    rects = np.zeros((cap, 4), np.int32)
    means = np.zeros(cap, np.float32)
    var = np.zeros(cap, np.float32)
    ge = (ctypes.c_double * 2)()
    n = ctypes.c_int()
    L.check(L.load().sred_quadtree_gat_2d(ctx.handle, L.fptr(L.f32(img)), w, h, min_size, alpha, max_iter, frac, ge,
                                          ctypes.byref(n), cap, rects.ctypes.data, L.fptr(means), L.fptr(var)))
    k = n.value
    return rects[:k], means[:k], var[:k], ge[0], ge[1]


@pytest.mark.parametrize("shape,min_size,alpha,max_iter,frac", [
    ((128, 128), 4, 0.01, 50, 0.75),
    ((256, 192), 4, 0.01, 50, 0.75),
    ((97, 131), 4, 0.01, 50, 0.75),     # odd sides drop columns and rows
    ((160, 160), 2, 0.05, 20, 0.5),
    ((300, 260), 8, 0.001, 50, 1.0),
    ((512, 512), 4, 0.01, 50, 0.75),
])
def test_quadtree_estimate_bitexact(ctx, shape, min_size, alpha, max_iter, frac):
    img = structured(*shape, seed=sum(shape))
    rects, means, var, g0, edc = run_device(ctx, img, min_size, alpha, max_iter, frac)
    ref = O.quadtree_gat(img, min_size, alpha, max_iter, frac)
    print(f"quadtree {shape} min {min_size}: {len(ref.means)} leaves, g0 {ref.g0:.6g} eDC {ref.edc:.6g}", flush=True)
    np.testing.assert_array_equal(rects, ref.rects)
    np.testing.assert_array_equal(means, ref.means)
    np.testing.assert_array_equal(var, ref.variances)
    assert (g0, edc) == (ref.g0, ref.edc)


def test_quadtree_full_size_and_tiles(ctx):
    # the benchmark's 1024 x 1024 image, and constant-intensity Poisson tiles
    # where the estimate recovers the gain statistically (c g, c the LTS factor)
    img = structured(1024, 1024, seed=1)
    rects, means, var, g0, edc = run_device(ctx, img)
    ref = O.quadtree_gat(img)
    np.testing.assert_array_equal(rects, ref.rects)
    np.testing.assert_array_equal(means, ref.means)
    np.testing.assert_array_equal(var, ref.variances)
    assert (g0, edc) == (ref.g0, ref.edc)
    st = pytest.importorskip("scipy.stats")
    c = st.chi2.cdf(st.chi2.ppf(0.75, 1), 3) / 0.75
    rng = np.random.default_rng(7)
    tiles = (3.0 * rng.poisson(np.kron(rng.uniform(20, 400, (8, 8)), np.ones((64, 64))))).astype(np.float32)
    _, _, _, g0, _ = run_device(ctx, tiles)
    assert g0 == pytest.approx(3.0 * c, rel=0.08)


def test_gat_tree_and_stabilised_image(ctx):
    img = structured(150, 170, seed=4)
    img[:2] = -800.0  # zero branch
    for g0, edc in [(0.74, -20.0), (1.3, 5.5)]:
        got = utils.QuadTree(170, 150, 4, 0.01, ctx).apply_gat_tree(img, img.size, g0, edc)
        np.testing.assert_array_equal(got, O.gat_tree(img, g0, edc).reshape(-1))
    got = sred_amd.get_input_image_2d(img, stabilise_noise_variance=True, gat_method="Quad/Octree",
                                      resources=ctx).image_array
    ref = O.quadtree_gat(img)
    want = O.normalize_input(O.gat_tree(img, ref.g0, ref.edc))
    np.testing.assert_array_equal(got, want)
    # the mirror's step-by-step interface gives the same pairs and line
    qt = utils.QuadTree(170, 150, 4, 0.01, ctx)
    qt.build_tree(img)
    qt.calculate_robust_means(img, 50)
    qt.calculate_lts_variances(img, 0.75)
    pairs = qt.collect_mean_variance_pairs()
    assert pairs == [[float(m), float(v)] for m, v in zip(ref.means, ref.variances)]
    assert qt.perform_linear_regression(pairs) == [ref.g0, ref.edc]
    np.testing.assert_array_equal(qt.leaves(), ref.rects)


def test_quadtree_device_entry_and_timing(ctx):
    torch = pytest.importorskip("torch")
    img = structured(200, 240, seed=9)
    d = torch.from_numpy(img).cuda()
    ge = (ctypes.c_double * 2)()
    n = ctypes.c_int()
    L.check(L.load().sred_quadtree_gat_2d_dev(ctx.handle, d.data_ptr(), 240, 200, 4, 0.01, 50, 0.75, ge,
                                              ctypes.byref(n), 0, None, None, None))
    ref = O.quadtree_gat(img)
    assert (ge[0], ge[1]) == (ref.g0, ref.edc) and n.value == len(ref.means)
    ms, launches, units = ctx.last_contraction()
    assert ms > 0.0 and launches == 1 and units > 0
    L.check(L.load().sred_gat_tree_2d_dev(ctx.handle, d.data_ptr(), img.size, ge[0], ge[1]))
    np.testing.assert_array_equal(d.cpu().numpy(), O.gat_tree(img, ref.g0, ref.edc))


def test_quadtree_errors(ctx):
    bad = structured(64, 64, seed=2)
    bad[10, 20] = np.inf
    with pytest.raises(sred_amd.SredError) as e:
        run_device(ctx, bad)
    assert e.value.code == L.SRED_ERR_NONFINITE
    with pytest.raises(sred_amd.SredError):  # one pixel: FDistribution(0, 0)
        run_device(ctx, np.ones((1, 1), np.float32))
    with pytest.raises(sred_amd.SredError):  # min_size < 1
        run_device(ctx, structured(32, 32, seed=3), min_size=0)
    # a 4-wide image is a single leaf: the regression needs 3 (SimpleRegression.regress)
    thin = structured(40, 4, seed=5)
    qt = utils.QuadTree(4, 40, 4, 0.01, ctx)
    qt.build_tree(thin)
    pairs = qt.collect_mean_variance_pairs()
    ref_m = O.qt_robust_mean(thin, 0, 0, 4, 40)
    assert len(pairs) == 1 and pairs[0][0] == float(ref_m)
    with pytest.raises(ValueError):
        qt.perform_linear_regression(pairs)
    with pytest.raises(sred_amd.SredError):
        run_device(ctx, thin)
