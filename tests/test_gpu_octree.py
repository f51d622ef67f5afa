"""GPU octree GAT estimate of a volume (OcTree_.java:136-486, the "Quad/Octree"
branch of Utils.getInputImage3D) against the oracle, through the C ABI: the
leaves (pre-order boxes), their robust means and LTS variances bit for bit,
g0 and eDC identical (float values); applyGATtree bit-exact; the stabilised
input volume identical; the mirror's step-by-step interface; the reference's
error cases."""
import ctypes

import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from sred_amd import utils

pytestmark = pytest.mark.gpu


def structured(d, h, w, seed, gain=2.0, sigma=3.0, offset=10.0, n=40):
    """Poisson counts of Gaussian spheres plus Gaussian read noise."""
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.mgrid[0:d, 0:h, 0:w]
    lam = np.full((d, h, w), 5.0)
    for _ in range(n):
        c = rng.uniform(0, [d, h, w])
        lam += rng.uniform(20, 80) * np.exp(-((zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2) / (2 * 2.5 ** 2))
    return (gain * rng.poisson(lam) + rng.normal(offset, sigma, (d, h, w))).astype(np.float32)


def run_device(ctx, vol, min_size=4, alpha=0.01, max_iter=50, frac=0.75):
    d, h, w = vol.shape
    cap = max(16, vol.size // 2)
    boxes = np.zeros((cap, 6), np.int32)
    means = np.zeros(cap, np.float32)
    var = np.zeros(cap, np.float32)
    ge = (ctypes.c_double * 2)()
    n = ctypes.c_int()
    L.check(L.load().sred_octree_gat_3d(ctx.handle, L.fptr(L.f32(vol)), w, h, d, min_size, alpha, max_iter, frac, ge,
                                        ctypes.byref(n), cap, boxes.ctypes.data, means.ctypes.data, var.ctypes.data))
    k = n.value
    return boxes[:k], means[:k], var[:k], ge[0], ge[1]


@pytest.mark.parametrize("shape,min_size,alpha,max_iter,frac", [
    ((32, 64, 64), 4, 0.01, 50, 0.75),
    ((24, 40, 56), 4, 0.01, 50, 0.75),
    ((21, 37, 45), 3, 0.05, 20, 0.5),     # odd sides drop columns, rows and slices
    ((16, 96, 80), 2, 0.001, 50, 1.0),
    ((64, 128, 128), 4, 0.01, 50, 0.75),
])
def test_octree_estimate_bitexact(ctx, shape, min_size, alpha, max_iter, frac):
    vol = structured(*shape, seed=sum(shape))
    print(f"octree {shape} min {min_size}", flush=True)
    boxes, means, var, g0, edc = run_device(ctx, vol, min_size, alpha, max_iter, frac)
    ref = O.octree_gat(vol, min_size, alpha, max_iter, frac)
    print(f"  {len(ref.means)} leaves, g0 {ref.g0:.6g} eDC {ref.edc:.6g}", flush=True)
    np.testing.assert_array_equal(boxes, ref.boxes)
    np.testing.assert_array_equal(means, ref.means)
    np.testing.assert_array_equal(var, ref.variances)
    assert (g0, edc) == (ref.g0, ref.edc)


def test_tiles_recover_the_gain(ctx):
    st = pytest.importorskip("scipy.stats")
    c = st.chi2.cdf(st.chi2.ppf(0.75, 1), 3) / 0.75
    rng = np.random.default_rng(7)
    lam = np.kron(rng.uniform(20, 400, (4, 4, 4)), np.ones((16, 16, 16)))
    tiles = (3.0 * rng.poisson(lam)).astype(np.float32)
    boxes, means, var, g0, edc = run_device(ctx, tiles)
    ref = O.octree_gat(tiles)
    np.testing.assert_array_equal(boxes, ref.boxes)
    assert (g0, edc) == (ref.g0, ref.edc)
    assert g0 == pytest.approx(3.0 * c, rel=0.05)


def test_gat_tree_and_stabilised_volume(ctx):
    vol = structured(12, 40, 36, seed=4)
    vol[:2] = -900.0  # zero branch
    for g0, edc in [(np.float32(0.74), np.float32(-20.0)), (np.float32(1.3), np.float32(5.5)),
                    (np.float32(-0.9), np.float32(30.0))]:
        got = utils.OcTree(36, 40, 12, 4, 0.01, ctx).apply_gat_tree(vol, vol.size, g0, edc)
        np.testing.assert_array_equal(got, O.gat_tree3d(vol, g0, edc).reshape(-1))
    inp = sred_amd.get_input_image_3d(vol, True, "Quad/Octree", resources=ctx)
    ref = O.octree_gat(vol)
    want = O.normalize_input(O.gat_tree3d(vol, ref.g0, ref.edc))
    np.testing.assert_array_equal(inp.image_array, want)
    # the mirror's step-by-step interface gives the same pairs, line and boxes
    oc = utils.OcTree(36, 40, 12, 4, 0.01, ctx)
    oc.build_tree(vol)
    oc.calculate_robust_means(vol, 50)
    oc.calculate_lts_variances(vol, 0.75)
    pairs = oc.collect_mean_variance_pairs()
    assert pairs == [[float(m), float(v)] for m, v in zip(ref.means, ref.variances)]
    assert oc.perform_linear_regression(pairs) == [ref.g0, ref.edc]
    np.testing.assert_array_equal(oc.leaves(), ref.boxes)
    # a stabilised volume through a 3D command
    vol = structured(12, 40, 36, seed=5)
    got = sred_amd.global_repetition_3d(vol, 3, 3, 3, 0.0, "Pearson's R", True, "fp32", resources=ctx,
                                        stabilise_noise_variance=True, gat_method="Quad/Octree")
    ref = O.octree_gat(vol)
    want = O.normalize_input(O.gat_tree3d(vol, ref.g0, ref.edc))
    np.testing.assert_allclose(got, O.global_pearson3d(want, 3, 3, 3, fp64=True).out, atol=1e-4)


def test_device_entry_and_timing(ctx):
    torch = pytest.importorskip("torch")
    vol = structured(20, 48, 40, seed=9)
    d, h, w = vol.shape
    dv = torch.from_numpy(vol).cuda()
    ge = (ctypes.c_double * 2)()
    n = ctypes.c_int()
    L.check(L.load().sred_octree_gat_3d_dev(ctx.handle, dv.data_ptr(), w, h, d, 4, 0.01, 50, 0.75, ge,
                                            ctypes.byref(n), 0, None, None, None))
    ref = O.octree_gat(vol)
    assert (ge[0], ge[1]) == (ref.g0, ref.edc) and n.value == len(ref.means)
    ms, launches, units = ctx.last_contraction()
    assert ms > 0.0 and launches == 1 and units > 0
    L.check(L.load().sred_gat_tree_3d_dev(ctx.handle, dv.data_ptr(), vol.size, ge[0], ge[1]))
    np.testing.assert_array_equal(dv.cpu().numpy(), O.gat_tree3d(vol, ref.g0, ref.edc))


def test_errors(ctx):
    with pytest.raises(sred_amd.SredError):  # one pixel per slice: FDistribution(0, 0)
        run_device(ctx, np.random.default_rng(0).normal(100, 10, (8, 1, 1)).astype(np.float32), 1)
    with pytest.raises(sred_amd.SredError):  # min_size < 1
        run_device(ctx, structured(8, 16, 16, seed=3), min_size=0)
    flat = np.random.default_rng(0).normal(0, 10, (16, 16, 16)).astype(np.float32)
    with pytest.raises(sred_amd.SredError):  # one leaf: the regression needs 3
        run_device(ctx, flat)
    oc = utils.OcTree(16, 16, 16, 4, 0.01, ctx)
    oc.build_tree(flat)
    assert len(oc.collect_mean_variance_pairs()) == 1
    with pytest.raises(ValueError):
        oc.perform_linear_regression(oc.collect_mean_variance_pairs())


def test_tree_plugins(ctx):
    img = structured(1, 150, 170, seed=12)[0]
    out, (g0, edc), rects = sred_amd.quadtree_stabilise_2d(img, 4, 0.01, 50, 0.5, resources=ctx)
    ref = O.quadtree_gat(img, 4, 0.01, 50, 0.5)
    assert (g0, edc) == (ref.g0, ref.edc)
    np.testing.assert_array_equal(rects, ref.rects)
    np.testing.assert_array_equal(out, O.gat_tree(img, ref.g0, ref.edc))
    vol = structured(12, 40, 36, seed=13)
    out, (g0, edc), boxes = sred_amd.octree_stabilise_3d(vol, 4, 0.01, 50, 0.5, resources=ctx)  # 0.5 is ignored
    ref = O.octree_gat(vol, 4, 0.01, 50, 0.75)
    assert (g0, edc) == (ref.g0, ref.edc)
    np.testing.assert_array_equal(boxes, ref.boxes)
    np.testing.assert_array_equal(out, O.gat_tree3d(vol, ref.g0, ref.edc))
    with pytest.raises(ValueError):
        sred_amd.octree_stabilise_3d(vol[:2], resources=ctx)
