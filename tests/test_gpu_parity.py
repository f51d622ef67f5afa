"""GPU parity: the HIP path (through the C ABI) against the oracle.

Bars: bit-exact for statistics, relevance masks and the block kernels'
float pipeline; for global maps the raw accumulators within a relative
tolerance and the normalised maps within 1e-4 absolute of both the float32
restatement (refsim) and the FP64 oracle.
"""
import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from sred_amd import clutils, synthetic, utils
from tests.helpers import kat_reference_block, kats
from tests.test_oracle import checkerboard, checkerboard_closed_form

pytestmark = pytest.mark.gpu
K = kats()
GM = sred_amd.GLOBAL_METRICS
MAP_TOL = 1e-4          # normalised map, absolute (the parity bound of the maps)
S_RTOL = {"fp32": 2e-5, "bf16x3": 2e-4}  # raw accumulators vs FP64 (S of Pearson can be small: atol too)


def stats_gpu(ctx, img, bw, bh):
    inp = utils.InputImage2D(np.ascontiguousarray(img, np.float32), img.shape[1], img.shape[0], img.size)
    st = clutils.get_local_statistics_2d(ctx, inp, bw, bh)
    return st.local_means.reshape(img.shape), st.local_stds.reshape(img.shape)


def test_local_statistics_kat(ctx):
    k = K["local_statistics_2d"]
    m, s = stats_gpu(ctx, np.array(k["image"], np.float32), *k["block"])
    np.testing.assert_allclose(m, np.array(k["expected_means"]), atol=k["tol"])
    np.testing.assert_allclose(s, np.array(k["expected_stds"], np.float32), atol=k["tol"])


@pytest.mark.parametrize("shape,blk", [((64, 64), (5, 5)), ((37, 53), (3, 3)), ((100, 71), (11, 11)),
                                       ((40, 90), (7, 5)), ((130, 129), (17, 17)), ((20, 20), (9, 15))])
def test_local_statistics_bitexact(ctx, shape, blk):
    rng = np.random.default_rng(shape[0] * 1000 + blk[0])
    img = O.normalize_input(rng.random(shape).astype(np.float32))
    img[5:12, 3:9] = img[5, 3]  # flat patch
    m, s = stats_gpu(ctx, img, *blk)
    mo, so = O.local_stats_f32(img, *blk)
    np.testing.assert_array_equal(m, mo)
    np.testing.assert_array_equal(s, so)


def global_inputs(ctx, img, b, c):
    inp = utils.get_input_image_2d(img)
    st = clutils.get_local_statistics_2d(ctx, inp, b, b)
    rm = utils.get_relevance_mask(inp.image_array, inp.width, inp.height, b // 2, b // 2, st.local_stds, c)
    npx = utils.count_relevant_pixels(rm.relevance_mask, inp.width, inp.height, b // 2, b // 2)
    return inp, st, rm, npx


def run_global(ctx, metric, img, b, c, precision, normalize=True):
    inp, st, rm, npx = global_inputs(ctx, img, b, c)
    out = clutils.calculate_global_repetition_map_2d(GM[metric], inp, b, b, sred_amd.block_size_2d(b, b), st, rm, npx,
                                                     normalize, ctx, precision)
    return out.reshape(img.shape), inp, st, rm, npx


def test_relevance_mask_bitexact(ctx):
    img = synthetic.spots(96, 80, 60, seed=9)
    for c in (0.0, 0.7, 2.0):
        inp, st, rm, npx = global_inputs(ctx, img, 5, c)
        mo, so = O.local_stats_f32(inp.image_array, 5, 5)
        mask, thr, np_o = O.relevance_mask(so, 2, 2, O.noise_var(inp.image_array), c)
        np.testing.assert_array_equal(rm.relevance_mask.reshape(mask.shape), mask)
        assert npx == np_o and rm.relevance_threshold == thr


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("metric", [0, 1, 2, 3])
@pytest.mark.parametrize("c", [0.0, 1.0])
def test_global_maps_small(ctx, precision, metric, c):
    img = synthetic.spots(72, 61, 40, seed=metric + 10)
    b = 5
    out, inp, st, rm, npx = run_global(ctx, metric, img, b, c, precision)
    ref32 = O.global_map(metric, inp.image_array, b, b, c, True)
    ref64 = O.global_map(metric, inp.image_array, b, b, c, True, fp64=True)
    assert np.abs(out - ref32).max() <= MAP_TOL
    assert np.abs(out - ref64).max() <= MAP_TOL
    assert (out[:2] == 0).all() and (out[-2:] == 0).all() and (out[:, :2] == 0).all() and (out[:, -2:] == 0).all()


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("metric", [0, 1, 2, 3])
def test_global_accumulators_vs_fp64(ctx, precision, metric):
    img = O.normalize_input(synthetic.spots(64, 64, 30, seed=3))
    b = 7
    m, s = O.local_stats_f32(img, b, b)
    nv = O.noise_var(img)
    inp = utils.InputImage2D(img, 64, 64, 64 * 64)
    S, W = clutils.get_global_accumulators_2d(ctx, metric, inp, clutils.LocalStatistics(m.ravel(), s.ravel()), b, b,
                                              nv, 0.0, precision)
    S64, W64 = O.global_accum_f64(metric, img, s, b, b, nv, 0.0)
    S, W = S.reshape(64, 64), W.reshape(64, 64)
    np.testing.assert_allclose(W, W64, rtol=2e-6, atol=1e-6)
    scale = np.abs(S64).max()
    np.testing.assert_allclose(S, S64, rtol=S_RTOL[precision], atol=S_RTOL[precision] * scale)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("metric", [0, 1, 2, 3])
def test_global_checkerboard_closed_form(ctx, precision, metric):
    w, h, b = 47, 38, 5
    img = O.normalize_input(checkerboard(w, h))
    m, s = O.local_stats_f32(img, b, b)
    inp = utils.InputImage2D(img, w, h, w * h)
    S, W = clutils.get_global_accumulators_2d(ctx, metric, inp, clutils.LocalStatistics(m.ravel(), s.ravel()), b, b,
                                              O.noise_var(img), 0.0, precision)
    Sx, Wx = checkerboard_closed_form(metric, w, h, b)
    np.testing.assert_allclose(W.reshape(h, w), Wx, rtol=1e-6)
    np.testing.assert_allclose(S.reshape(h, w), Sx, rtol=5e-5 if precision == "fp32" else 5e-4, atol=1e-3)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
@pytest.mark.parametrize("metric", [0, 1, 2, 3])
def test_global_constant_image(ctx, precision, metric):
    w, h, b = 40, 33, 5
    out, *_ = run_global(ctx, metric, np.full((h, w), 2.5, np.float32), b, 0.0, precision)
    ref = O.global_map(metric, O.normalize_input(np.full((h, w), 2.5, np.float32)), b, b, 0.0, True)
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_global_flat_patches_special_cases(ctx, precision):
    rng = np.random.default_rng(21)
    img = rng.random((50, 44)).astype(np.float32)
    img[10:25, 5:20] = 0.25   # flat region: sigma == 0 blocks (both-flat and one-flat pairs)
    img[30:45, 25:40] = 0.75
    for metric in (0, 1, 2, 3):
        out, inp, *_ = run_global(ctx, metric, img, 5, 0.0, precision)
        ref64 = O.global_map(metric, inp.image_array, 5, 5, 0.0, True, fp64=True)
        ref32 = O.global_map(metric, inp.image_array, 5, 5, 0.0, True)
        assert np.abs(out - ref32).max() <= MAP_TOL, metric
        assert np.abs(out - ref64).max() <= MAP_TOL, metric


def test_global_empty_interior_and_edge_sizes(ctx):
    # block wider than the image: no interior pixel, the map is all zeros
    img = synthetic.spots(12, 12, 3, seed=1)
    out, *_ = run_global(ctx, 0, img, 13, 0.0, "bf16x3")
    assert (out == 0).all()
    # ragged sizes, tiny and non-square
    for (h, w) in [(9, 9), (17, 31), (130, 67)]:
        img = synthetic.spots(w, h, 10, seed=h)
        for precision in ("fp32", "bf16x3"):
            out, inp, *_ = run_global(ctx, 0, img, 3, 0.0, precision)
            ref = O.global_map(0, inp.image_array, 3, 3, 0.0, True, fp64=True)
            assert np.abs(out - ref).max() <= MAP_TOL


def test_global_errors(ctx):
    img = synthetic.spots(32, 32, 5, seed=2)
    inp, st, rm, npx = global_inputs(ctx, img, 5, 0.0)
    assert clutils.calculate_global_repetition_map_2d("Hausdorff", inp, 5, 5, 13, st, rm, npx, True, ctx) is None
    bad = utils.RelevanceMask(rm.relevance_mask * 0.5, 0.0, 0.0, rm.noise_mean_variance)
    with pytest.raises(sred_amd.NonBinaryMaskError):
        clutils.calculate_global_repetition_map_2d("Pearson's R", inp, 5, 5, 13, st, bad, npx, True, ctx)
    with pytest.raises(ValueError):
        sred_amd.global_repetition_2d(img, 4, 5)
    with pytest.raises(sred_amd.SredError):
        clutils.get_local_statistics_2d(ctx, inp, 6, 5)


def test_buffer_reuse_large_then_small(ctx):
    big = synthetic.spots(200, 150, 80, seed=5)
    small = synthetic.spots(40, 30, 10, seed=6)
    for precision in ("fp32", "bf16x3"):
        run_global(ctx, 0, big, 7, 0.0, precision)
        a, *_ = run_global(ctx, 0, small, 5, 0.0, precision)
        with sred_amd.Context(0) as fresh:
            b, *_ = run_global(fresh, 0, small, 5, 0.0, precision)
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("bsize", [9, 11, 13, 17])
def test_global_block_sizes_mfma_vs_fp32(ctx, bsize):
    img = synthetic.spots(90, 84, 50, seed=bsize)
    a, inp, *_ = run_global(ctx, 0, img, bsize, 0.0, "bf16x3")
    b, *_ = run_global(ctx, 0, img, bsize, 0.0, "fp32")
    ref = O.global_map(0, inp.image_array, bsize, bsize, 0.0, True, fp64=True)
    assert np.abs(a - ref).max() <= MAP_TOL
    assert np.abs(b - ref).max() <= MAP_TOL


def test_global_device_pointer_path(ctx):
    torch = pytest.importorskip("torch")
    img = O.normalize_input(synthetic.spots(128, 96, 50, seed=8))
    d_img = torch.from_numpy(img).cuda()
    outs = {}
    for prec_name, prec in (("fp32", L.FP32), ("bf16x3", L.BF16X3)):
        d_out = torch.zeros_like(d_img)
        d_mask = torch.zeros_like(d_img)
        sred_amd.commands.global_map_device(ctx, 0, prec, d_img.data_ptr(), 128, 96, 5, 5, 0.0, True,
                                            d_out.data_ptr(), d_mask=d_mask.data_ptr())
        torch.cuda.synchronize()
        outs[prec_name] = d_out.cpu().numpy()
        m, s = O.local_stats_f32(img, 5, 5)
        mask, _, _ = O.relevance_mask(s, 2, 2, O.noise_var(img), 0.0)
        np.testing.assert_array_equal(d_mask.cpu().numpy(), mask)
    ref = O.global_map(0, img, 5, 5, 0.0, True, fp64=True)
    for k, v in outs.items():
        assert np.abs(v - ref).max() <= MAP_TOL, k


# ---------------------------------------------------------------------------
# block repetition
# ---------------------------------------------------------------------------
BM = sred_amd.BLOCK_METRICS


@pytest.mark.parametrize("name", ["block_pearson_2d", "block_ssim_2d", "block_nrmse_2d", "block_absdiff_2d"])
def test_block_kats(ctx, name):
    k = K[name]
    px, m, s = kat_reference_block(k["block"], O.footprint)
    rb = utils.ReferenceBlock2D(px, 3, 3, 1, 1, px.size, m, s)
    inp = utils.InputImage2D(np.array(k["image"], np.float32), 10, 10, 100)
    out = clutils.calculate_block_repetition_map_2d(k["metric"], inp, rb, 0.0, True, ctx).reshape(10, 10)
    assert ((out >= 0) & (out <= 1)).all()
    np.testing.assert_allclose(out, np.array(k["expected"], np.float32), atol=k["tol"], rtol=0)


@pytest.mark.parametrize("metric", [0, 1, 2, 3])
@pytest.mark.parametrize("c", [0.0, 1.0])
def test_block_maps_bitexact(ctx, metric, c):
    img = synthetic.rings_and_blobs(256, 40, 40, seed=metric)
    tmpl = synthetic.ring_template(15)
    rb = utils.get_reference_block_2d(tmpl)
    inp = utils.get_input_image_2d(img)
    out = clutils.calculate_block_repetition_map_2d(BM[metric], inp, rb, c, True, ctx).reshape(img.shape)
    ref = O.block_map(metric, inp.image_array, rb.pixels, 15, 15, rb.mean, rb.std, c, True)
    np.testing.assert_array_equal(out, ref)


def test_block_unknown_metric_and_even_block(ctx):
    img = synthetic.spots(30, 30, 5, seed=1)
    inp = utils.get_input_image_2d(img)
    rb = utils.get_reference_block_2d(np.ones((3, 3), np.float32))
    assert clutils.calculate_block_repetition_map_2d("NRMSE", inp, rb, 0.0, True, ctx) is None
    with pytest.raises(ValueError):
        sred_amd.block_repetition_2d(np.ones((4, 3), np.float32), img)


@pytest.mark.parametrize("metric", [0, 1, 2, 3])
def test_block_full_size_2048(ctx, metric):
    # full-size workload: GPU vs oracle on the whole 2048x2048 map
    img = synthetic.rings_and_blobs(2048, 3000, 3000, seed=1)
    tmpl = synthetic.ring_template(15)
    rb = utils.get_reference_block_2d(tmpl)
    inp = utils.get_input_image_2d(img)
    out = clutils.calculate_block_repetition_map_2d(BM[metric], inp, rb, 0.0, True, ctx).reshape(img.shape)
    ref = O.block_map(metric, inp.image_array, rb.pixels, 15, 15, rb.mean, rb.std, 0.0, True)
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("metric", [0, 1, 2])
@pytest.mark.parametrize("shape,b", [((61, 72), 5), ((150, 97), 7), ((300, 260), 11), ((33, 41), 3)])
def test_global_symmetric_matches_full_pairs(ctx, metric, shape, b):
    img = synthetic.spots(shape[1], shape[0], 30, seed=shape[0] + metric)
    img[5:15, 5:15] = 0.3  # flat patch: special cases on both sides of the diagonal
    inp = utils.get_input_image_2d(img)
    m, s = O.local_stats_f32(inp.image_array, b, b)
    nv = O.noise_var(inp.image_array)
    st = clutils.LocalStatistics(m.ravel(), s.ravel())
    res = {}
    try:
        for sym in (1, 0):
            ctx.set_option(L.OPT_SYMMETRIC, sym)
            res[sym] = clutils.get_global_accumulators_2d(ctx, metric, inp, st, b, b, nv, 0.0, "bf16x3")
    finally:
        ctx.set_option(L.OPT_SYMMETRIC, 1)
    S64, W64 = O.global_accum_f64(metric, inp.image_array, s, b, b, nv, 0.0)
    for sym in (1, 0):
        S, W = (a.reshape(shape) for a in res[sym])
        # FP32 pair weights: on these inputs the reference's own FP32 formula
        # exp(-(s_i-s_j)^2/(h+eps)) gives row sums up to 3.8e-6 (relative) off FP64
        np.testing.assert_allclose(W, W64, rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(S, S64, rtol=2e-4, atol=2e-4 * np.abs(S64).max())
    np.testing.assert_allclose(res[1][1], res[0][1], rtol=2e-6)


@pytest.mark.parametrize("metric", [0, 1])
def test_symmetric_contraction_deterministic(ctx, metric):
    """Repeated runs of the double-buffered symmetric kernel (several panels,
    column sums staged through LDS) give bit-identical accumulators."""
    shape, b = (300, 260), 11
    img = synthetic.spots(shape[1], shape[0], 30, seed=shape[0] + metric)
    img[5:15, 5:15] = 0.3
    inp = utils.get_input_image_2d(img)
    m, s = O.local_stats_f32(inp.image_array, b, b)
    nv = O.noise_var(inp.image_array)
    st = clutils.LocalStatistics(m.ravel(), s.ravel())
    runs = [clutils.get_global_accumulators_2d(ctx, metric, inp, st, b, b, nv, 0.0, "bf16x3")[0] for _ in range(4)]
    for r in runs[1:]:
        np.testing.assert_array_equal(r, runs[0])


@pytest.mark.parametrize("metric", [0, 1])
def test_global_skips_vanishing_weights(ctx, metric):
    """Two textures of very different block sigma (low-noise background |
    uniform random texture): the symmetric kernel skips the cross pairs,
    whose weights are below 2^-40, and the accumulators and the normalised
    map still match FP64."""
    rng = np.random.default_rng(17)
    h, w, b = 160, 160, 5
    img = (0.1 + 0.02 * rng.standard_normal((h, w))).astype(np.float32)
    img[:, w // 2:] = rng.random((h, w - w // 2)).astype(np.float32)
    inp = utils.get_input_image_2d(img)
    m, s = O.local_stats_f32(inp.image_array, b, b)
    nv = O.noise_var(inp.image_array)
    st = clutils.LocalStatistics(m.ravel(), s.ravel())
    S, W = clutils.get_global_accumulators_2d(ctx, metric, inp, st, b, b, nv, 0.0, "bf16x3")
    n = (h - b + 1) * (w - b + 1)
    skipped = ctx.last_skipped()
    assert skipped > 0.3 * n * n, skipped
    S64, W64 = O.global_accum_f64(metric, inp.image_array, s, b, b, nv, 0.0)
    np.testing.assert_allclose(W.reshape(h, w), W64, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(S.reshape(h, w), S64, rtol=2e-4, atol=2e-4 * np.abs(S64).max())
    out = sred_amd.global_repetition_2d(img, b, b, 0.0, GM[metric], resources=ctx)
    assert ctx.last_skipped() > 0.3 * n * n
    ref = O.global_map(metric, inp.image_array, b, b, 0.0, True, fp64=True)
    assert np.abs(out - ref).max() <= MAP_TOL
