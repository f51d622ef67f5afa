"""GPU relevance-mask tools against the oracle, through the C ABI: the
shared input step (fit from gain 0, GAT with the fitted doubles, double
normalisation) bit for bit; RelevanceMask2D_'s mask, noise variance and
threshold; OptimiseRelevanceMask_'s constant range (the device reduces the
squared variances once, the oracle rebuilds every mask); the statistics
CalculateLocalMeans_ sets out to compute; the errors."""
import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import synthetic, tools

pytestmark = pytest.mark.gpu


def _image(h, w, seed):
    rng = np.random.default_rng(seed)
    base = synthetic.rings_and_blobs(max(h, w), 30, 30, seed=seed)[:h, :w]
    return (3.0 * rng.poisson(5 + 40 * base / base.max()) + rng.normal(8, 2, (h, w))).astype(np.float32)


@pytest.mark.parametrize("shape", [(64, 80), (128, 128), (97, 61)])
def test_prepare_input_bitexact(ctx, shape):
    img = _image(*shape, seed=shape[0])
    got, p = tools.prepare_input(img, 300, resources=ctx)
    want, q = O.tool_prep(img, 300)
    assert p == tuple(q)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("b,c", [(3, 0.0), (3, 1.5), (5, 0.7), (7, 3.0)])
def test_relevance_mask_tool_bitexact(ctx, b, c):
    img = _image(90, 110, 7)
    mask, nv, thr = tools.relevance_mask_2d(img, b, b, c, max_iter=300, resources=ctx)
    prep, _ = O.tool_prep(img, 300)
    wmask, wnv, wthr = O.relevance_mask_tool(prep, b, b, c)
    assert (nv, thr) == (wnv, wthr)
    np.testing.assert_array_equal(mask, wmask)


@pytest.mark.parametrize("b,step", [(3, 0.001), (5, 0.002), (3, 0.0004)])
def test_optimise_relevance_mask_bitexact(ctx, b, step):
    img = _image(40, 56, 4 + b)
    got = tools.optimise_relevance_mask(img, b, b, step, max_iter=200, resources=ctx)
    prep, _ = O.tool_prep(img, 200)
    assert got == O.optimise_relevance_mask(prep, b, b, step)


def test_image_statistics_and_errors(ctx):
    img = _image(72, 64, 9)
    st = tools.image_statistics_2d(img, 5, 3, max_iter=200, resources=ctx)
    prep, _ = O.tool_prep(img, 200)
    m, s = O.local_stats_f32(prep, 5, 3)
    np.testing.assert_array_equal(st[0], m)
    np.testing.assert_array_equal(st[1], s)
    with pytest.raises(sred_amd.SredError):  # non-square blocks overrun the reference's patch array
        tools.relevance_mask_2d(img, 3, 5, 1.0, max_iter=50, resources=ctx)
    with pytest.raises(sred_amd.SredError):  # the reference loops for ever
        tools.optimise_relevance_mask(img, 3, 3, 0.0, max_iter=50, resources=ctx)
    with pytest.raises(ValueError):
        tools.image_statistics_2d(img, 4, 3, resources=ctx)


@pytest.mark.parametrize("b,lower,upper,step", [(3, 0.0, 1.0, 0.1), (5, 0.25, 3.0, 0.35), (3, 0.0, 40.0, 2.5)])
def test_mask_stack_bitexact(ctx, b, lower, upper, step):
    raw = _image(96, 80, 21)
    masks, consts, prepared = tools.relevance_mask_stack_2d(raw, b, b, lower, upper, step, 300, ctx)
    want_img = O.tool_prep(raw, 300)[0]
    np.testing.assert_array_equal(prepared, want_img)
    wm, wc = O.relevance_mask_stack(want_img, b, b, lower, upper, step)
    np.testing.assert_array_equal(consts, wc)
    np.testing.assert_array_equal(masks, wm)
    with pytest.raises(sred_amd.SredError):
        tools.relevance_mask_stack_2d(raw, 3, 5, 0.0, 1.0, 0.1, 50, ctx)
