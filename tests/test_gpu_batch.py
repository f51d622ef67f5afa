"""Batched global repetition (sred_global_batch_2d): images of uneven sizes
and block sizes in one graph launch, against the single-image path and the
FP64 oracle."""
import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from sred_amd import synthetic
from sred_amd.commands import global_batch_device, global_map_device

pytestmark = pytest.mark.gpu
MAP_TOL = 1e-4

SHAPES = [((61, 72), 5), ((97, 150), 9), ((64, 64), 5), ((80, 70), 7), ((120, 100), 13), ((40, 45), 3),
          ((130, 90), 9)]


def images(seed0=20):
    return [O.normalize_input(synthetic.spots(s[1], s[0], 25, seed=seed0 + i)) for i, (s, _) in enumerate(SHAPES)]


def run_batch(ctx, metric, imgs, c=0.0):
    torch = pytest.importorskip("torch")
    d_imgs = [torch.from_numpy(im).cuda() for im in imgs]
    d_outs = [torch.full_like(d, 7.0) for d in d_imgs]
    probs = [(d.data_ptr(), im.shape[1], im.shape[0], b, b, o.data_ptr())
             for d, im, o, (_, b) in zip(d_imgs, imgs, d_outs, SHAPES)]
    global_batch_device(ctx, metric, probs, c, True)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in d_outs]


def run_single(ctx, metric, im, b, c=0.0):
    torch = pytest.importorskip("torch")
    d = torch.from_numpy(im).cuda()
    o = torch.zeros_like(d)
    global_map_device(ctx, metric, L.BF16X3, d.data_ptr(), im.shape[1], im.shape[0], b, b, c, True, o.data_ptr())
    torch.cuda.synchronize()
    return o.cpu().numpy()


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_batch_matches_single_and_oracle(metric):
    imgs = images()
    with sred_amd.Context(0) as ctx:
        outs = run_batch(ctx, metric, imgs)
        for (shape, b), im, out in zip(SHAPES, imgs, outs):
            single = run_single(ctx, metric, im, b)
            assert np.abs(out - single).max() <= 2e-6, (shape, b)
            ref = O.global_map(metric, im, b, b, 0.0, True, fp64=True)
            assert np.abs(out - ref).max() <= MAP_TOL, (shape, b)
        # every class got a panel size from the autotuner
        for ks in {(sred_amd.block_size_2d(b, b) + 31) // 32 for _, b in SHAPES}:
            assert ctx.tuned_panel(metric, ks) in (256, 512, 1024, 2048)


def test_batch_forced_panel_is_the_single_path_bit_for_bit():
    im = images()[1]
    b = SHAPES[1][1]
    with sred_amd.Context(0) as ctx:
        nv = (im.shape[0] - b + 1) * (im.shape[1] - b + 1)
        P = L.global_plan(nv, sred_amd.block_size_2d(b, b), 0, L.BF16X3)["rows"]
        ctx.set_option(L.OPT_BATCH_PANEL, P)
        torch = pytest.importorskip("torch")
        d = torch.from_numpy(im).cuda()
        o = torch.zeros_like(d)
        global_batch_device(ctx, 0, [(d.data_ptr(), im.shape[1], im.shape[0], b, b, o.data_ptr())])
        torch.cuda.synchronize()
        np.testing.assert_array_equal(o.cpu().numpy(), run_single(ctx, 0, im, b))


def test_batch_waves_repeat_and_relevance():
    imgs = images(40)
    with sred_amd.Context(0) as ctx:
        ctx.set_option(L.OPT_BATCH_PANEL, 256)
        one = run_batch(ctx, 0, imgs, c=0.5)
        ctx.set_option(L.OPT_BATCH_MEMORY_MB, 1)  # several memory waves inside the graph
        waves = run_batch(ctx, 0, imgs, c=0.5)
        again = run_batch(ctx, 0, imgs, c=0.5)
        for a, b_, c_ in zip(one, waves, again):
            np.testing.assert_array_equal(a, b_)
            np.testing.assert_array_equal(b_, c_)
        for (shape, b), im, out in zip(SHAPES, imgs, one):
            ref = O.global_map(0, im, b, b, 0.5, True, fp64=True)
            assert np.abs(out - ref).max() <= MAP_TOL, (shape, b)


def test_batch_errors():
    with sred_amd.Context(0) as ctx:
        with pytest.raises(sred_amd.SredError):
            global_batch_device(ctx, 3, [(0, 10, 10, 3, 3, 0)])  # AbsDiff is not batched
        with pytest.raises(sred_amd.SredError):
            ctx.set_option(L.OPT_BATCH_PANEL, 300)
        global_batch_device(ctx, 0, [])


BIG = [((640, 1000), 17), ((1024, 520), 13), ((700, 512), 5), ((512, 900), 9)]


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_batch_large_nonsquare_images_with_b17(metric):
    """Images of 512-1024 px per side (non-square) at block sizes 5..17 in
    one batch: each map equals the single-image path's and matches FP64 on
    seeded rows plus its extremes."""
    from tests.helpers import fp64_map_at, sample_pixels
    torch = pytest.importorskip("torch")
    imgs = [O.normalize_input(synthetic.spot_ring_mixture(s[1], s[0], seed=60 + i)) for i, (s, _) in enumerate(BIG)]
    with sred_amd.Context(0) as ctx:
        ctx.set_option(L.OPT_AUTOTUNE, 0)  # largest panel that fits (the autotuner is covered above)
        d_imgs = [torch.from_numpy(im).cuda() for im in imgs]
        d_outs = [torch.full_like(d, 7.0) for d in d_imgs]
        probs = [(d.data_ptr(), im.shape[1], im.shape[0], b, b, o.data_ptr())
                 for d, im, o, (_, b) in zip(d_imgs, imgs, d_outs, BIG)]
        global_batch_device(ctx, metric, probs, 0.0, True)
        torch.cuda.synchronize()
        outs = [o.cpu().numpy() for o in d_outs]
        for i, ((shape, b), im, out) in enumerate(zip(BIG, imgs, outs)):
            single = run_single(ctx, metric, im, b)
            assert np.abs(out - single).max() <= 2e-6, (shape, b)
            m32, s32 = O.local_stats_f32(im, b, b)
            mask, _, _ = O.relevance_mask(s32, b // 2, b // 2, O.noise_var(im), 0.0)
            px = sample_pixels(out, mask, 40, seed=70 + i)
            ref, _ = fp64_map_at(O, metric, im, b, b, 0.0, px)
            assert np.abs(out.ravel()[px] - ref).max() <= MAP_TOL, (shape, b)


LARGE_BLOCKS = [((84, 90), 21, 21), ((120, 80), 11, 51), ((96, 100), 23, 23), ((60, 70), 5, 5)]


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_batch_blocks_beyond_256_pixels(metric):
    """Blocks of 317 (21x21), 385 (11x51, reflected into 384 dimensions) and
    377 (23x23) pixels in one batch beside a 5x5 class: the 10- and 12-k-step
    classes of the symmetric kernel; each map equals the single-image path's
    and matches FP64."""
    torch = pytest.importorskip("torch")
    imgs = [O.normalize_input(synthetic.spots(s[1], s[0], 14, seed=80 + i, sigma=4.0))
            for i, (s, _, _) in enumerate(LARGE_BLOCKS)]
    with sred_amd.Context(0) as ctx:
        d_imgs = [torch.from_numpy(im).cuda() for im in imgs]
        d_outs = [torch.full_like(d, 7.0) for d in d_imgs]
        probs = [(d.data_ptr(), im.shape[1], im.shape[0], bw, bh, o.data_ptr())
                 for d, im, o, (_, bw, bh) in zip(d_imgs, imgs, d_outs, LARGE_BLOCKS)]
        global_batch_device(ctx, metric, probs, 0.0, True)
        torch.cuda.synchronize()
        for (shape, bw, bh), d, im, o in zip(LARGE_BLOCKS, d_imgs, imgs, d_outs):
            single = torch.zeros_like(d)
            global_map_device(ctx, metric, L.BF16X3, d.data_ptr(), im.shape[1], im.shape[0], bw, bh, 0.0, True,
                              single.data_ptr())
            torch.cuda.synchronize()
            out = o.cpu().numpy()
            assert np.abs(out - single.cpu().numpy()).max() <= 2e-6, (shape, bw, bh)
            ref = O.global_map(metric, im, bw, bh, 0.0, True, fp64=True)
            assert np.abs(out - ref).max() <= MAP_TOL, (shape, bw, bh)
