"""sred_block_map_2d_dev_noalloc (the block pipeline on the library's static
scratch array) against sred_block_map_2d_dev through the C ABI: bit-identical
maps on two calls in a row, the second with a smaller image; the context's
own buffers and template cache unaffected; oversized calls refused before
any device work."""
import numpy as np
import pytest

import sred_amd
from sred_amd import _lib as L
from sred_amd import synthetic, utils

pytestmark = pytest.mark.gpu

CASES = [(0, 0.0), (1, 0.0), (2, 0.0), (3, 0.0), (0, 0.7), (1, 0.7), (2, 0.7), (3, 0.7)]


def _call(fn, ctx, metric, d_img, w, h, rb, c, d_out):
    torch = pytest.importorskip("torch")
    px = L.f32(rb.pixels)
    L.check(fn(ctx.handle, metric, d_img.data_ptr(), w, h, L.fptr(px), rb.width, rb.height, rb.mean, rb.std, c, 1,
               d_out.data_ptr()))
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


def _maps(ctx, metric, img, rb, c):
    torch = pytest.importorskip("torch")
    lib = L.load()
    h, w = img.shape
    d_img = torch.from_numpy(img).cuda()
    ref = _call(lib.sred_block_map_2d_dev, ctx, metric, d_img, w, h, rb, c, torch.full_like(d_img, 3.0))
    got = _call(lib.sred_block_map_2d_dev_noalloc, ctx, metric, d_img, w, h, rb, c, torch.full_like(d_img, 5.0))
    return ref, got


@pytest.mark.parametrize("metric,c", CASES)
def test_noalloc_matches_block_map_two_calls(metric, c):
    tmpl = synthetic.ring_template(15)
    rb = utils.get_reference_block_2d(tmpl)
    big = utils.get_input_image_2d(synthetic.rings_and_blobs(512, 120, 120, seed=metric)).image_array
    small = utils.get_input_image_2d(synthetic.rings_and_blobs(256, 30, 30, seed=metric + 9)).image_array[:150, :200]
    small = np.ascontiguousarray(small)
    with sred_amd.Context(0) as ctx:
        for img in (big, small):
            ref, got = _maps(ctx, metric, img, rb, c)
            np.testing.assert_array_equal(got, ref)
    # a fresh context's first call is the no-allocation one
    with sred_amd.Context(0) as ctx:
        torch = pytest.importorskip("torch")
        d_img = torch.from_numpy(small).cuda()
        got = _call(L.load().sred_block_map_2d_dev_noalloc, ctx, metric, d_img, 200, 150, rb, c,
                    torch.zeros_like(d_img))
        np.testing.assert_array_equal(got, ref)


def test_noalloc_leaves_the_context_template_cache_valid():
    """A regular call with template A, a no-allocation call with template B,
    then the regular call with A again (its cached upload is reused) equals a
    fresh context's map with A."""
    torch = pytest.importorskip("torch")
    lib = L.load()
    ra = utils.get_reference_block_2d(synthetic.ring_template(15))
    rb = utils.get_reference_block_2d(synthetic.ring_template(9, radius=3.0))
    img = utils.get_input_image_2d(synthetic.rings_and_blobs(256, 40, 40, seed=3)).image_array
    d_img = torch.from_numpy(img).cuda()
    with sred_amd.Context(0) as ctx:
        first = _call(lib.sred_block_map_2d_dev, ctx, 0, d_img, 256, 256, ra, 0.0, torch.zeros_like(d_img))
        _call(lib.sred_block_map_2d_dev_noalloc, ctx, 0, d_img, 256, 256, rb, 0.0, torch.zeros_like(d_img))
        again = _call(lib.sred_block_map_2d_dev, ctx, 0, d_img, 256, 256, ra, 0.0, torch.zeros_like(d_img))
    with sred_amd.Context(0) as fresh:
        want = _call(lib.sred_block_map_2d_dev, fresh, 0, d_img, 256, 256, ra, 0.0, torch.zeros_like(d_img))
    np.testing.assert_array_equal(first, want)
    np.testing.assert_array_equal(again, want)


def test_noalloc_refuses_calls_beyond_its_scratch():
    torch = pytest.importorskip("torch")
    rb = utils.get_reference_block_2d(synthetic.ring_template(15))
    d_img = torch.zeros((4096, 4096), dtype=torch.float32, device="cuda")
    d_out = torch.zeros_like(d_img)
    with sred_amd.Context(0) as ctx:
        with pytest.raises(sred_amd.SredError) as e:
            _call(L.load().sred_block_map_2d_dev_noalloc, ctx, 0, d_img, 4096, 4096, rb, 1.0, d_out)
        assert e.value.code == -1 and "scratch" in str(e.value)  # SRED_ERR_ARG
        # relevance 0 Pearson needs only the template and the partials: accepted at this size
        out = _call(L.load().sred_block_map_2d_dev_noalloc, ctx, 0, d_img, 4096, 4096, rb, 0.0, d_out)
        assert np.isfinite(out).all()
