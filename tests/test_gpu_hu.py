"""GPU Hu template map (kernelGetPatchHu2D.cl) against the oracle, bit for bit."""
import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from sred_amd import synthetic, utils

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,b,seed", [((61, 72), 5, 1), ((97, 150), 9, 2), ((64, 64), 15, 3), ((33, 41), 3, 4),
                                          ((80, 70), (7, 11), 5), ((130, 140), 27, 6)])
def test_block_hu_bitexact(ctx, shape, b, seed):
    bw, bh = (b, b) if isinstance(b, int) else b
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 25, seed=seed))
    rng = np.random.default_rng(seed)
    tmpl = O.hu_prepare_template(rng.random((bh, bw)).astype(np.float32))
    inp = utils.InputImage2D(img, shape[1], shape[0], img.size)
    out = sred_amd.calculate_block_hu_map_2d(inp, tmpl, ctx).reshape(shape)
    np.testing.assert_array_equal(out, O.block_hu(img, tmpl))


def test_block_hu_ring_template_full_size(ctx):
    """2048x2048 rings-and-blobs image, 15x15 ring template (the block-repetition workload)."""
    img = O.normalize_input(synthetic.rings_and_blobs(2048, seed=1))
    tmpl = sred_amd.get_hu_template_2d(synthetic.ring_template(15))
    inp = utils.InputImage2D(img, 2048, 2048, img.size)
    out = sred_amd.calculate_block_hu_map_2d(inp, tmpl, ctx).reshape(img.shape)
    np.testing.assert_array_equal(out, O.block_hu(img, tmpl))


def test_block_hu_device_path(ctx):
    torch = pytest.importorskip("torch")
    img = O.normalize_input(synthetic.spots(50, 40, 10, seed=8))
    tmpl = sred_amd.get_hu_template_2d(synthetic.ring_template(7))
    d_img = torch.from_numpy(img).cuda()
    d_out = torch.full_like(d_img, 5.0)
    L.check(L.load().sred_block_hu_2d_dev(ctx.handle, d_img.data_ptr(), 50, 40, L.fptr(L.f32(tmpl).reshape(-1)), 7,
                                         7, d_out.data_ptr()))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_out.cpu().numpy(), O.block_hu(img, tmpl))


def test_block_hu_errors(ctx):
    img = np.zeros((20, 20), np.float32)
    inp = utils.InputImage2D(img, 20, 20, 400)
    with pytest.raises(sred_amd.SredError):
        sred_amd.calculate_block_hu_map_2d(inp, np.zeros((4, 5), np.float32), ctx)
