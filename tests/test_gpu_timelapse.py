"""GPU time-lapse block repetition (BlockRepetition2D_.java:95-137) against
the oracle's composition of the same steps, through the C ABI: every map of
the stack bit for bit, for the four metrics, with and without a relevance
constant and output normalisation; the device entry; the frame-count error."""
import ctypes

import numpy as np
import pytest

import oracle as O
from collections import namedtuple

RB = namedtuple("RB", "pixels mean std")
import sred_amd
from sred_amd import _lib as L
from sred_amd import synthetic

pytestmark = pytest.mark.gpu

METRICS = [("Pearson's R", O.PEARSON), ("SSIM", O.SSIM), ("NRMSE (inverted)", O.NRMSE),
           ("Abs. Diff. of StdDevs.", O.ABSDIFF)]


def _stack(d, h, w, seed):
    s = max(h, w)
    fr = [synthetic.rings_and_blobs(s, 60, 60, seed=seed + t)[:h, :w] * (1.0 + 0.3 * t) for t in range(d)]
    return np.stack(fr).astype(np.float32)


@pytest.mark.parametrize("name,metric", METRICS)
@pytest.mark.parametrize("c,normalize", [(0.0, True), (0.8, True), (0.0, False)])
def test_timelapse_bitexact(ctx, name, metric, c, normalize):
    st = _stack(5, 72, 90, seed=20)
    tmpl = synthetic.ring_template(9)
    rb = RB(*O.reference_block(tmpl))
    got = sred_amd.block_repetition_timelapse_2d(tmpl, st, c, name, normalize, resources=ctx)
    want = O.block_timelapse(metric, st, rb.pixels, 9, 9, rb.mean, rb.std, c, normalize)
    np.testing.assert_array_equal(got, want)


def test_timelapse_device_entry_and_errors(ctx):
    torch = pytest.importorskip("torch")
    st = _stack(6, 128, 128, seed=30)
    tmpl = synthetic.ring_template(15)
    rb = RB(*O.reference_block(tmpl))
    d = torch.from_numpy(st).cuda()
    out = torch.empty_like(d)
    L.check(L.load().sred_block_timelapse_2d_dev(ctx.handle, O.PEARSON, d.data_ptr(), 128, 128, 6,
                                                 L.fptr(L.f32(rb.pixels)), 15, 15, rb.mean, rb.std, 0.0, 1,
                                                 out.data_ptr()))
    want = O.block_timelapse(O.PEARSON, st, rb.pixels, 15, 15, rb.mean, rb.std, 0.0, True)
    np.testing.assert_array_equal(out.cpu().numpy(), want)
    np.testing.assert_array_equal(d.cpu().numpy(), st)  # the input stack is not modified
    with pytest.raises(sred_amd.SredError):
        sred_amd.block_repetition_timelapse_2d(tmpl, st[:2], 0.0, "Pearson's R", True, resources=ctx)


def test_timelapse_stabilised(ctx):
    """getInputImage3D's stabilisation of the stack (:97): "Quad/Octree" is the
    octree GAT of the whole stack before the maps, "Simplex" changes nothing."""
    rng = np.random.default_rng(21)  # gain-2 counts, constant on the octree's 28 x 24 x 4 leaves
    lam = np.kron(rng.uniform(20, 400, (4, 4, 4)), np.ones((4, 24, 28)))
    st = (2.0 * rng.poisson(lam) + rng.normal(10.0, 3.0, lam.shape)).astype(np.float32)
    tmpl = st[5, 20:29, 24:33].copy()
    rb = RB(*O.reference_block(tmpl))
    oc = O.octree_gat(st, 4, 0.01, 50, 0.75)
    stab = O.gat_tree3d(st, oc.g0, oc.edc)
    got = sred_amd.block_repetition_timelapse_2d(tmpl, st, 0.0, "Pearson's R", True, True, "Quad/Octree",
                                                 resources=ctx)
    want = O.block_timelapse(O.PEARSON, stab, rb.pixels, 9, 9, rb.mean, rb.std, 0.0, True)
    np.testing.assert_array_equal(got, want)
    plain = sred_amd.block_repetition_timelapse_2d(tmpl, st, 0.0, "Pearson's R", True, resources=ctx)
    simplex = sred_amd.block_repetition_timelapse_2d(tmpl, st, 0.0, "Pearson's R", True, True, "Simplex",
                                                     resources=ctx)
    np.testing.assert_array_equal(simplex, plain)
    assert not np.array_equal(got, plain)
