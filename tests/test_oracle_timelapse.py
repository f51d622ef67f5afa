"""Pins the oracle's time-lapse composition (BlockRepetition2D_.java:95-137,
no stabilisation); no GPU needed.  The per-frame steps are the oracle's
block maps and input normalisation, pinned elsewhere; here the loop itself:

* identical frames give identical maps (the stack normalisation is shared);
* each frame is normalised on its own before mapping: an affine change of
  one frame leaves its map unchanged up to float rounding, while without
  the per-frame step a brighter frame would map differently;
* the stacked maps are normalised as a whole: min 0 and max 1 over the
  stack when normalize is on; otherwise the inverted NRMSE maps keep their
  range;
* fewer than 3 frames is an error (getInputImage3D returns null).
"""
import numpy as np
import pytest

import oracle as O
from collections import namedtuple

RB = namedtuple("RB", "pixels mean std")
from sred_amd import synthetic


def _stack(d, s, seed):
    return np.stack([synthetic.rings_and_blobs(s, 40, 40, seed=seed + t) for t in range(d)]).astype(np.float32)


def _template():
    rb = RB(*O.reference_block(synthetic.ring_template(9)))
    return rb


def test_identical_frames_and_affine_invariance():
    rb = _template()
    f = synthetic.rings_and_blobs(64, 30, 30, seed=3)
    st = np.stack([f, f, f, 3.0 * f + 7.0]).astype(np.float32)
    out = O.block_timelapse(O.PEARSON, st, rb.pixels, 9, 9, rb.mean, rb.std, 0.0, True)
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[1], out[2])
    np.testing.assert_allclose(out[3], out[0], atol=2e-5)
    # without the per-frame normalisation the affine frame maps differently under SSIM (it is not scale-free)
    flat = O.normalize_input(st.reshape(-1)).reshape(st.shape)
    raw3 = O.block_map(O.SSIM, flat[3], rb.pixels, 9, 9, rb.mean, rb.std, 0.0, True)
    raw0 = O.block_map(O.SSIM, flat[0], rb.pixels, 9, 9, rb.mean, rb.std, 0.0, True)
    out = O.block_timelapse(O.SSIM, st, rb.pixels, 9, 9, rb.mean, rb.std, 0.0, True)
    assert np.abs(raw3 - raw0).max() > 1e-3
    np.testing.assert_allclose(out[3], out[0], atol=2e-5)


@pytest.mark.parametrize("normalize", [True, False])
def test_stack_normalisation(normalize):
    rb = _template()
    st = _stack(4, 48, seed=10)
    out = O.block_timelapse(O.NRMSE, st, rb.pixels, 9, 9, rb.mean, rb.std, 0.0, normalize)
    if normalize:
        assert out.min() == 0.0 and out.max() == pytest.approx(1.0, abs=1e-6)
    else:  # the inverted NRMSE maps stay unnormalised (CLUtils.java:834-841)
        assert out.min() == 0.0 and out.max() > 1.0


def test_too_few_frames():
    rb = _template()
    with pytest.raises(ValueError):
        O.block_timelapse(O.PEARSON, _stack(2, 32, 1), rb.pixels, 9, 9, rb.mean, rb.std, 0.0, True)
