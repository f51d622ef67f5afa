"""Pins the oracle's Hu template metric (kernelGetPatchHu2D.cl) and the
library's host-side template preparation (no GPU needed).

* an independent numpy restatement (masks and array sums instead of the
  kernel's loops) on small images;
* closed forms: a window holding one bright pixel (the 1-based raw / 0-based
  central moment offset makes hu1 = 2 (1+eps)-ish instead of 0), a constant
  window (hu1 = 0);
* the template prepared from a block reproduces the window's own hu1 at that
  block's position.
"""
import numpy as np
import pytest

import oracle as O
from sred_amd import synthetic

EPS = O.hu_eps()


def masks(bw, bh):
    rw, rh = bw // 2, bh // 2
    i = np.arange(bw)
    j = np.arange(bh)
    dxf = (np.float32(i - rw) / np.float32(rw)).astype(np.float32)
    dyf = (np.float32(j - rh) / np.float32(rh)).astype(np.float32)
    ell_f = ((dxf[None, :] * dxf[None, :]).astype(np.float32) + (dyf[:, None] * dyf[:, None]).astype(np.float32)
             ) <= np.float32(1)
    ell_d = ((i[None, :] - rw) / rw) ** 2 + ((j[:, None] - rh) / rh) ** 2 <= 1.0
    corner = (np.abs(i[None, :] - rw) == rw) & (np.abs(j[:, None] - rh) == rh)
    return ell_f, ell_d, ~corner


def hu1_np(p, m):
    jj, ii = np.nonzero(m)
    v = p[m]
    M00 = v.sum()
    cx = ((ii + 1) * v).sum() / (M00 + EPS)
    cy = ((jj + 1) * v).sum() / (M00 + EPS)
    mu20 = (((ii - cx) ** 2) * v).sum()
    mu02 = (((jj - cy) ** 2) * v).sum()
    return mu20 / (M00 ** 2 + EPS) + mu02 / (M00 ** 2 + EPS)


def window_np(patch, K, ell_f, sq):
    p = patch.astype(np.float64).copy()
    mean = p[sq].sum() / K
    v = p[ell_f] - mean
    p[ell_f] = (v - v.min()) / (v.max() - v.min() + EPS)
    return p


def map_np(img, tmpl):
    bh, bw = tmpl.shape
    rw, rh = bw // 2, bh // 2
    ell_f, ell_d, sq = masks(bw, bh)
    K = int(ell_f.sum())
    t = np.where(ell_f, tmpl.astype(np.float64), 0.0)
    ref = hu1_np(t, ell_d)
    h, w = img.shape
    out = np.zeros((h, w))
    for y in range(rh, h - rh):
        for x in range(rw, w - rw):
            p = window_np(img[y - rh:y + rh + 1, x - rw:x + rw + 1], K, ell_f, sq)
            out[y, x] = np.sqrt(ref * hu1_np(p, ell_f))
    return out


@pytest.mark.parametrize("shape,b,seed", [((20, 23), 5, 1), ((22, 19), 7, 2), ((17, 25), 3, 3)])
def test_hu_oracle_matches_numpy(shape, b, seed):
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 6, seed=seed))
    rng = np.random.default_rng(seed)
    tmpl = O.hu_prepare_template(rng.random((b, b)).astype(np.float32))
    out = O.block_hu(img, tmpl)
    ref = map_np(img, tmpl)
    np.testing.assert_allclose(out, ref, rtol=2e-7, atol=1e-30)
    r = b // 2
    assert (out[:r] == 0).all() and (out[:, :r] == 0).all() and (out[-r:] == 0).all() and (out[:, -r:] == 0).all()


@pytest.mark.parametrize("b", [5, 9, 15])
def test_hu_closed_forms(b):
    r = b // 2
    K = O.footprint(b, b)[0]
    # one bright pixel in the middle of the window: after the window steps it
    # is a = 1/(1+eps), every other ellipse cell 0
    img = np.zeros((b + 4, b + 4), np.float32)
    img[2 + r, 2 + r] = 1.0
    tmpl = np.ones((b, b), np.float32)
    a = 1.0 / (1.0 + EPS)
    c = (r + 1) * a / (a + EPS)
    hu_win = 2 * ((r - c) ** 2 * a) / (a * a + EPS)
    ell_f, ell_d, _ = masks(b, b)
    hu_t = hu1_np(np.where(ell_f, 1.0, 0.0), ell_d)
    out = O.block_hu(img, tmpl)
    assert abs(out[2 + r, 2 + r] - np.float32(np.sqrt(hu_t * hu_win))) <= 1e-6 * abs(out[2 + r, 2 + r])
    assert hu_win > 1.9  # the 1-based raw moments; 0-based ones would give 0
    # constant window: all-zero normalised patch, hu1 = 0
    flat = np.full((b + 2, b + 2), 0.3, np.float32)
    assert (O.block_hu(flat, tmpl) == 0).all()
    assert K == int(ell_f.sum())


def test_hu_template_reproduces_window():
    b = 9
    img = O.normalize_input(synthetic.spots(40, 40, 12, seed=5))
    y, x = 20, 17
    blk = img[y - 4:y + 5, x - 4:x + 5]
    tmpl = O.hu_prepare_template(blk)
    ell_f, _, sq = masks(b, b)
    K = O.footprint(b, b)[0]
    hw = hu1_np(window_np(blk, K, ell_f, sq), ell_f)
    out = O.block_hu(img, tmpl)
    assert abs(out[y, x] - np.float32(hw)) <= 2e-7 * hw
    assert abs(O.hu1_template(tmpl) - hw) <= 1e-6 * hw


def test_hu_template_host_entry_matches_oracle():
    """sred_hu_prepare_template_2d (library host code) equals the oracle bit for bit."""
    import sred_amd
    rng = np.random.default_rng(7)
    for b in (3, 5, 11, 15):
        blk = rng.random((b, b)).astype(np.float32)
        np.testing.assert_array_equal(sred_amd.get_hu_template_2d(blk), O.hu_prepare_template(blk))
    with pytest.raises(sred_amd.SredError):
        sred_amd.get_hu_template_2d(np.zeros((4, 4), np.float32))
