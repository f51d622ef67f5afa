"""Pins the oracle's relevance-mask tools (RelevanceMask2D_,
OptimiseRelevanceMask_, the input step they share with CalculateLocalMeans_);
no GPU needed.

* the input step: the GAT with the fitted doubles followed by a min-max
  normalisation in double (numpy float64 arithmetic is the same IEEE double:
  a float-precision normalisation would differ), output in [0, 1] with min
  exactly 0; the fit starts at gain 0, an invalid point (NaN objective),
  and still reaches a finite objective;
* the double-sum noise variance against an FP64 numpy restatement and
  against the float-sum variant (they differ only by float rounding);
* the tool mask: a constant image has zero local deviation, so its mask is
  all zeros, and its noise variance is 0, so the threshold is the constant
  itself; with c = 0 the mask is 1 exactly where the deviation is nonzero;
* the constant scan: the literal loop (masks rebuilt for every constant)
  against a reduction formulation in numpy (min and max of the squared
  variances walked along the float sequence 0, step, 2 step, ...); the
  all-ones constant precedes the all-zeros one;
* the mask stack (RelevanceMaskStack2D_): the count (int) ceil((upper -
  lower) / step) + 1 in float, the constants as float sums, each mask
  against the numpy rule on the squared local variances, and, with lower 0,
  the last all-ones and first all-zeros masks at the constants the
  already-pinned scan reports.
"""
import numpy as np
import pytest

import oracle as O
from sred_amd import synthetic


def _image(h, w, seed):
    rng = np.random.default_rng(seed)
    base = synthetic.rings_and_blobs(max(h, w), 30, 30, seed=seed)[:h, :w]
    return (3.0 * rng.poisson(5 + 40 * base / base.max()) + rng.normal(8, 2, (h, w))).astype(np.float32)


def test_tool_prep_double_normalisation():
    img = _image(64, 80, 1)
    prep, p = O.tool_prep(img, 400)
    x, f = O.gat_minimize(img, 0.0, 100.0, 0.0, 400)
    assert tuple(p) == tuple(x) and np.isfinite(f)
    g = O.gat(img, *p).astype(np.float64)  # the fitted doubles, not rounded to float
    mn, mx = g.min(), g.max()
    want = ((g - mn) / (mx - mn + np.float64(np.float32(1e-7)))).astype(np.float32)
    np.testing.assert_array_equal(prep, want)
    assert prep.min() == 0.0 and prep.max() <= 1.0


def test_noise_var_double():
    img = _image(96, 64, 2)
    t = 8
    tiles = img[: 96 // t * t, : 64 // t * t].reshape(96 // t, t, 64 // t, t).swapaxes(1, 2).reshape(-1, t * t)
    d = tiles.astype(np.float64)
    var = (np.float32(tiles * tiles).astype(np.float64).mean(1) - d.mean(1) ** 2).astype(np.float32)
    v = np.sort(var)
    n = int(np.float32(0.03) * np.float32(len(v)) + np.float32(1.0))
    nv = np.float32(abs(np.float64(v[:n].astype(np.float64).sum()) / n))
    nv = (1.0 + 0.001 * (nv - 40.0)) * nv
    assert O.noise_var_double(img) == pytest.approx(nv, rel=1e-5)
    assert O.noise_var_double(img) == pytest.approx(O.noise_var(img), rel=1e-4)


def test_relevance_mask_tool_closed_forms():
    mask, nv, thr = O.relevance_mask_tool(np.full((40, 48), 0.25, np.float32), 3, 3, 2.5)
    assert not mask.any() and nv == 0.0 and thr == np.float32(2.5)
    img, _ = O.tool_prep(_image(48, 40, 3), 200)
    mask, nv, thr = O.relevance_mask_tool(img, 5, 5, 0.0)
    _, s = O.local_stats_f32(img, 5, 5)
    inner = np.zeros_like(mask, bool)
    inner[2:-2, 2:-2] = True
    np.testing.assert_array_equal(mask.astype(bool), inner & (s * s != 0))
    with pytest.raises(ValueError):
        O.relevance_mask_tool(img, 3, 5, 1.0)


def _scan_by_extremes(img, b, step):
    var = O.local_var_f32(img, b, b)
    r = b // 2
    v2 = (var[r:-r, r:-r] * var[r:-r, r:-r]).astype(np.float32).ravel()
    nv = np.float32(O.noise_var(img))
    vmin, nz = v2.min(), v2[v2 != 0]
    vmax, any_zero = (nz.max() if nz.size else None), bool((v2 == 0).any())
    c, ones = np.float32(0.0), np.float32(0.0)
    step = np.float32(step)
    while True:
        thr = c if nv == 0 else np.float32(nv * c)
        if not any_zero and vmin >= thr:
            ones = c
        if vmax is None or vmax < thr:
            return float(ones), float(c)
        c = np.float32(c + step)


@pytest.mark.parametrize("b,step", [(3, 0.001), (5, 0.002), (3, 0.0004)])
def test_optimise_scan_against_extremes(b, step):
    img, _ = O.tool_prep(_image(40, 56, 4 + b), 200)
    ones, zeros = O.optimise_relevance_mask(img, b, b, step)
    assert (ones, zeros) == _scan_by_extremes(img, b, step)
    assert ones <= zeros
    with pytest.raises(ValueError):
        O.optimise_relevance_mask(img, b, b, 0.0)


@pytest.mark.parametrize("b,lower,upper,step", [(3, 0.0, 1.0, 0.1), (5, 0.25, 3.0, 0.35), (3, 0.0, 40.0, 2.5)])
def test_mask_stack(b, lower, upper, step):
    img = O.tool_prep(_image(60, 72, 11), 300)[0]
    masks, consts = O.relevance_mask_stack(img, b, b, lower, upper, step)
    f = np.float32
    n = int(np.ceil(np.float64((f(upper) - f(lower)) / f(step)))) + 1
    assert len(consts) == n
    c = f(lower)
    for i in range(n):
        assert consts[i] == c
        c = f(c + f(step))
    var2 = O.local_var_f32(img, b, b) ** 2
    nv = f(O.noise_var(img))
    br = b // 2
    inner = np.zeros(img.shape, bool)
    inner[br:-br, br:-br] = True
    for m, c in zip(masks, consts):
        thr = c if nv == 0 else f(nv * c)
        want = inner & ~((var2 < thr) | (var2 == 0))
        np.testing.assert_array_equal(m, want.astype(np.float32))
    if lower == 0.0:
        ones, zeros = O.optimise_relevance_mask(img, b, b, step)
        full = [i for i, m in enumerate(masks) if m[inner].all()]
        empty = [i for i, m in enumerate(masks) if not m[inner].any()]
        if full and full[-1] + 1 < n:
            assert consts[full[-1]] == ones
        if empty:
            assert consts[empty[0]] == zeros
    with pytest.raises(ValueError):
        O.relevance_mask_stack(img, 3, 5, 0.0, 1.0, 0.1)
    with pytest.raises(ValueError):
        O.relevance_mask_stack(img, 3, 3, 1.0, 0.0, 0.1)
