"""Pins the oracle's all-pairs vector maps (no GPU needed): the global Hu map
(kernelGetHuMap.cl) and the phase correlation map (textbook specification of
kernelGetPhaseCorrelationMap.cl, which does not compile in the reference).

Hu invariant (getInvariant, :146-172):
* closed form: a window holding one pixel v at offset (ox, oy) from the centre
  with mean 0 has mu00 = v, mu20 = ox^2 v, mu02 = oy^2 v;
* closed form: a constant window;
* an independent numpy restatement (array sums in FP64) and invariance under
  a 90-degree rotation of image and means (square blocks).
Hu weight and sums (:87-144):
* closed forms of the weight: equal stds -> 1, (1, 3) -> (100^0.5 - 1)/99 = 1/11;
* equal stds make the map sum_j |hu_j - hu_i|, computed by sorting;
* the reference-pixel list computes exactly the listed interior pixels.
Phase correlation:
* a circular shift s of a random block peaks at -s (all shifts, several
  block shapes); the block against itself peaks at 0; a zero block has no
  positive peak;
* r equals numpy.fft.ifft2 of the unit-spectrum cross power (a library
  routine, so the DFT conventions and the inverse are pinned);
* the weight's closed forms (kernelGetPhaseCorrelationMap.cl:167-178);
* closed form of the whole map: on an image periodic with the block size,
  every block is a circular shift of every other block, so the map is
  sum_j |wrap(q_j - q_i)| w(s_i, s_j) with q the pixel position;
* a numpy brute force (FFT per pair) on a small random image.
"""
import numpy as np
import pytest

import oracle as O

EPS = O.hu_eps()


# ---------------------------------------------------------------- Hu invariant

def test_hu_invariant_single_pixel_closed_form():
    for bw, bh, ox, oy, v in [(5, 5, 1, 2, 0.5), (7, 3, -3, 1, 0.75), (9, 9, 0, -4, 0.25), (3, 11, 1, 5, 1.0)]:
        img = np.zeros((bh + 4, bw + 6), np.float32)
        x0, y0 = 3 + bw // 2, 2 + bh // 2
        img[y0 + oy, x0 + ox] = v
        got = O.lib().or_hu_invariant(img.reshape(-1), img.shape[1], bw, bh, x0, y0, 0.0, EPS)
        m2 = np.float32(v) * np.float32(v)
        want = np.float32(np.float32(ox * ox * v) / (np.float64(m2) + EPS)) + \
            np.float32(np.float32(oy * oy * v) / (np.float64(m2) + EPS))
        assert np.float32(got) == np.float32(want), (bw, bh, ox, oy, got, want)


def test_hu_invariant_constant_window():
    bw, bh, c, m = 7, 5, 0.8, 0.3
    img = np.full((bh, bw), c, np.float32)
    got = O.lib().or_hu_invariant(img.reshape(-1), bw, bw, bh, bw // 2, bh // 2, m, EPS)
    v = np.float32(c) - np.float32(m)
    sx = 2 * sum(k * k for k in range(1, bw // 2 + 1))
    sy = 2 * sum(k * k for k in range(1, bh // 2 + 1))
    mu00 = bw * bh * float(v)
    want = (bh * sx * float(v)) / (mu00 ** 2 + EPS) + (bw * sy * float(v)) / (mu00 ** 2 + EPS)
    assert got == pytest.approx(want, rel=2e-6)


def hu_numpy(img, means, bw, bh):
    h, w = img.shape
    rw, rh = bw // 2, bh // 2
    dx2 = (np.arange(bw) - rw) ** 2.0
    dy2 = (np.arange(bh) - rh) ** 2.0
    out = np.zeros((h, w))
    for y in range(rh, h - rh):
        for x in range(rw, w - rw):
            p = img[y - rh:y + rh + 1, x - rw:x + rw + 1].astype(np.float64) - float(means[y, x])
            m00 = p.sum()
            out[y, x] = ((p * dx2[None, :]).sum() + (p * dy2[:, None]).sum()) / (m00 * m00 + EPS)
    return out


@pytest.mark.parametrize("bw,bh", [(5, 5), (3, 7), (9, 5)])
def test_hu_invariants_numpy_restatement(bw, bh):
    rng = np.random.default_rng(bw * 10 + bh)
    img = rng.random((23, 29)).astype(np.float32)
    means, _ = O.local_stats_f32(img, bw, bh)
    got = O.hu_invariants(img, means, bw, bh)
    want = hu_numpy(img, means, bw, bh)
    # hu is ill-conditioned where mu00 nearly cancels; compare relative to the window's scale
    scale = np.maximum(np.abs(want), 1.0)
    assert np.max(np.abs(got - want) / scale) < 1e-3
    assert np.median(np.abs(got - want) / scale) < 1e-6


def test_hu_invariants_rotation():
    rng = np.random.default_rng(5)
    img = rng.random((21, 21)).astype(np.float32)
    means = rng.random((21, 21)).astype(np.float32) * 0.2 + 0.3
    a = O.hu_invariants(img, means, 5, 5)
    b = O.hu_invariants(np.ascontiguousarray(np.rot90(img)), np.ascontiguousarray(np.rot90(means)), 5, 5)
    np.testing.assert_allclose(np.rot90(a), b, rtol=1e-4, atol=1e-4)


# ---------------------------------------------------------------- Hu weight/sums

def test_hu_weight_closed_forms():
    assert O.hu_weight(0.4, 0.4) == pytest.approx(1.0, abs=1e-15)
    assert O.hu_weight(0.0, 0.0) == pytest.approx(1.0, abs=1e-15)
    assert O.hu_weight(1.0, 3.0) == pytest.approx(1.0 / 11.0, rel=1e-6)   # sim = 1/2, 100^(1/2) = 10
    assert O.hu_weight(3.0, 1.0) == O.hu_weight(1.0, 3.0)
    assert O.hu_weight(0.0, 2.0) == pytest.approx(0.0, abs=1e-7)         # sim = eps/2


def test_hu_sums_equal_stds_sorted_closed_form():
    h, w, bw, bh = 13, 17, 5, 3
    rng = np.random.default_rng(1)
    hu = rng.normal(size=(h, w)).astype(np.float32)
    stds = np.full((h, w), 0.37, np.float32)
    got = O.global_hu_sums(hu, stds, bw, bh)
    inner = hu[bh // 2:h - bh // 2, bw // 2:w - bw // 2].astype(np.float64)
    v = np.sort(inner.reshape(-1))
    cs = np.concatenate([[0.0], np.cumsum(v)])
    n = v.size
    want = np.zeros((h, w))
    for y in range(bh // 2, h - bh // 2):
        for x in range(bw // 2, w - bw // 2):
            a = float(hu[y, x])
            k = np.searchsorted(v, a, side="left")
            want[y, x] = (a * k - cs[k]) + (cs[n] - cs[k] - a * (n - k))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


def test_hu_sums_rows_list():
    h, w = 15, 12
    rng = np.random.default_rng(2)
    hu = rng.normal(size=(h, w)).astype(np.float32)
    stds = rng.random((h, w)).astype(np.float32)
    full = O.global_hu_sums(hu, stds, 3, 5)
    rows = [0, 5 * w + 4, 7 * w + 1, 2 * w + 2, 10 * w + 9, h * w + 3, -1]   # border/out-of-range skipped
    sub = O.global_hu_sums(hu, stds, 3, 5, rows=rows)
    mask = np.zeros(h * w, bool)
    mask[[5 * w + 4, 7 * w + 1, 2 * w + 2, 10 * w + 9]] = True
    mask = mask.reshape(h, w)
    np.testing.assert_array_equal(sub[mask], full[mask])
    assert not sub[~mask].any()


# ---------------------------------------------------------------- phase correlation

@pytest.mark.parametrize("bw,bh", [(3, 3), (5, 5), (7, 3), (5, 9), (11, 11)])
def test_phase_shift_peaks_at_minus_shift(bw, bh):
    rng = np.random.default_rng(bw * 100 + bh)
    b = rng.normal(size=(bh, bw))
    for sx in range(-(bw // 2), bw // 2 + 1):
        for sy in range(-(bh // 2), bh // 2 + 1):
            _, d, found = O.phase_peak_blocks(b, np.roll(b, (sy, sx), axis=(0, 1)))
            assert found and d == (-sx, -sy)


@pytest.mark.parametrize("bw,bh", [(3, 5), (5, 5), (9, 7)])
def test_phase_r_matches_numpy_fft(bw, bh):
    rng = np.random.default_rng(bw + bh)
    for _ in range(5):
        bi, bj = rng.normal(size=(bh, bw)), rng.normal(size=(bh, bw))
        r, d, found = O.phase_peak_blocks(bi, bj)
        Fi, Fj = np.fft.fft2(bi), np.fft.fft2(bj)
        Ui, Uj = Fi / (np.abs(Fi) + EPS), Fj / (np.abs(Fj) + EPS)
        rn = np.fft.ifft2(Ui * np.conj(Uj)).real
        np.testing.assert_allclose(r, rn, atol=1e-12)
        k = int(np.argmax(rn))
        py, px = divmod(k, bw)
        assert found == (rn.max() > 0)
        assert d == ((px if px <= bw // 2 else px - bw), (py if py <= bh // 2 else py - bh))


def test_phase_self_and_zero_blocks():
    rng = np.random.default_rng(3)
    b = rng.normal(size=(5, 7))
    _, d, found = O.phase_peak_blocks(b, b)
    assert found and d == (0, 0)
    r, d, found = O.phase_peak_blocks(np.zeros((5, 7)), b)
    assert not found and d == (0, 0) and not r.any()


def test_phase_weight_closed_forms():
    assert O.phase_weight(0.3, 0.3) == 1.0
    assert O.phase_weight(0.0, 0.0) == 1.0
    assert O.phase_weight(1.0, 3.0) == 0.0
    assert O.phase_weight(0.0, 2.0) == -1.0
    assert O.phase_weight(1.0, 2.0) == pytest.approx(1.0 / 3.0, abs=1e-7)


def periodic_image(w, h, bw, bh, seed):
    rng = np.random.default_rng(seed)
    pat = rng.random((bh, bw)).astype(np.float32)
    return np.tile(pat, (h // bh + 1, w // bw + 1))[:h, :w].copy()


def periodic_closed_form(img, stds, bw, bh):
    h, w = img.shape
    rw, rh = bw // 2, bh // 2
    ys, xs = np.mgrid[rh:h - rh, rw:w - rw]
    ys, xs = ys.reshape(-1), xs.reshape(-1)
    sj = stds[ys, xs]
    out = np.zeros((h, w))
    for y, x in zip(ys, xs):
        dx = (xs - x) % bw
        dy = (ys - y) % bh
        dx = np.where(dx <= rw, dx, dx - bw)
        dy = np.where(dy <= rh, dy, dy - bh)
        si = np.float32(stds[y, x])
        wgt = np.where(sj == si, np.float32(1),
                       np.float32(1) - np.float32(2) * (np.abs(si - sj) / (np.abs(si) + np.abs(sj))).astype(np.float32))
        out[y, x] = np.sum(np.sqrt(dx * dx + dy * dy) * wgt.astype(np.float64))
    return out


@pytest.mark.parametrize("bw,bh", [(5, 5), (3, 7)])
def test_phase_map_periodic_closed_form(bw, bh):
    img = periodic_image(19, 17, bw, bh, seed=bw + bh)
    means, stds = O.local_stats_f32(img, bw, bh)
    got = O.phase_map(img, means, stds, bw, bh)
    np.testing.assert_allclose(got, periodic_closed_form(img, stds, bw, bh), rtol=1e-12, atol=1e-9)


def test_phase_map_numpy_bruteforce():
    bw, bh = 3, 5
    rng = np.random.default_rng(9)
    img = rng.random((12, 11)).astype(np.float32)
    means, stds = O.local_stats_f32(img, bw, bh)
    h, w = img.shape
    rw, rh = bw // 2, bh // 2
    pix = [(y, x) for y in range(rh, h - rh) for x in range(rw, w - rw)]
    blocks = np.stack([(img[y - rh:y + rh + 1, x - rw:x + rw + 1].astype(np.float64) - float(means[y, x]))
                       / (float(stds[y, x]) + EPS) for y, x in pix])
    F = np.fft.fft2(blocks)
    U = F / (np.abs(F) + EPS)
    want = np.zeros((h, w))
    idx = np.arange(len(pix))
    for i, (y, x) in enumerate(pix):
        # each pair's lower interior index leads: R = U_lo conj(U_hi)
        lo_first = np.fft.ifft2(U[i][None] * np.conj(U)).real.reshape(len(pix), -1)
        hi_first = np.fft.ifft2(U * np.conj(U[i])[None]).real.reshape(len(pix), -1)
        r = np.where((idx >= i)[:, None], lo_first, hi_first)
        k = np.argmax(r, axis=1)
        pos = r[np.arange(len(pix)), k] > 0
        py, px = np.divmod(k, bw)
        dx = np.where(px <= rw, px, px - bw) * pos
        dy = np.where(py <= rh, py, py - bh) * pos
        wgt = np.array([O.phase_weight(stds[y, x], stds[q]) for q in pix])
        want[y, x] = np.sum(np.sqrt(dx * dx + dy * dy) * wgt)
    np.testing.assert_allclose(O.phase_map(img, means, stds, bw, bh), want, rtol=1e-12, atol=1e-12)
