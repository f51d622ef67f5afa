"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

* known-answer vectors of the reference's own JUnit tests (tests/golden/kats.json);
* closed forms: lattice-point counts, constant images, a checkerboard whose
  block pairs have exact correlations for every metric;
* an independent brute force (numpy slicing, FP64) on tiny images;
* invariances (transpose, 180 degree rotation, affine intensity change).
"""
import numpy as np
import pytest

import oracle as O
from tests.helpers import JavaRandom, kat_reference_block, kats, random42_inputs

K = kats()
METRIC_IDS = {"Pearson's R": O.PEARSON, "SSIM": O.SSIM, "NRMSE (inverted)": O.NRMSE, "Abs. Diff. of StdDevs.": O.ABSDIFF}


def test_footprint_sizes_are_gauss_circle_numbers():
    for b, k in K["block_size_table"]["sizes"].items():
        assert O.footprint(int(b), int(b))[0] == k
    # brute-force lattice count for circles and ellipses
    for bw in (3, 5, 7, 9, 11, 13, 15, 17, 21, 31):
        for bh in (3, 5, 9, 15):
            rw, rh = bw // 2, bh // 2
            n = sum(1 for y in range(-rh, rh + 1) for x in range(-rw, rw + 1)
                    if np.float32(x * x) / np.float32(rw * rw) + np.float32(y * y) / np.float32(rh * rh) <= 1)
            assert O.footprint(bw, bh)[0] == n


def test_footprint_order_is_row_major():
    k, dx, dy = O.footprint(5, 5)
    keys = list(zip(dy, dx))
    assert keys == sorted(keys)
    assert (dx[0], dy[0]) == (0, -2) and (dx[-1], dy[-1]) == (0, 2)


def test_java_float_literal_round_trips():
    rng = np.random.default_rng(7)
    for f in list(rng.uniform(1e-6, 10, 200).astype(np.float32)) + [np.float32(1e-7), np.float32(0.1)]:
        d = O.java_float_literal(f)
        assert np.float32(d) == f
        assert len(repr(d).replace("e-0", "e-").rstrip("0")) <= 16
    assert O.java_float_literal(np.float32(1e-7)) == 1e-7
    assert O.java_float_literal(np.float32(0.1)) == 0.1
    assert O.java_float_literal(0.0) == 0.0


def test_local_statistics_kat():
    k = K["local_statistics_2d"]
    m, s = O.local_stats_f32(np.array(k["image"], np.float32), *k["block"])
    np.testing.assert_allclose(m, np.array(k["expected_means"], np.float32), atol=k["tol"])
    np.testing.assert_allclose(s, np.array(k["expected_stds"], np.float32), atol=k["tol"])


@pytest.mark.parametrize("name", ["block_pearson_2d", "block_ssim_2d", "block_nrmse_2d", "block_absdiff_2d"])
def test_block_map_kats(name):
    k = K[name]
    px, m, s = kat_reference_block(k["block"], O.footprint)
    out = O.block_map(METRIC_IDS[k["metric"]], np.array(k["image"], np.float32), px, 3, 3, m, s, 0.0, True)
    exp = np.array(k["expected"], np.float32)
    assert ((out >= 0) & (out <= 1)).all()
    np.testing.assert_allclose(out, exp, atol=k["tol"], rtol=0)


def test_reference_block_kat():
    k = K["reference_block_2d"]
    px, m, s = O.reference_block(np.array(k["block"], np.float32))
    assert abs(m - k["expected_mean"]) <= k["tol"]
    assert abs(s - k["expected_std"]) <= k["tol"]
    assert abs(s - np.sqrt(5.0 / 12.0)) < 1e-6  # closed form of the double mean subtraction
    assert O.reference_block(np.array(k["even_block"], np.float32)) is None


def test_input_normalisation_kat():
    k = K["input_image_2d_normalize"]
    out = O.normalize_input(np.array(k["image"], np.float32))
    np.testing.assert_allclose(out, np.array(k["expected"]), atol=k["tol"])


def test_mean_var_block_kat():
    k = K["mean_var_block_2d"]
    img = np.add.outer(np.arange(100), np.arange(100)).astype(np.float32)
    m, v = O.mean_var_block(img, *k["block"])
    assert abs(m - k["expected_mean"]) <= k["tol"] and abs(v - k["expected_var"]) <= k["tol"]
    with pytest.raises(ValueError):
        O.mean_var_block(img, 95, 95, 105, 105)
    bad = img.copy()
    bad[12, 12] = np.nan
    with pytest.raises(ValueError):
        O.mean_var_block(bad, *k["block"])


def test_apply_mask_kat():
    k = K["apply_mask_2d"]
    out = O.apply_mask(np.array(k["image"], np.float32), np.array(k["mask"], np.float32))
    np.testing.assert_array_equal(out, np.array(k["expected"], np.float32))


def test_relevance_mask_random42_kat():
    k = K["relevance_mask_random42"]
    img, stds = random42_inputs(k)
    nv = O.noise_var(img)
    mask, thr, npx = O.relevance_mask(stds, *k["block_radius"], nv, k["relevance_constant"])
    assert np.float64(np.float32(nv)) == k["expected_noise_var"]
    assert np.float32(thr) == np.float32(k["expected_threshold_f32"])
    assert set(np.unique(mask)) <= {0.0, 1.0}
    assert mask[:2].sum() == 0 and mask[:, :2].sum() == 0  # border
    assert npx == mask.sum()


def test_java_random_matches_documented_lcg():
    # first nextInt() of new Random(42) is -1170105035 (documented LCG)
    assert JavaRandom(42).next(32) == -1170105035


def test_normalise_with_mask_kat():
    k = K["normalize_image_2d_mask"]
    out = O.normalize_map(np.array(k["image"], np.float32), *k["border"], np.array(k["mask"], np.float32))
    np.testing.assert_allclose(out, np.array(k["expected"], np.float32), atol=k["tol"])
    k2 = K["normalize_image_2d_nonbinary"]
    with pytest.raises(ValueError):
        O.normalize_map(np.array(k2["image"], np.float32), *k2["border"], np.array(k2["mask"], np.float32))


# ---------------------------------------------------------------------------
# global maps
# ---------------------------------------------------------------------------
def checkerboard(w, h):
    return ((np.add.outer(np.arange(h), np.arange(w)) % 2) == 1).astype(np.float32)


def checkerboard_closed_form(metric, w, h, b):
    """Exact S, Wsum of the checkerboard (values 0 and a = 1/(1+1e-7) after
    input normalisation): every block has the same std, same-parity blocks
    are identical, opposite-parity blocks are complements (r = -1)."""
    a = 1.0 / (1.0 + 1e-7)
    r = b // 2
    k, dx, dy = O.footprint(b, b)
    par = (dx + dy) % 2
    n_same_phase_ones = np.count_nonzero(par == 1)  # block centred on a 0 pixel: ones where dx+dy odd
    ones0 = n_same_phase_ones
    mu0 = a * ones0 / k
    mu1 = a * (k - ones0) / k
    var = (ones0 * (a - mu0) ** 2 + (k - ones0) * mu0 ** 2) / (k - 1)
    sd = np.sqrt(var)
    ys, xs = np.mgrid[r:h - r, r:w - r]
    phase = (xs + ys) % 2
    n_even, n_odd = np.count_nonzero(phase == 0), np.count_nonzero(phase == 1)
    N = n_even + n_odd
    S = np.zeros((h, w))
    W = np.zeros((h, w))
    for ph, mu_s, mu_o, n_s, n_o in ((0, mu0, mu1, n_even, n_odd), (1, mu1, mu0, n_odd, n_even)):
        sel = np.zeros((h, w), bool)
        sel[r:h - r, r:w - r] = phase == ph
        W[sel] = N
        if metric == O.PEARSON:
            S[sel] = n_s * var / (var + 1e-7)
        elif metric == O.SSIM:
            c1, c2 = np.float32(1e-4), np.float32(9e-4)
            same = (2 * mu_s * mu_s + c1) * (2 * var + c2) / ((mu_s ** 4 + c1) * (2 * var + c2))
            opp = (2 * mu_s * mu_o + c1) * (-2 * var + c2) / ((mu_s ** 2 * mu_o ** 2 + c1) * (2 * var + c2))
            S[sel] = n_s * same + n_o * max(opp, 0.0)
        elif metric == O.NRMSE:
            S[sel] = n_o * 2.0 * sd * np.sqrt((k - 1) / k)
        else:
            S[sel] = 0.0
    return S, W


@pytest.mark.parametrize("metric", [O.PEARSON, O.SSIM, O.NRMSE, O.ABSDIFF])
@pytest.mark.parametrize("b", [3, 5])
def test_global_checkerboard_closed_form(metric, b):
    w, h = 15, 12
    img = O.normalize_input(checkerboard(w, h))
    m, s = O.local_stats_f32(img, b, b)
    nv = O.noise_var(img)
    S64, W64 = O.global_accum_f64(metric, img, s, b, b, nv, 0.0)
    S32, W32 = O.global_accum_f32(metric, img, m, s, b, b, nv, 0.0)
    Sx, Wx = checkerboard_closed_form(metric, w, h, b)
    np.testing.assert_allclose(S64, Sx, rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(W64, Wx, rtol=1e-12)
    np.testing.assert_allclose(S32, Sx, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(W32, Wx, rtol=1e-6)


@pytest.mark.parametrize("metric", [O.PEARSON, O.SSIM, O.NRMSE, O.ABSDIFF])
def test_global_constant_image_closed_form(metric):
    w, h, b = 20, 17, 5
    img = O.normalize_input(np.full((h, w), 3.0, np.float32))
    out = O.global_map(metric, img, b, b, 0.0, True)
    n = (w - 4) * (h - 4)
    inner = out[2:h - 2, 2:w - 2]
    if metric in (O.PEARSON, O.SSIM):
        # every pair is both-flat (weight 1), nPixels = 0, normalisation selects nothing
        assert (inner == np.float32(n) / np.float32(1e-7)).all()
    else:
        assert (inner == 0).all()
    assert (out[:2] == 0).all() and (out[:, :2] == 0).all()


def brute_force_f64(metric, img, b, thr=0.0):
    """Independent FP64 restatement on tiny images (numpy slicing of square
    windows and an explicit ellipse mask, no offset tables)."""
    h, w = img.shape
    r = b // 2
    yy, xx = np.mgrid[-r:r + 1, -r:r + 1]
    ell = (np.float32(1) * xx * xx / np.float32(r * r) + np.float32(1) * yy * yy / np.float32(r * r)) <= 1
    k = ell.sum()
    img64 = img.astype(np.float64)
    blocks, sd32, pos = [], [], []
    m32, s32 = O.local_stats_f32(img, b, b)
    for y in range(r, h - r):
        for x in range(r, w - r):
            v = img64[y - r:y + r + 1, x - r:x + r + 1][ell]
            blocks.append(v - v.mean())
            sd32.append(s32[y, x])
            pos.append((y, x))
    B = np.array(blocks)
    sd = np.sqrt((B * B).sum(1) / (k - 1))
    mu = np.array([img64[y - r:y + r + 1, x - r:x + r + 1][ell].mean() for y, x in pos])
    sd32 = np.array(sd32, np.float32)
    valid = ~((sd32 * sd32).astype(np.float64) < thr)
    nv = O.noise_var(img)
    h_ = O.java_float_literal(nv) + 1e-7
    S = np.zeros((h, w))
    W = np.zeros((h, w))
    cov = B @ B.T / (k - 1)
    wt = np.exp(-np.subtract.outer(sd, sd) ** 2 / h_)
    for i, (y, x) in enumerate(pos):
        if not valid[i]:
            continue
        j = valid
        W[y, x] = wt[i, j].sum()
        if metric == O.PEARSON:
            t = np.where((sd32[i] == 0) & (sd32 == 0), 1.0,
                         np.where((sd32[i] == 0) | (sd32 == 0), 0.0, np.maximum(cov[i] / (sd[i] * sd + 1e-7), 0)))
            S[y, x] = (t * wt[i])[j].sum()
        elif metric == O.SSIM:
            c1, c2 = float(np.float32(1e-4)), float(np.float32(9e-4))
            ss = (2 * mu[i] * mu + c1) * (2 * cov[i] + c2) / ((mu[i] ** 2 * mu ** 2 + c1) * (sd[i] ** 2 + sd ** 2 + c2))
            t = np.where((sd32[i] == 0) & (sd32 == 0), 1.0,
                         np.where((sd32[i] == 0) | (sd32 == 0), 0.0, np.maximum(ss, 0)))
            S[y, x] = (t * wt[i])[j].sum()
        elif metric == O.NRMSE:
            S[y, x] = np.sqrt(((B[i] - B) ** 2).sum(1) / k)[j].sum()
        else:
            S[y, x] = (np.abs(sd[i] - sd) * wt[i])[j].sum()
    return S, W


@pytest.mark.parametrize("metric", [O.PEARSON, O.SSIM, O.NRMSE, O.ABSDIFF])
@pytest.mark.parametrize("c", [0.0, 1.5])
def test_global_oracle64_matches_brute_force(metric, c):
    rng = np.random.default_rng(11)
    img = rng.random((13, 16)).astype(np.float32)
    img[3:8, 4:10] = 0.5  # a flat patch: exercises both sigma == 0 special cases
    img = O.normalize_input(img)
    b = 3
    m, s = O.local_stats_f32(img, b, b)
    nv = O.noise_var(img)
    thr = np.float32(np.float32(nv) * np.float32(c))
    S64, W64 = O.global_accum_f64(metric, img, s, b, b, nv, thr)
    Sb, Wb = brute_force_f64(metric, img, b, O.java_float_literal(thr))
    np.testing.assert_allclose(S64, Sb, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(W64, Wb, rtol=1e-9, atol=1e-12)
    # refsim (float32, sequential) stays close to FP64
    S32, W32 = O.global_accum_f32(metric, img, m, s, b, b, nv, thr)
    np.testing.assert_allclose(S32, S64, rtol=2e-4, atol=2e-5)
    np.testing.assert_allclose(W32, W64, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("metric", [O.PEARSON, O.SSIM, O.NRMSE, O.ABSDIFF])
def test_global_invariances(metric):
    rng = np.random.default_rng(5)
    img = O.normalize_input(rng.random((24, 24)).astype(np.float32))
    base = O.global_map(metric, img, 5, 5, 0.0, True, fp64=True)
    tr = O.global_map(metric, np.ascontiguousarray(img.T), 5, 5, 0.0, True, fp64=True)
    rot = O.global_map(metric, np.ascontiguousarray(img[::-1, ::-1]), 5, 5, 0.0, True, fp64=True)
    aff = O.global_map(metric, O.normalize_input(3.0 * img + 7.0), 5, 5, 0.0, True, fp64=True)
    np.testing.assert_allclose(tr, base.T, atol=1e-6)
    np.testing.assert_allclose(rot, base[::-1, ::-1], atol=1e-6)
    np.testing.assert_allclose(aff, base, atol=1e-4)


def test_refsim_vs_oracle64_normalised_maps_are_close():
    rng = np.random.default_rng(3)
    img = O.normalize_input(rng.random((40, 36)).astype(np.float32))
    for metric in (O.PEARSON, O.SSIM, O.NRMSE, O.ABSDIFF):
        a = O.global_map(metric, img, 5, 5, 0.0, True)
        b = O.global_map(metric, img, 5, 5, 0.0, True, fp64=True)
        assert np.abs(a - b).max() < 1e-4


@pytest.mark.parametrize("metric", [O.PEARSON, O.SSIM, O.NRMSE, O.ABSDIFF])
def test_global_oracle_large_blocks_match_brute_force(metric):
    """Blocks beyond 1024 pixels (41x41: 1257 pixels) against the numpy brute
    force, through the same per-thread buffers the parity tests use."""
    b = 41
    assert O.footprint(b, b)[0] > 1024
    rng = np.random.default_rng(41)
    img = O.normalize_input(rng.random((b + 7, b + 5)).astype(np.float32))
    m, s = O.local_stats_f32(img, b, b)
    nv = O.noise_var(img)
    S64, W64 = O.global_accum_f64(metric, img, s, b, b, nv, 0.0)
    Sb, Wb = brute_force_f64(metric, img, b)
    np.testing.assert_allclose(S64, Sb, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(W64, Wb, rtol=1e-9, atol=1e-12)
    S32, W32 = O.global_accum_f32(metric, img, m, s, b, b, nv, 0.0)
    np.testing.assert_allclose(S32, S64, rtol=2e-4, atol=2e-5)


def test_oracle_refuses_footprints_beyond_its_buffers():
    img = np.zeros((70, 70), np.float32)
    assert O.footprint(65, 65)[0] > O.OR_MAX_K
    with pytest.raises(ValueError):
        O.local_stats_f32(img, 65, 65)
    with pytest.raises(ValueError):
        O.global_map(O.PEARSON, img, 65, 65)
    with pytest.raises(ValueError):
        O.global_accum_f64(O.PEARSON, img, img, 65, 65, 0.01, 0.0)
