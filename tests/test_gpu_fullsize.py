"""GPU parity at the benchmark workloads' full sizes, block shapes beyond the
split-BF16 range, non-square global blocks and the device noise variance.

Bars: the normalised global maps within 1e-4 absolute of the FP64 oracle
(oracle64) and, where the whole float32 restatement (refsim) finishes in
seconds, of refsim too; block maps and the noise variance bit-exact.  Maps
too large for every oracle row are checked on seeded rows plus the map's
extremes (tests.helpers.fp64_map_at), which fixes the normalisation.
"""
import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from sred_amd import synthetic, utils
from tests.helpers import fp64_map_at, sample_pixels

pytestmark = pytest.mark.gpu
GM = sred_amd.GLOBAL_METRICS
BM = sred_amd.BLOCK_METRICS
MAP_TOL = 1e-4


def noise_var_dev(ctx, img):
    torch = pytest.importorskip("torch")
    d = torch.from_numpy(np.ascontiguousarray(img, np.float32)).cuda()
    out = np.zeros(1, np.float32)
    L.check(L.load().sred_noise_var_2d_dev(ctx.handle, d.data_ptr(), img.shape[1], img.shape[0], L.fptr(out)))
    return out[0]


@pytest.mark.parametrize("shape,seed", [((256, 256), 0), ((288, 352), 1), ((289, 352), 2), ((131, 97), 3),
                                        ((1024, 1024), 4), ((8, 8), 5), ((17, 23), 6)])
def test_noise_variance_device_bitexact(ctx, shape, seed):
    """The device noise variance every _dev and batched pipeline uses
    (tile variances on the GPU, sort and lowest-3% mean on the host) against
    the oracle's Utils.getMeanNoiseVar2D, across the 8x8 / 16x16 tile switch
    at 352 x 288 pixels."""
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], max(3, shape[0] * shape[1] // 300), seed=seed))
    got = noise_var_dev(ctx, img)
    want = np.float32(O.noise_var(img))
    assert got.tobytes() == want.tobytes(), (got, want)
    assert got.tobytes() == np.float32(utils.get_mean_noise_var_2d(img, img.shape[1], img.shape[0])).tobytes()


# ---------------------------------------------------------------------------
# 256 x 256, 5 x 5 global Pearson: the reference's CPU-runnable case, whole map
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def spots256():
    img = O.normalize_input(synthetic.spots(256, 256, 200, seed=0))
    return img, O.global_map(O.PEARSON, img, 5, 5, 0.0, True), O.global_map(O.PEARSON, img, 5, 5, 0.0, True,
                                                                           fp64=True)


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
def test_spots256_b5_full_map(ctx, spots256, precision):
    img, ref32, ref64 = spots256
    out = sred_amd.global_repetition_2d(img, 5, 5, 0.0, "Pearson's R", True, precision, resources=ctx)
    assert np.abs(out - ref64).max() <= MAP_TOL
    assert np.abs(out - ref32).max() <= MAP_TOL


# ---------------------------------------------------------------------------
# 1024 x 1024, 11 x 11: Pearson, SSIM and NRMSE in both precisions
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def lattice1024():
    return O.normalize_input(synthetic.lattice_and_blobs(1024, seed=2))


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
@pytest.mark.parametrize("metric", [0, 1, 2])
def test_lattice1024_b11_sampled_rows(ctx, lattice1024, metric, precision):
    img = lattice1024
    out = sred_amd.global_repetition_2d(img, 11, 11, 0.0, GM[metric], True, precision, resources=ctx)
    m32, s32 = O.local_stats_f32(img, 11, 11)
    mask, _, _ = O.relevance_mask(s32, 5, 5, O.noise_var(img), 0.0)
    px = sample_pixels(out, mask, 96, seed=100 + metric)
    ref, _ = fp64_map_at(O, metric, img, 11, 11, 0.0, px)
    err = np.abs(out.ravel()[px] - ref)
    assert err.max() <= MAP_TOL, (err.max(), px[np.argmax(err)])
    # the minimum maps to 0; the maximum stays below 1 because the reference
    # divides by (max - min + EPSILON) with a range of ~1e-6 here (maps scale
    # as 1/nPixels)
    assert out[mask > 0].min() == 0.0


# ---------------------------------------------------------------------------
# block shapes: beyond K = 256 and non-square
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("metric", [0, 1, 2, 3])
@pytest.mark.parametrize("b,shape", [(21, (90, 80)), (23, (84, 92)), (31, (72, 66))])
def test_global_large_blocks_default_precision(ctx, metric, b, shape):
    """Blocks the reference dialog accepts beyond K = 256 through the default
    (split-BF16) precision: 21x21 (317 px) and 23x23 (377 px) take the
    symmetric split-BF16 contraction with one row group (Pearson, SSIM,
    NRMSE), 31x31 (709 px) and AbsDiffStds the FP32 contraction; every map
    matches the oracle."""
    K = sred_amd.block_size_2d(b, b)
    assert K > 256
    assert bool(L.global_plan(10000, K, metric, L.BF16X3)["flags"] & 1) == (K <= 384 and metric != L.ABSDIFF)
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 12, seed=b + metric, sigma=4.0))
    out = sred_amd.global_repetition_2d(img, b, b, 0.0, GM[metric], resources=ctx)
    ref64 = O.global_map(metric, img, b, b, 0.0, True, fp64=True)
    ref32 = O.global_map(metric, img, b, b, 0.0, True)
    assert np.abs(out - ref64).max() <= MAP_TOL
    assert np.abs(out - ref32).max() <= MAP_TOL


def test_global_large_block_device_entry(ctx):
    """A 21x21 map through sred_global_map_2d_dev with BF16X3 requested."""
    torch = pytest.importorskip("torch")
    img = O.normalize_input(synthetic.spots(96, 84, 15, seed=7, sigma=4.0))
    d_img = torch.from_numpy(img).cuda()
    d_out = torch.zeros_like(d_img)
    sred_amd.commands.global_map_device(ctx, L.PEARSON, L.BF16X3, d_img.data_ptr(), 96, 84, 21, 21, 0.0, True,
                                        d_out.data_ptr())
    torch.cuda.synchronize()
    ref = O.global_map(O.PEARSON, img, 21, 21, 0.0, True, fp64=True)
    assert np.abs(d_out.cpu().numpy() - ref).max() <= MAP_TOL


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
@pytest.mark.parametrize("metric", [0, 1, 2, 3])
@pytest.mark.parametrize("bw,bh", [(9, 5), (5, 11), (3, 7)])
def test_global_non_square_blocks(ctx, metric, precision, bw, bh):
    img = O.normalize_input(synthetic.spots(83, 71, 30, seed=bw * 10 + bh))
    out = sred_amd.global_repetition_2d(img, bw, bh, 0.0, GM[metric], True, precision, resources=ctx)
    ref64 = O.global_map(metric, img, bw, bh, 0.0, True, fp64=True)
    ref32 = O.global_map(metric, img, bw, bh, 0.0, True)
    assert np.abs(out - ref64).max() <= MAP_TOL
    assert np.abs(out - ref32).max() <= MAP_TOL


@pytest.mark.parametrize("metric", [0, 1, 2])
@pytest.mark.parametrize("bw,bh", [(3, 31), (15, 7), (9, 17)])
def test_global_reflected_blocks(ctx, metric, bw, bh):
    """Blocks of K = 32m + 1 pixels (3x31: 33, 15x7: 65, 9x17: 97), which
    the split-BF16 gather reflects into K - 1 dimensions (one k-step fewer):
    the maps match FP64 and the FP32 contraction, which keeps all K."""
    K = sred_amd.block_size_2d(bw, bh)
    assert (K - 1) % 32 == 0 and L.global_plan(10000, K, metric, L.BF16X3)["flags"] & 4
    img = O.normalize_input(synthetic.spots(90, 84, 30, seed=bw * 100 + bh, sigma=3.0))
    out = sred_amd.global_repetition_2d(img, bw, bh, 0.0, GM[metric], True, "bf16x3", resources=ctx)
    f32 = sred_amd.global_repetition_2d(img, bw, bh, 0.0, GM[metric], True, "fp32", resources=ctx)
    ref64 = O.global_map(metric, img, bw, bh, 0.0, True, fp64=True)
    assert np.abs(out - ref64).max() <= MAP_TOL
    assert np.abs(out - f32).max() <= MAP_TOL


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_global_385_pixel_blocks(ctx, metric):
    """11x51 blocks hold 385 pixels: reflected into 384 dimensions they take
    the symmetric split-BF16 contraction at its largest size (12 k-steps);
    the maps match FP64 and the FP32 contraction, which keeps all 385."""
    K = sred_amd.block_size_2d(11, 51)
    assert K == 385 and L.global_plan(10000, K, metric, L.BF16X3)["flags"] & 7 == 7
    img = O.normalize_input(synthetic.spots(80, 120, 16, seed=metric + 51, sigma=4.0))
    out = sred_amd.global_repetition_2d(img, 11, 51, 0.0, GM[metric], True, "bf16x3", resources=ctx)
    f32 = sred_amd.global_repetition_2d(img, 11, 51, 0.0, GM[metric], True, "fp32", resources=ctx)
    ref64 = O.global_map(metric, img, 11, 51, 0.0, True, fp64=True)
    assert np.abs(out - ref64).max() <= MAP_TOL
    assert np.abs(out - f32).max() <= MAP_TOL


@pytest.mark.parametrize("metric", [0, 1, 2, 3])
@pytest.mark.parametrize("b", [21, 31])
def test_block_maps_large_templates_bitexact(ctx, metric, b):
    img = synthetic.rings_and_blobs(240, 30, 30, seed=metric + b)
    tmpl = synthetic.ring_template(b, radius=b / 3.0)
    rb = utils.get_reference_block_2d(tmpl)
    inp = utils.get_input_image_2d(img)
    out = sred_amd.clutils.calculate_block_repetition_map_2d(BM[metric], inp, rb, 0.0, True, ctx)
    ref = O.block_map(metric, inp.image_array, rb.pixels, b, b, rb.mean, rb.std, 0.0, True)
    np.testing.assert_array_equal(out.reshape(img.shape), ref)


# ---------------------------------------------------------------------------
# SSIM with skipped pairs where the luminance factor is large
# ---------------------------------------------------------------------------
def test_ssim_skip_with_means_near_one_tenth(ctx):
    """Two textures with local means near 0.1 (where the product-denominator
    luminance factor (2 m_i m_j + c1)/(m_i^2 m_j^2 + c1) peaks near 100) and
    block sigmas far apart (a quiet background and a sparse bright texture):
    the symmetric kernel skips the cross pairs and the map matches FP64."""
    rng = np.random.default_rng(29)
    h, w, b = 150, 170, 5
    img = (0.1 + 0.003 * rng.standard_normal((h, w))).astype(np.float32)
    right = np.where(rng.random((h, w - w // 2)) < 0.05, 1.0, 0.05).astype(np.float32)
    img[:, w // 2:] = right
    img[0, 0], img[-1, -1] = 0.0, 1.0  # fixes the input normalisation to the identity
    inp = utils.get_input_image_2d(img)
    out = sred_amd.global_repetition_2d(img, b, b, 0.0, "SSIM", resources=ctx)
    n = (h - b + 1) * (w - b + 1)
    assert ctx.last_skipped() > 0.3 * n * n
    m32, _ = O.local_stats_f32(inp.image_array, b, b)
    inner = m32[2:-2, 2:-2]
    assert np.median(inner) == pytest.approx(0.1, abs=0.03)
    ref = O.global_map(O.SSIM, inp.image_array, b, b, 0.0, True, fp64=True)
    assert np.abs(out - ref).max() <= MAP_TOL


# ---------------------------------------------------------------------------
# weight sums with more than 2^21 blocks in one sigma bin
# ---------------------------------------------------------------------------
def test_weight_sums_many_blocks_in_one_bin(ctx):
    """1600 x 1600 pure noise with 11 x 11 blocks: 2.53e6 valid blocks whose
    sigmas (t = sigma sqrt(log2 e / h) close to 1.2) fall in one or two bins
    of the weight-sum expansion.  The bin's count moment then exceeds 2^21
    blocks, where 2^-42 fixed point would pass 2^63; the scale steps up to
    2^-40 and the Pearson map still matches FP64 on sampled rows."""
    rng = np.random.default_rng(31)
    img = O.normalize_input((0.5 + 0.05 * rng.standard_normal((1600, 1600))).astype(np.float32))
    out = sred_amd.global_repetition_2d(img, 11, 11, 0.0, "Pearson's R", True, "bf16x3", resources=ctx)
    m32, s32 = O.local_stats_f32(img, 11, 11)
    t = s32[5:-5, 5:-5].astype(np.float64) * np.sqrt(np.log2(np.e) / (O.noise_var(img) + 1e-7))
    counts = np.bincount(np.floor(t.ravel() / 0.5 + 0.5).astype(np.int64))
    assert counts.max() > 2 ** 21
    mask, _, _ = O.relevance_mask(s32, 5, 5, O.noise_var(img), 0.0)
    px = sample_pixels(out, mask, 24, seed=131)
    ref, _ = fp64_map_at(O, 0, img, 11, 11, 0.0, px)
    err = np.abs(out.ravel()[px] - ref)
    assert err.max() <= MAP_TOL, (err.max(), px[np.argmax(err)])


# ---------------------------------------------------------------------------
# super-panels: a call whose partial sums exceed the memory budget
# ---------------------------------------------------------------------------
def _superpanel_run(img, metric, c, budget_mb):
    """(one-launch map, super-panel map, launches) on a fresh context."""
    with sred_amd.Context(0) as c2:
        one = sred_amd.global_repetition_2d(img, 5, 5, c, GM[metric], resources=c2)
        _, launches1, pairs1 = c2.last_contraction()
        skip1 = c2.last_skipped()
        c2.set_option(L.OPT_PARTIAL_MEMORY_MB, budget_mb)
        out = sred_amd.global_repetition_2d(img, 5, 5, c, GM[metric], resources=c2)
        _, launches, pairs = c2.last_contraction()
        assert launches1 == 1 and pairs == pairs1
        assert c2.last_skipped() == skip1
    return one, out, launches


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_global_superpanels_small(metric):
    """Partial-sum budget 1 MB on a 300 x 260 image (up to 60 panels of 1280
    blocks, up to 17.6 MB of partial sums in one launch): super-panels of 10
    panels, so a few tens of launches of triangles and rectangles.  The map
    equals the
    one-launch map to FP32 rounding (only the FP64 order of the row totals
    differs), skips the same pairs, and matches FP64 (relevance constant 0.7)."""
    img = O.normalize_input(synthetic.spots(300, 260, 120, seed=90 + metric))
    one, out, launches = _superpanel_run(img, metric, 0.7, 1)
    assert launches > 1
    assert np.abs(out - one).max() <= 2e-6
    ref64 = O.global_map(metric, img, 5, 5, 0.7, True, fp64=True)
    assert np.abs(out - ref64).max() <= MAP_TOL


@pytest.mark.parametrize("metric", [0, 1, 2])
def test_global_superpanels_uneven(metric):
    """700 x 700, 5x5 (484,416 blocks in 237 panels of 2048; one launch needs
    439 MB of partial sums): with a 1 MB budget the blocks run as 30
    super-panels of 8 panels, the last of 5 (465 launches); the map equals
    the one-launch map to FP32 rounding and skips the same pairs."""
    img = O.normalize_input(synthetic.spot_ring_mixture(700, 700, seed=95 + metric))
    one, out, launches = _superpanel_run(img, metric, 0.0, 1)
    assert launches == 30 * 31 // 2
    assert np.abs(out - one).max() <= 2e-6


def test_global_2048_b11_default_budget():
    """2048 x 2048, 11x11 (4.16e6 blocks): its partial sums (34 GB in one
    launch) exceed the default 16 GB budget, so the call runs as two
    super-panels (three launches); the map matches FP64 on 8 seeded pixels
    plus its extremes."""
    from tests.helpers import fp64_map_at, sample_pixels
    img = O.normalize_input(synthetic.lattice_and_blobs(2048, seed=2, n_blobs=3600))
    with sred_amd.Context(0) as c2:
        out = sred_amd.global_repetition_2d(img, 11, 11, 0.0, "Pearson's R", resources=c2)
        assert c2.last_contraction()[1] == 3
    m32, s32 = O.local_stats_f32(img, 11, 11)
    mask, _, _ = O.relevance_mask(s32, 5, 5, O.noise_var(img), 0.0)
    px = sample_pixels(out, mask, 8, seed=11)
    ref, _ = fp64_map_at(O, 0, img, 11, 11, 0.0, px)
    err = np.abs(out.ravel()[px] - ref).max()
    print(f"2048^2 b11 Pearson (super-panels) vs FP64 at {px.size} pixels: max {err:.2e}")
    assert err <= MAP_TOL
