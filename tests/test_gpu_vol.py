"""GPU volume path (GlobalRepetition3D_ with its four metrics) against the
oracle, through the C ABI: the 3D statistics bit for bit; the relevance mask
and nPixels bit for bit; the normalised maps within 1e-4 of the float
restatement and of the FP64 form for both contraction precisions, with and
without a relevance constant, with flat regions (flat rows score 0, flat
columns leave AbsDiffStds); the 2D geometry untouched (covered by the 2D
tests); the errors."""
import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from sred_amd import clutils

pytestmark = pytest.mark.gpu


def _blobs(d, h, w, seed, n=24):
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.mgrid[0:d, 0:h, 0:w]
    v = np.full((d, h, w), 0.1)
    for _ in range(n):
        cz, cy, cx = rng.uniform(0, d), rng.uniform(0, h), rng.uniform(0, w)
        v += rng.uniform(0.5, 1.0) * np.exp(-((zz - cz) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2) / (2 * 1.8 ** 2))
    v += 0.03 * rng.standard_normal(v.shape)
    return ((v - v.min()) / (v.max() - v.min())).astype(np.float32)


@pytest.mark.parametrize("b", [(3, 3, 3), (5, 5, 5), (7, 5, 3)])
def test_stats3d_bitexact(ctx, b):
    vol = _blobs(14, 30, 34, 1)
    st = clutils.get_local_statistics_3d(ctx, vol, *b)
    m, s = O.local_stats3d_f32(vol, *b)
    np.testing.assert_array_equal(st.local_means, m)
    np.testing.assert_array_equal(st.local_stds, s)


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
@pytest.mark.parametrize("b,c", [((3, 3, 3), 0.0), ((5, 5, 5), 0.0), ((5, 5, 3), 0.8)])
def test_global_pearson3d(ctx, precision, b, c):
    vol = _blobs(16, 26, 30, 2)
    got = sred_amd.global_repetition_3d(vol, *b, c, "Pearson's R", True, precision, resources=ctx)
    inp = O.normalize_input(vol)  # the command's input step (Utils.getInputImage3D)
    ref = O.global_pearson3d(inp, *b, c=c)
    r64 = O.global_pearson3d(inp, *b, c=c, fp64=True)
    e32, e64 = np.abs(got - ref.out).max(), np.abs(got - r64.out).max()
    print(f"3d {b} c={c} {precision}: max|err| vs float {e32:.2e}, vs FP64 {e64:.2e}", flush=True)
    assert e32 <= 1e-4 and e64 <= 1e-4


@pytest.mark.parametrize("precision", ["bf16x3", "fp32"])
@pytest.mark.parametrize("metric", ["SSIM", "NRMSE", "Abs. diff. of StdDevs"])
@pytest.mark.parametrize("b,c", [((3, 3, 3), 0.0), ((5, 5, 3), 0.8), ((5, 5, 5), 0.0)])
def test_global_metrics3d(ctx, precision, metric, b, c):
    vol = _blobs(16, 26, 30, 5)
    vol[4:12, 6:16, 8:20] = 0.5  # flat region: flat rows, flat columns
    got = sred_amd.global_repetition_3d(vol, *b, c, metric, True, precision, resources=ctx)
    m = {"SSIM": O.SSIM, "NRMSE": O.NRMSE, "Abs. diff. of StdDevs": O.ABSDIFF}[metric]
    inp = O.normalize_input(vol)
    ref = O.global_repetition3d(inp, m, *b, c=c)
    r64 = O.global_repetition3d(inp, m, *b, c=c, fp64=True)
    e32, e64 = np.abs(got - ref.out).max(), np.abs(got - r64.out).max()
    print(f"3d {metric} {b} c={c} {precision}: max|err| vs float {e32:.2e}, vs FP64 {e64:.2e}", flush=True)
    assert e32 <= 1e-4 and e64 <= 1e-4
    assert sred_amd.global_repetition_3d(vol, *b, c, "Hu moments", resources=ctx) is None


def test_mask_and_device_entry(ctx):
    torch = pytest.importorskip("torch")
    vol = _blobs(12, 24, 28, 3)
    d, h, w = vol.shape
    dv = torch.from_numpy(vol).cuda()
    bufs = [torch.zeros_like(dv) for _ in range(6)]
    out, means, stds, mask, S, W = bufs
    L.check(L.load().sred_global_repetition_3d_dev(ctx.handle, L.PEARSON, L.BF16X3, dv.data_ptr(), w, h, d, 3, 3, 3,
                                                   0.6, 1, out.data_ptr(), means.data_ptr(), stds.data_ptr(),
                                                   mask.data_ptr(), S.data_ptr(), W.data_ptr()))
    ref = O.global_pearson3d(vol, 3, 3, 3, c=0.6)
    np.testing.assert_array_equal(mask.cpu().numpy(), ref.mask)
    m, s = O.local_stats3d_f32(vol, 3, 3, 3)
    np.testing.assert_array_equal(stds.cpu().numpy(), s)
    assert np.abs(out.cpu().numpy() - ref.out).max() <= 1e-4
    # raw sums: S and Wsum within the contraction's precision
    Sg, Wg = S.cpu().numpy(), W.cpu().numpy()
    sel = ref.W > 0
    assert np.abs(Wg[sel] - ref.W[sel]).max() <= 2e-5 * np.abs(ref.W[sel]).max()
    assert np.abs(Sg[sel] - ref.S[sel]).max() <= 2e-4 * np.abs(ref.S[sel]).max()


def test_errors(ctx):
    vol = _blobs(8, 16, 16, 4)
    with pytest.raises(sred_amd.SredError):  # block depth below 3
        L.check(L.load().sred_global_repetition_3d(ctx.handle, L.PEARSON, L.BF16X3, L.fptr(vol), 16, 16, 8, 3, 3, 1,
                                                   0.0, 1, L.fptr(np.zeros_like(vol))))
    with pytest.raises(sred_amd.SredError) as e:  # unknown metric id
        L.check(L.load().sred_global_repetition_3d(ctx.handle, 7, L.BF16X3, L.fptr(vol), 16, 16, 8, 3, 3, 3,
                                                   0.0, 1, L.fptr(np.zeros_like(vol))))
    assert e.value.code == L.SRED_ERR_UNKNOWN_METRIC
    assert L.load().sred_block_size_3d(5, 5, 5) == 33


@pytest.mark.parametrize("metric", ["Pearson's R", "SSIM", "NRMSE"])
def test_global3d_257_pixel_blocks(ctx, metric):
    """9x9x9 blocks hold 257 pixels: reflected into 256 dimensions they stay
    on the split-BF16 contraction; the map matches FP64."""
    assert O.footprint3d(9, 9, 9)[0] == 257
    vol = _blobs(18, 22, 20, 7)
    got = sred_amd.global_repetition_3d(vol, 9, 9, 9, 0.0, metric, True, "bf16x3", resources=ctx)
    m = {"Pearson's R": O.PEARSON, "SSIM": O.SSIM, "NRMSE": O.NRMSE}[metric]
    inp = O.normalize_input(vol)
    r64 = (O.global_pearson3d(inp, 9, 9, 9, fp64=True) if m == O.PEARSON
           else O.global_repetition3d(inp, m, 9, 9, 9, fp64=True))
    e64 = np.abs(got - r64.out).max()
    print(f"3d {metric} 9x9x9 bf16x3: max|err| vs FP64 {e64:.2e}", flush=True)
    assert e64 <= 1e-4


@pytest.mark.parametrize("metric", ["Pearson's R", "SSIM", "NRMSE"])
def test_global3d_superpanels(metric):
    """A volume's map with the partial-sum budget cut to 1 MB (9,408 blocks in
    37 panels of 256: super-panels of 21 panels, so two triangles and one
    rectangle; NRMSE takes the weighted 3D epilogue) equals the one-launch
    map to FP32 rounding and matches FP64."""
    vol = _blobs(16, 26, 30, 7)
    with sred_amd.Context(0) as c2:
        one = sred_amd.global_repetition_3d(vol, 3, 3, 3, 0.0, metric, True, "bf16x3", resources=c2)
        c2.set_option(L.OPT_PARTIAL_MEMORY_MB, 1)
        got = sred_amd.global_repetition_3d(vol, 3, 3, 3, 0.0, metric, True, "bf16x3", resources=c2)
        assert c2.last_contraction()[1] == 3
    assert np.abs(got - one).max() <= 2e-6
    m = {"Pearson's R": O.PEARSON, "SSIM": O.SSIM, "NRMSE": O.NRMSE}[metric]
    inp = O.normalize_input(vol)
    r64 = (O.global_pearson3d(inp, 3, 3, 3, fp64=True) if m == O.PEARSON
           else O.global_repetition3d(inp, m, 3, 3, 3, fp64=True))
    assert np.abs(got - r64.out).max() <= 1e-4
