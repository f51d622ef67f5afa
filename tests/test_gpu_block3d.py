"""GPU block repetition of volumes (BlockRepetition3D_ ->
CLUtils.calculateBlockRepetitionMap3D, as intended) against the oracle,
through the C ABI: the maps of all four metrics bit for bit, with and
without a relevance constant and normalisation, for several block shapes
(Pearson with and without its LDS cache),
with flat regions (test std 0) and a flat reference block; the device entry
with its statistics and mask outputs; the benchmark's 128 x 128 x 64 volume;
the errors."""
import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L

pytestmark = pytest.mark.gpu

NAMES = {O.PEARSON: "Pearson's R", O.SSIM: "SSIM", O.NRMSE: "NRMSE (inverted)", O.ABSDIFF: "Abs. Diff. of StdDevs."}


def _vol(d, h, w, seed, n=16):
    rng = np.random.default_rng(seed)
    zz, yy, xx = np.mgrid[0:d, 0:h, 0:w]
    v = np.full((d, h, w), 0.1)
    for _ in range(n):
        cz, cy, cx = rng.uniform(0, d), rng.uniform(0, h), rng.uniform(0, w)
        v += rng.uniform(0.5, 1.0) * np.exp(-((zz - cz) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2) / (2 * 1.7 ** 2))
    v += 0.03 * rng.standard_normal(v.shape)
    return (200.0 * v + 10.0).astype(np.float32)


def _cut(vol, z, y, x, bw, bh, bd, seed):
    blk = vol[z - bd // 2:z + bd // 2 + 1, y - bh // 2:y + bh // 2 + 1, x - bw // 2:x + bw // 2 + 1].copy()
    return blk + np.random.default_rng(seed).normal(0.0, 2.0, blk.shape).astype(np.float32)


def _oracle_map(metric, vol, blk, c, normalize):
    bd, bh, bw = blk.shape
    px, m, s = O.reference_block3d(blk)
    return O.block_map3d(metric, O.normalize_input(vol), px, bw, bh, bd, m, s, c, normalize)


@pytest.mark.parametrize("metric", [O.PEARSON, O.SSIM, O.NRMSE, O.ABSDIFF])
@pytest.mark.parametrize("b,c,normalize", [((3, 3, 3), 0.0, True), ((5, 5, 5), 0.0, True), ((7, 5, 3), 0.8, True),
                                           ((3, 5, 7), 0.0, False), ((5, 3, 3), 1.5, False),
                                           ((9, 7, 5), 0.0, True)])  # K = 99: Pearson without the LDS cache
def test_block3d_bitexact(ctx, metric, b, c, normalize):
    vol = _vol(14, 30, 37, 1)
    vol[5:10, 10:20, 12:24] = 60.0  # flat region: test std 0
    blk = _cut(vol, 7, 12, 20, *b, 3)
    print(f"block3d {NAMES[metric]} {b} c={c} normalize={normalize}", flush=True)
    got = sred_amd.block_repetition_3d(blk, vol, c, NAMES[metric], normalize, resources=ctx)
    want = _oracle_map(metric, vol, blk, c, normalize)
    np.testing.assert_array_equal(got, want)
    assert got.any()


def test_flat_reference_block(ctx):
    vol = _vol(10, 20, 24, 2)
    vol[3:7, 5:15, 5:15] = 60.0
    blk = np.full((3, 3, 3), 5.0, np.float32)  # std 0 after the two normalisations
    for metric in (O.PEARSON, O.SSIM, O.NRMSE, O.ABSDIFF):
        got = sred_amd.block_repetition_3d(blk, vol, 0.0, NAMES[metric], False, resources=ctx)
        np.testing.assert_array_equal(got, _oracle_map(metric, vol, blk, 0.0, False))


def test_device_entry_outputs(ctx):
    torch = pytest.importorskip("torch")
    vol = O.normalize_input(_vol(12, 26, 30, 3))
    d, h, w = vol.shape
    blk = _cut(vol, 6, 12, 15, 5, 5, 3, 4)
    px, m, s = O.reference_block3d(blk)
    dv = torch.from_numpy(vol).cuda()
    out, means, stds, mask = (torch.zeros_like(dv) for _ in range(4))
    L.check(L.load().sred_block_repetition_3d_dev(ctx.handle, L.NRMSE, dv.data_ptr(), w, h, d, L.fptr(px), 5, 5, 3,
                                                  m, s, 0.6, 1, out.data_ptr(), means.data_ptr(), stds.data_ptr(),
                                                  mask.data_ptr()))
    mo, so = O.local_stats3d_f32(vol, 5, 5, 3)
    np.testing.assert_array_equal(means.cpu().numpy(), mo)
    np.testing.assert_array_equal(stds.cpu().numpy(), so)
    np.testing.assert_array_equal(mask.cpu().numpy(), O.global_pearson3d(vol, 5, 5, 3, c=0.6).mask)
    np.testing.assert_array_equal(out.cpu().numpy(), O.block_map3d(O.NRMSE, vol, px, 5, 5, 3, m, s, 0.6))
    ms, launches, units = ctx.last_contraction()
    assert ms > 0.0 and launches == 1 and units == vol.size


def test_bench_volume(ctx):
    # the benchmark's volume: 128 x 128 x 64 voxels, 5 x 5 x 5 blocks
    vol = _vol(64, 128, 128, 5, n=300)
    blk = _cut(vol, 30, 60, 70, 5, 5, 5, 6)
    for metric in (O.PEARSON, O.SSIM):
        got = sred_amd.block_repetition_3d(blk, vol, 0.0, NAMES[metric], True, resources=ctx)
        np.testing.assert_array_equal(got, _oracle_map(metric, vol, blk, 0.0, True))


def test_errors(ctx):
    vol = _vol(6, 12, 12, 7)
    blk = _cut(vol, 3, 6, 6, 3, 3, 3, 1)
    assert sred_amd.block_repetition_3d(blk, vol, 0.0, "NRMSE", resources=ctx) is None  # the global list's name
    with pytest.raises(ValueError):  # block deeper than the volume
        sred_amd.block_repetition_3d(np.ones((7, 3, 3), np.float32), vol, resources=ctx)
    with pytest.raises(ValueError):
        sred_amd.block_repetition_3d(np.ones((3, 4, 3), np.float32), vol, resources=ctx)
    px, m, s = O.reference_block3d(blk)
    with pytest.raises(sred_amd.SredError):
        L.check(L.load().sred_block_repetition_3d(ctx.handle, L.PEARSON, L.fptr(vol), 12, 12, 6, L.fptr(px), 3, 3, 7,
                                                  m, s, 0.0, 1, L.fptr(np.zeros_like(vol))))
    with pytest.raises(sred_amd.SredError) as e:
        L.check(L.load().sred_block_repetition_3d(ctx.handle, 9, L.fptr(vol), 12, 12, 6, L.fptr(px), 3, 3, 3, m, s,
                                                  0.0, 1, L.fptr(np.zeros_like(vol))))
    assert e.value.code == L.SRED_ERR_UNKNOWN_METRIC
