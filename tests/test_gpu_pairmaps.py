"""GPU all-pairs vector maps against the oracle: the global Hu map
(kernelGetHuMap.cl) and the phase correlation map (textbook specification of
kernelGetPhaseCorrelationMap.cl).

* Hu invariants: bit for bit (same float order, no contraction).
* Global Hu map: pair terms in FP32 with hardware reciprocal/exp2, row sums in
  FP64, against the oracle's FP64 sums: rtol 2e-5.
* Phase correlation: FP64 with the same unit spectra; the peak decisions are
  identical away from exact ties, so the maps agree to the FP64 summation
  order: rtol 1e-12.
* Full sizes (1024x1024 b11 for Hu, 256x256 b5 for phase) are checked by
  size-independent properties: the sorted closed form of the Hu map when all
  stds are equal, the periodic-image closed form of the phase map, and
  sampled reference pixels against the oracle.
"""
import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from sred_amd import synthetic, utils

pytestmark = pytest.mark.gpu


def _inp(img):
    h, w = img.shape
    return utils.InputImage2D(img, w, h, img.size)


def _stats(img, bw, bh):
    m, s = O.local_stats_f32(img, bw, bh)
    return sred_amd.clutils.LocalStatistics(m.reshape(-1), s.reshape(-1)), m, s


# ---------------------------------------------------------------- Hu

@pytest.mark.parametrize("shape,bw,bh,seed", [((40, 52), 5, 5, 1), ((61, 47), 3, 9, 2), ((128, 96), 11, 11, 3),
                                               ((33, 70), 7, 3, 4), ((90, 90), 21, 15, 5)])
def test_hu_invariants_bitexact(ctx, shape, bw, bh, seed):
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 30, seed=seed))
    st, m, _ = _stats(img, bw, bh)
    got = sred_amd.get_hu_invariants_2d(_inp(img), st, bw, bh, ctx).reshape(shape)
    np.testing.assert_array_equal(got, O.hu_invariants(img, m, bw, bh))


@pytest.mark.parametrize("shape,bw,bh,seed", [((48, 40), 5, 5, 1), ((37, 64), 3, 7, 2), ((70, 66), 11, 11, 3)])
def test_global_hu_vs_oracle(ctx, shape, bw, bh, seed):
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 20, seed=seed))
    st, m, s = _stats(img, bw, bh)
    got = sred_amd.calculate_global_hu_map_2d(_inp(img), st, bw, bh, ctx).reshape(shape)
    want = O.global_hu_map(img, m, s, bw, bh)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-6 * np.abs(want).max())


def test_global_hu_rows_list(ctx):
    shape, bw, bh = (50, 45), 5, 7
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 20, seed=7))
    st, m, s = _stats(img, bw, bh)
    rng = np.random.default_rng(7)
    rows = np.concatenate([rng.integers(0, img.size, 40), [0, -3, img.size + 5, img.size - 1]]).astype(np.int32)
    got = sred_amd.calculate_global_hu_map_2d(_inp(img), st, bw, bh, ctx, unique_coords=rows).reshape(shape)
    want = O.global_hu_map(img, m, s, bw, bh, rows=rows)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-6 * np.abs(want).max())
    assert not got[want == 0].any()


def test_global_hu_full_size_equal_stds(ctx):
    """1024x1024, 11x11, all stds equal: w = 1, map = sum_j |hu_j - hu_i| (sorted prefix sums)."""
    n, b = 1024, 11
    img = O.normalize_input(synthetic.lattice_and_blobs(n, seed=2))
    st, m, _ = _stats(img, b, b)
    st = sred_amd.clutils.LocalStatistics(st.local_means, np.full(n * n, 0.25, np.float32))
    got = sred_amd.calculate_global_hu_map_2d(_inp(img), st, b, b, ctx).reshape(n, n)
    hu = O.hu_invariants(img, m, b, b)
    r = b // 2
    inner = hu[r:n - r, r:n - r].astype(np.float64).reshape(-1)
    v = np.sort(inner)
    cs = np.concatenate([[0.0], np.cumsum(v)])
    k = np.searchsorted(v, inner, side="left")
    want = (inner * k - cs[k]) + (cs[-1] - cs[k] - inner * (v.size - k))
    np.testing.assert_allclose(got[r:n - r, r:n - r].reshape(-1), want, rtol=2e-5, atol=1e-6 * want.max())


def test_global_hu_full_size_sampled_rows(ctx):
    n, b = 1024, 11
    img = O.normalize_input(synthetic.lattice_and_blobs(n, seed=2))
    st, m, s = _stats(img, b, b)
    full = sred_amd.calculate_global_hu_map_2d(_inp(img), st, b, b, ctx).reshape(n, n)
    rng = np.random.default_rng(11)
    rows = (rng.integers(b, n - b, 24) * n + rng.integers(b, n - b, 24)).astype(np.int32)
    want = O.global_hu_map(img, m, s, b, b, rows=rows).reshape(-1)[rows]
    np.testing.assert_allclose(full.reshape(-1)[rows], want, rtol=2e-5)



def test_global_hu_full_map_ordered_fallback():
    """A full Hu map whose panel-pair partial sums (npanels x N doubles, panels
    of 2048 interior pixels) exceed a 1 MB budget runs the ordered-pair
    kernel over every row instead of the symmetric one: the two maps agree to
    the FP32 sub-sum order, and the ordered map matches the oracle on sampled
    rows."""
    shape, b = (300, 280), 5
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 60, seed=31))
    st, m, s = _stats(img, b, b)
    n_int = (shape[0] - 2 * (b // 2)) * (shape[1] - 2 * (b // 2))
    assert -(-n_int // 2048) * n_int * 8 > (1 << 20)
    with sred_amd.Context(0) as c2:
        sym = sred_amd.calculate_global_hu_map_2d(_inp(img), st, b, b, c2).reshape(shape)
        c2.set_option(L.OPT_PARTIAL_MEMORY_MB, 1)
        ordered = sred_amd.calculate_global_hu_map_2d(_inp(img), st, b, b, c2).reshape(shape)
    np.testing.assert_allclose(ordered, sym, rtol=1e-5, atol=1e-6 * np.abs(sym).max())
    rng = np.random.default_rng(31)
    rows = (rng.integers(b, shape[0] - b, 16) * shape[1] + rng.integers(b, shape[1] - b, 16)).astype(np.int32)
    want = O.global_hu_map(img, m, s, b, b, rows=rows).reshape(-1)[rows]
    np.testing.assert_allclose(ordered.reshape(-1)[rows], want, rtol=2e-5)

# ---------------------------------------------------------------- phase correlation

@pytest.mark.parametrize("shape,bw,bh,seed", [((24, 22), 3, 3, 1), ((30, 27), 5, 5, 2), ((26, 31), 3, 7, 3),
                                               ((28, 25), 7, 5, 4), ((25, 25), 9, 9, 5), ((24, 26), 11, 11, 6),
                                               ((22, 30), 11, 3, 7)])
def test_phase_map_vs_oracle(ctx, shape, bw, bh, seed):
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 6, seed=seed))
    st, m, s = _stats(img, bw, bh)
    got = sred_amd.calculate_phase_correlation_map_2d(_inp(img), st, bw, bh, ctx).reshape(shape)
    want = O.phase_map(img, m, s, bw, bh)
    np.testing.assert_allclose(got, want.astype(np.float32), rtol=1e-6, atol=0)
    np.testing.assert_allclose(got, want, rtol=1e-7, atol=1e-7 * np.abs(want).max())


def test_phase_map_rows_list(ctx):
    shape, bw, bh = (29, 34), 5, 3
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 8, seed=12))
    st, m, s = _stats(img, bw, bh)
    rows = np.array([0, 5 * 34 + 5, 6 * 34 + 30, 28 * 34 + 2, 10 * 34 + 17, 10 * 34 + 17, 999], np.int32)
    got = sred_amd.calculate_phase_correlation_map_2d(_inp(img), st, bw, bh, ctx, unique_coords=rows).reshape(shape)
    want = O.phase_map(img, m, s, bw, bh, rows=rows)
    np.testing.assert_allclose(got, want, rtol=1e-7)
    assert not got[want == 0].any()


def _periodic(w, h, bw, bh, seed):
    rng = np.random.default_rng(seed)
    pat = rng.random((bh, bw)).astype(np.float32)
    return np.tile(pat, (h // bh + 1, w // bw + 1))[:h, :w].copy()


def test_phase_full_size_periodic_closed_form(ctx):
    """256x256, 5x5 on an image periodic with the block: every block is a
    circular shift of every other, so the displacement of (i, j) is the
    wrapped position difference."""
    n, b = 256, 5
    img = _periodic(n, n, b, b, seed=21)
    st, m, s = _stats(img, b, b)
    got = sred_amd.calculate_phase_correlation_map_2d(_inp(img), st, b, b, ctx).reshape(n, n)
    r = b // 2
    ys, xs = np.mgrid[r:n - r, r:n - r]
    ys, xs = ys.reshape(-1), xs.reshape(-1)
    sj = s[ys, xs]
    rng = np.random.default_rng(21)
    for t in rng.integers(0, ys.size, 12):
        y, x = ys[t], xs[t]
        dx = (xs - x) % b
        dy = (ys - y) % b
        dx = np.where(dx <= r, dx, dx - b)
        dy = np.where(dy <= r, dy, dy - b)
        si = np.float32(s[y, x])
        wgt = np.where(sj == si, np.float32(1),
                       np.float32(1) - np.float32(2) * (np.abs(si - sj) / (np.abs(si) + np.abs(sj))).astype(np.float32))
        want = np.sum(np.sqrt(dx * dx + dy * dy) * wgt.astype(np.float64))
        assert got[y, x] == pytest.approx(want, rel=1e-7)


def test_phase_full_size_sampled_rows(ctx):
    n, b = 256, 5
    img = O.normalize_input(synthetic.spots(n, n, 200, seed=0))
    st, m, s = _stats(img, b, b)
    full = sred_amd.calculate_phase_correlation_map_2d(_inp(img), st, b, b, ctx).reshape(-1)
    rng = np.random.default_rng(5)
    rows = (rng.integers(b, n - b, 8) * n + rng.integers(b, n - b, 8)).astype(np.int32)
    want = O.phase_map(img, m, s, b, b, rows=rows).reshape(-1)[rows]
    np.testing.assert_allclose(full[rows], want, rtol=1e-7)



def test_phase_full_map_ordered_fallback():
    """A full phase correlation map beyond a 1 MB partial-sum budget (panels
    of 256 interior pixels) runs the ordered-pair kernel, each lane taking
    the lower interior index of its pair as the leading spectrum: the same
    per-pair values as the symmetric kernel, so the maps agree to the FP64
    summation order; sampled rows match the oracle."""
    shape, b = (150, 160), 5
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 30, seed=33))
    st, m, s = _stats(img, b, b)
    n_int = (shape[0] - 2 * (b // 2)) * (shape[1] - 2 * (b // 2))
    assert -(-n_int // 256) * n_int * 8 > (1 << 20)
    with sred_amd.Context(0) as c2:
        sym = sred_amd.calculate_phase_correlation_map_2d(_inp(img), st, b, b, c2).reshape(shape)
        c2.set_option(L.OPT_PARTIAL_MEMORY_MB, 1)
        ordered = sred_amd.calculate_phase_correlation_map_2d(_inp(img), st, b, b, c2).reshape(shape)
    np.testing.assert_allclose(ordered, sym, rtol=1e-12, atol=0)
    rng = np.random.default_rng(33)
    rows = (rng.integers(b, shape[0] - b, 8) * shape[1] + rng.integers(b, shape[1] - b, 8)).astype(np.int32)
    want = O.phase_map(img, m, s, b, b, rows=rows).reshape(-1)[rows]
    np.testing.assert_allclose(ordered.reshape(-1)[rows], want, rtol=1e-7)

def test_pairmaps_device_path(ctx):
    torch = pytest.importorskip("torch")
    img = O.normalize_input(synthetic.spots(40, 36, 8, seed=30))
    m, s = O.local_stats_f32(img, 5, 5)
    d = [torch.from_numpy(a).cuda() for a in (img, m, s)]
    out = torch.full_like(d[0], 7.0)
    rows = np.array([5 * 40 + 5, 20 * 40 + 30], np.int32)
    for fn, ref in [(L.load().sred_global_hu_2d_dev, O.global_hu_map), (L.load().sred_phase_correlation_2d_dev,
                                                                        O.phase_map)]:
        L.check(fn(ctx.handle, d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr(), 40, 36, 5, 5, rows.ctypes.data, 2,
                   out.data_ptr()))
        torch.cuda.synchronize()
        want = ref(img, m, s, 5, 5, rows=rows)
        np.testing.assert_allclose(out.cpu().numpy(), want, rtol=2e-5)


def test_pairmaps_errors(ctx):
    img = np.zeros((30, 30), np.float32)
    st = sred_amd.clutils.LocalStatistics(np.zeros(900, np.float32), np.zeros(900, np.float32))
    with pytest.raises(sred_amd.SredError):
        sred_amd.calculate_phase_correlation_map_2d(_inp(img), st, 13, 13, ctx)
    with pytest.raises(sred_amd.SredError):
        sred_amd.calculate_global_hu_map_2d(_inp(img), st, 4, 5, ctx)
    small = np.zeros((4, 30), np.float32)
    st2 = sred_amd.clutils.LocalStatistics(np.zeros(120, np.float32), np.zeros(120, np.float32))
    with pytest.raises(sred_amd.SredError):
        sred_amd.calculate_global_hu_map_2d(_inp(small), st2, 5, 5, ctx)


def test_phase_zero_image(ctx):
    """All blocks zero after normalisation: no positive peak anywhere, map 0."""
    img = np.full((20, 20), 0.5, np.float32)
    st, m, s = _stats(img, 3, 3)
    got = sred_amd.calculate_phase_correlation_map_2d(_inp(img), st, 3, 3, ctx)
    assert not got.any()
