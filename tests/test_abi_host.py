"""CPU-only checks of the C ABI: the library loads and exports every symbol
include/sred.h declares, and the host-side Utils mirror (plain C++, no GPU
call) matches the oracle and the reference's known answers."""
import os
import re

import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from tests.helpers import kats, random42_inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
K = kats()


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sred_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = L.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert names == set(L.EXPORTS)


def test_version_and_metric_names():
    lib = L.load()
    assert lib.sred_version() >= 100
    for i, m in enumerate(sred_amd.GLOBAL_METRICS):
        assert lib.sred_metric_from_name(m.encode(), 0) == i
    for i, m in enumerate(sred_amd.BLOCK_METRICS):
        assert lib.sred_metric_from_name(m.encode(), 1) == i
    assert lib.sred_metric_from_name(b"Hausdorff", 0) == L.SRED_ERR_UNKNOWN_METRIC
    assert lib.sred_metric_from_name(b"NRMSE", 1) == L.SRED_ERR_UNKNOWN_METRIC  # block list says "NRMSE (inverted)"


def test_block_size_table():
    for b, k in K["block_size_table"]["sizes"].items():
        assert sred_amd.block_size_2d(int(b), int(b)) == k
    for bw, bh in [(3, 7), (9, 5), (21, 11), (63, 63)]:
        assert sred_amd.block_size_2d(bw, bh) == O.footprint(bw, bh)[0]


def test_input_normalisation_matches_oracle_bitwise():
    rng = np.random.default_rng(0)
    img = (rng.random((37, 29)) * 1000).astype(np.float32)
    a = sred_amd.get_input_image_2d(img).image_array
    np.testing.assert_array_equal(a, O.normalize_input(img))
    k = K["input_image_2d_normalize"]
    np.testing.assert_allclose(sred_amd.get_input_image_2d(k["image"]).image_array, k["expected"], atol=k["tol"])
    # a method string other than "Simplex" or "Quad/Octree" stabilises nothing
    # (Utils.java:674, :686 compare the strings by identity)
    np.testing.assert_array_equal(
        sred_amd.get_input_image_2d(img, stabilise_noise_variance=True, gat_method="none").image_array,
        O.normalize_input(img))


def test_reference_block_kat_and_oracle():
    k = K["reference_block_2d"]
    rb = sred_amd.get_reference_block_2d(np.array(k["block"], np.float32))
    assert abs(rb.mean - k["expected_mean"]) <= k["tol"] and abs(rb.std - k["expected_std"]) <= k["tol"]
    assert sred_amd.get_reference_block_2d(np.array(k["even_block"], np.float32)) is None
    rng = np.random.default_rng(1)
    blk = rng.random((15, 11)).astype(np.float32)
    rb = sred_amd.get_reference_block_2d(blk)
    px, m, s = O.reference_block(blk)
    np.testing.assert_array_equal(rb.pixels, px)
    assert (rb.mean, rb.std) == (m, s)


def test_reference_block_3d_and_input_3d_match_oracle():
    rng = np.random.default_rng(2)
    for shape in [(3, 3, 3), (5, 7, 5), (7, 5, 9), (3, 11, 3)]:
        blk = rng.random(shape).astype(np.float32)
        rb = sred_amd.get_reference_block_3d(blk)
        px, m, s = O.reference_block3d(blk)
        np.testing.assert_array_equal(rb.pixels, px)
        assert (rb.mean, rb.std, rb.size) == (m, s, len(px))
        assert (rb.width, rb.height, rb.depth) == (shape[2], shape[1], shape[0])
    for bad in [(3, 4, 3), (1, 3, 3)]:
        with pytest.raises(ValueError):
            sred_amd.get_reference_block_3d(np.ones(bad, np.float32))
    vol = (rng.random((4, 9, 7)) * 50 - 3).astype(np.float32)
    inp = sred_amd.get_input_image_3d(vol)
    np.testing.assert_array_equal(inp.image_array, O.normalize_input(vol))
    assert (inp.width, inp.height, inp.depth, inp.size) == (7, 9, 4, 252)
    np.testing.assert_array_equal(sred_amd.get_input_image_3d(vol, True, "Simplex").image_array, inp.image_array)
    assert sred_amd.get_input_image_3d(vol[:2]) is None


def test_noise_variance_and_mask_random42_kat():
    k = K["relevance_mask_random42"]
    img, stds = random42_inputs(k)
    rm = sred_amd.get_relevance_mask(img, 50, 50, 2, 2, stds, k["relevance_constant"])
    assert np.float64(np.float32(rm.noise_mean_variance)) == k["expected_noise_var"]
    assert np.float32(rm.relevance_threshold) == np.float32(k["expected_threshold_f32"])
    mask, thr, npx = O.relevance_mask(stds, 2, 2, O.noise_var(img), k["relevance_constant"])
    np.testing.assert_array_equal(rm.relevance_mask.reshape(50, 50), mask)


def test_noise_variance_matches_oracle_large_tiles():
    rng = np.random.default_rng(2)
    img = rng.random((400, 300)).astype(np.float32)  # > CIF: 16x16 tiles
    assert sred_amd.get_mean_noise_var_2d(img, 300, 400) == O.noise_var(img)
    bad = img.copy()
    bad[5, 5] = np.inf
    with pytest.raises(sred_amd.SredError):
        sred_amd.get_mean_noise_var_2d(bad, 300, 400)
    with pytest.raises(sred_amd.SredError):
        sred_amd.get_mean_noise_var_2d(np.zeros((5, 5), np.float32), 5, 5)  # no 8x8 tile


def test_apply_mask_and_normalise_kats():
    k = K["apply_mask_2d"]
    a = np.array(k["image"], np.float32)
    sred_amd.apply_mask_2d(a, 3, 3, np.array(k["mask"], np.float32))
    np.testing.assert_array_equal(a, np.array(k["expected"], np.float32))
    k = K["normalize_image_2d_mask"]
    out = sred_amd.normalize_image_2d(np.array(k["image"], np.float32), 4, 4, *k["border"], np.array(k["mask"], np.float32))
    np.testing.assert_allclose(out, np.array(k["expected"]), atol=k["tol"])
    k = K["normalize_image_2d_nonbinary"]
    with pytest.raises(ValueError):
        sred_amd.normalize_image_2d(np.array(k["image"], np.float32), 4, 4, *k["border"], np.array(k["mask"], np.float32))


def test_normalise_matches_oracle_bitwise():
    rng = np.random.default_rng(4)
    m = rng.random((30, 31)).astype(np.float32)
    mask = (rng.random((30, 31)) > 0.3).astype(np.float32)
    for mk in (None, mask):
        a = sred_amd.normalize_image_2d(m, 31, 30, 3, 2, mk)
        np.testing.assert_array_equal(a, O.normalize_map(m, 3, 2, mk))


# ---------------------------------------------------------------------------
# launch geometry of the global contraction (host only)
# ---------------------------------------------------------------------------
NV_CASES = [1, 2, 63, 64, 255, 256, 257, 1000, 3876, 8320, 27936, 63504, 254016, 1028196, 4137156]
K_CASES = [5, 13, 29, 33, 49, 65, 81, 113, 149, 197, 253, 257, 289, 317, 377, 384, 385, 386, 441, 709, 3001]


@pytest.mark.parametrize("precision", [L.FP32, L.BF16X3])
@pytest.mark.parametrize("sym", [0, 1])
def test_contraction_plan_invariants(precision, sym):
    for nv in NV_CASES:
        for k in K_CASES:
            for metric in range(4):
                p = L.global_plan(nv, k, metric, precision, sym)
                # split BF16 up to K_pad = 256 (384 for the symmetric contraction),
                # blocks of K = 32m + 1 reflected into K - 1 dimensions (257 and 385
                # included); larger blocks take the FP32 contraction
                split = precision == L.BF16X3 and metric != L.ABSDIFF
                kd = k - 1 if split and (k - 1) % 32 == 0 else k
                mfma = split and (kd + 31) // 32 <= (12 if sym else 8)
                assert bool(p["flags"] & 1) == mfma
                assert bool(p["flags"] & 4) == (mfma and kd == k - 1)
                is_sym = bool(p["flags"] & 2)
                assert is_sym == (mfma and sym == 1)
                assert p["nv_pad"] >= nv and p["nv_pad"] - nv < max(p["rows"], 256)
                assert p["partial_elems"] == p["nchunks"] * p["nv_pad"]
                if is_sym:
                    npan = p["nv_pad"] // p["rows"]
                    assert p["nv_pad"] % p["rows"] == 0 and p["rows"] % 256 == 0 and p["rows"] <= 2048
                    assert p["nchunks"] == npan and p["grid"] == npan * (npan + 1) // 2
                    assert 2 * p["pairs"] >= p["nv_pad"] ** 2
                else:
                    assert p["nv_pad"] % p["rows"] == 0 and p["cols_per_chunk"] % 64 == 0
                    assert p["nchunks"] * p["cols_per_chunk"] >= p["nv_pad"]
                    assert (p["nchunks"] - 1) * p["cols_per_chunk"] < p["nv_pad"]
                    assert p["grid"] == (p["nv_pad"] // p["rows"]) * p["nchunks"]
                    if mfma:
                        assert p["cols_per_chunk"] <= 2048


def _decode_panel_pair(q):
    """The symmetric kernel's (p, c) decode of a workgroup index."""
    c = int((np.sqrt(np.float32(8.0 * q + 1.0)) - 1.0) * 0.5)
    while (c + 1) * (c + 2) // 2 <= q:
        c += 1
    while c * (c + 1) // 2 > q:
        c -= 1
    return q - c * (c + 1) // 2, c


@pytest.mark.parametrize("nv,k,metric", [(49, 5, 0), (435, 5, 1), (3876, 13, 0), (8320, 5, 2), (27936, 29, 1),
                                         (63504, 13, 0), (254016, 49, 0), (1028196, 81, 0)])
def test_contraction_plan_covers_each_pair_once(nv, k, metric):
    for sym in (1, 0):
        p = L.global_plan(nv, k, metric, L.BF16X3, sym)
        npad, slots = p["nv_pad"], np.zeros((p["nchunks"], p["nv_pad"]), np.int32)
        if p["flags"] & 2:
            P = p["rows"]
            npan = npad // P
            seen = np.zeros((npan, npan), np.int32)
            for q in range(p["grid"]):
                a, c = _decode_panel_pair(q)
                assert 0 <= a <= c < npan
                seen[a, c] += 1
                slots[c, a * P:(a + 1) * P] += 1           # row sums of panel a
                if a != c:
                    slots[a, c * P:(c + 1) * P] += 1       # column sums of panel c
            assert (np.triu(seen) == np.triu(np.ones_like(seen))).all()
        else:
            rows, cpc, nrt = p["rows"], p["cols_per_chunk"], npad // p["rows"]
            for q in range(p["grid"]):
                rt, ch = q % nrt, q // nrt
                assert ch < p["nchunks"] and (rt + 1) * rows <= npad
                assert ch * cpc < npad
                slots[ch, rt * rows:(rt + 1) * rows] += 1
        assert (slots == 1).all()  # every partial slot has exactly one writer


def test_host_f_quantile_and_regression():
    # the library's host helpers of the quadtree GAT estimate (no GPU):
    # the F quantile against scipy and closed forms, the regression against
    # numpy's least squares and an exact line
    st = pytest.importorskip("scipy.stats")
    from sred_amd import utils
    p = float(np.float32(0.01))
    for d in [1, 2, 3, 15, 63, 1023, 16383, 262143]:
        q = utils.f_quantile(d, d, p)
        assert q == pytest.approx(st.f.ppf(p, d, d), rel=1e-11)
        assert np.float32(q) == np.float32(O.f_quantile(d, d, p))
    assert utils.f_quantile(2, 2, 0.3) == pytest.approx(0.3 / 0.7, rel=1e-14)  # F(2,2): x/(1+x)
    assert utils.f_quantile(40, 40, 0.5) == pytest.approx(1.0, abs=1e-14)
    assert utils.f_quantile(7, 12, 0.2) == pytest.approx(st.f.ppf(0.2, 7, 12), rel=1e-11)
    with pytest.raises(sred_amd.SredError):
        utils.f_quantile(0, 3, 0.5)
    qt = utils.QuadTree(8, 8, 4, 0.01)
    rng = np.random.default_rng(4)
    x = rng.random(200) * 50
    y = 1.7 * x - 3.25 + rng.normal(0, 0.5, 200)
    s, i = qt.perform_linear_regression(np.stack([x, y], 1))
    ps, pi = np.polyfit(x, y, 1)
    assert s == pytest.approx(ps, rel=1e-12) and i == pytest.approx(pi, rel=1e-11)
    assert (s, i) == O.simple_regression(x, y)
    s, i = qt.perform_linear_regression([[1.0, 5.0], [2.0, 7.0], [4.0, 11.0]])
    assert (s, i) == (2.0, 3.0)
    with pytest.raises(ValueError):
        qt.perform_linear_regression([[1.0, 2.0], [2.0, 3.0]])


def test_vst2d_roi_statistics():
    from sred_amd.commands import _roi_mean_std
    assert _roi_mean_std([]) == (0.0, 0.0)
    m, s = _roi_mean_std([2, 4, 4, 4, 5, 5, 7, 9])  # the textbook population std 2
    assert (m, s) == (5.0, 2.0)
    rng = np.random.default_rng(3)
    v = rng.integers(0, 4000, 500)
    m, s = _roi_mean_std(v)
    assert m == pytest.approx(v.mean(), rel=1e-15) and s == pytest.approx(v.std(), rel=1e-9)


def test_vst2d_roi_clipped_to_the_image():
    """ImageJ keeps the intersection of a ROI with the image (an empty one
    becomes (0, 0, 0, 0)); numpy slicing alone would wrap negative starts."""
    from sred_amd.commands import _clip_roi
    assert _clip_roi((4, 6, 20, 10), 64, 48) == (4, 6, 20, 10)
    assert _clip_roi((-5, -3, 20, 10), 64, 48) == (0, 0, 15, 7)
    assert _clip_roi((50, 40, 30, 30), 64, 48) == (50, 40, 14, 8)
    assert _clip_roi((-10, 5, 5, 5), 64, 48) == (0, 0, 0, 0)
    assert _clip_roi((70, 5, 5, 5), 64, 48) == (0, 0, 0, 0)


def test_tree_plugin_argument_errors():
    """QuadTree_ takes one 2D image; OcTree_ refuses fewer than 3 slices
    (OcTree_.java:64-67); both before any device work."""
    with pytest.raises(ValueError):
        sred_amd.quadtree_stabilise_2d(np.zeros((3, 8, 8), np.float32))
    with pytest.raises(ValueError):
        sred_amd.octree_stabilise_3d(np.zeros((2, 8, 8), np.float32))
    with pytest.raises(ValueError):
        sred_amd.octree_stabilise_3d(np.zeros((8, 8), np.float32))
