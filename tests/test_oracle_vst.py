"""Pins the oracle's variance stabilisation (no GPU needed):
VarianceStabilisingTransform2D_.getGAT and GATMinimizer2D's objective, and
the Nelder-Mead simplex the build uses in place of ImageJ's Minimizer.

* getGAT with gain 1, sigma 0, offset 0 is the textbook Anscombe transform
  2 sqrt(x + 3/8); its special case (g v + ref < 0 -> 0); on Poisson data
  with a large mean it stabilises the variance to 1;
* the objective of a constant image is (var - 1)^2 with var = fl32(A^2) - A^2
  (the float-squared mean square less the squared mean),
  and of an image whose blocks alternate two values a, b it is
  (((A - B)/2)^2 - 1)^2 with A, B the transformed values;
* an FP64 numpy restatement of the objective (blocks of different
  variance, 7 block columns or rows where w or h leaves a remainder, pixels
  in the zero branch) to 1e-12 relative (the summation orders differ);
* the simplex: a convex quadratic's analytic minimum, an immediate stop
  when all vertices tie (constant function), and the GAT fit lowering the
  objective of a Poisson-Gaussian image.
"""
import numpy as np
import pytest

import oracle as O


def test_gat_is_anscombe_for_unit_gain():
    x = np.linspace(0, 500, 1001, dtype=np.float32)
    got = O.gat(x, 1.0, 0.0, 0.0)
    want = (2.0 * np.sqrt(x.astype(np.float64) + 3.0 / 8.0)).astype(np.float32)
    np.testing.assert_array_equal(got, want)


def test_gat_special_case_and_general_parameters():
    g, s, o = 2.5, 3.0, 40.0
    ref = 3.0 / 8.0 * g * g + s * s - g * o
    x = np.array([-100.0, -ref / g - 1.0, -ref / g + 1.0, 0.0, 17.25, 1e4], np.float32)
    got = O.gat(x, g, s, o)
    v = x.astype(np.float64)
    want = np.where(g * v + ref < 0, 0.0, (2.0 / g) * np.sqrt(np.maximum(g * v + ref, 0.0))).astype(np.float32)
    np.testing.assert_array_equal(got, want)
    assert got[0] == 0.0 and got[1] == 0.0 and got[2] > 0.0


def test_anscombe_stabilises_poisson_variance():
    rng = np.random.default_rng(1)
    for lam in (30.0, 100.0, 1000.0):
        x = rng.poisson(lam, 200000).astype(np.float32)
        assert abs(np.var(O.gat(x, 1.0, 0.0, 0.0).astype(np.float64)) - 1.0) < 0.02


def test_objective_constant_image():
    # every block holds one value A: mean A, mean square fl32(A*A) (the
    # reference squares in float), so var = fl32(A^2) - A^2 (exact in double)
    for shape in [(60, 60), (37, 53), (6, 6), (13, 29)]:
        img = np.full(shape, 3.0, np.float32)
        A = float(O.gat(np.array([3.0], np.float32), 1.0, 10.0, 100.0)[0])
        var = float(np.float32(A) * np.float32(A)) - A * A
        assert O.gat_objective(img, 1.0, 10.0, 100.0) == pytest.approx((var - 1.0) ** 2, rel=1e-14)


def test_objective_two_value_blocks():
    h, w = 48, 72
    bw, bh = w // 6, h // 6
    img = np.zeros((h, w), np.float32)
    a, b = 7.0, 19.0
    img[:, 0::2] = a  # every block row alternates a, b with an even block width
    img[:, 1::2] = b
    g, s, o = 1.3, 2.0, 5.0
    A, B = (float(v) for v in O.gat(np.array([a, b], np.float32), g, s, o))
    # half the pixels A, half B: var = mean of the float squares - mean^2 = ((A - B)/2)^2 up to the float squares
    var = (float(np.float32(A) ** 2) + float(np.float32(B) ** 2)) / 2.0 - ((A + B) / 2.0) ** 2
    assert bw % 2 == 0 and var == pytest.approx(((A - B) / 2.0) ** 2, rel=1e-6)
    assert O.gat_objective(img, g, s, o) == pytest.approx((var - 1.0) ** 2, rel=1e-12)


def _objective_numpy(img, g, s, o):
    """GATMinimizer2D.userFunction (GATMinimizer2D.java:64-110) in numpy FP64:
    blocks of (w//6) x (h//6) pixels, w // (w//6) of them per row."""
    h, w = img.shape
    bw, bh = w // 6, h // 6
    nbx, nby = w // bw, h // bh
    ref = 3.0 / 8.0 * g * g + s * s - g * o
    d = img.astype(np.float64)
    v = np.where(d <= -ref / g, 0.0, (2.0 / g) * np.sqrt(np.maximum(g * d + ref, 0.0))).astype(np.float32)
    err = 0.0
    for by in range(nby):
        for bx in range(nbx):
            blk = v[by * bh:(by + 1) * bh, bx * bw:(bx + 1) * bw]
            mean = blk.astype(np.float64).sum() / blk.size
            sq = (blk * blk).astype(np.float64).sum() / blk.size
            err += ((sq - mean * mean) - 1.0) ** 2 / (nbx * nby)
    return err


@pytest.mark.parametrize("shape", [(60, 60), (37, 53), (13, 29), (100, 77), (7, 7), (90, 200)])
def test_objective_numpy_restatement(shape):
    # block-dependent brightness (every block a different variance), remainders
    # that give 7 block columns or rows, and pixels in the zero branch
    rng = np.random.default_rng(sum(shape))
    h, w = shape
    yy, xx = np.mgrid[0:h, 0:w]
    lam = 5.0 + 60.0 * (xx / w) + 30.0 * (yy / h) ** 2
    img = (1.7 * rng.poisson(lam) + rng.normal(4.0, 2.0, shape)).astype(np.float32)
    img[rng.random(shape) < 0.05] = -200.0
    for g, s, o in [(1.0, 10.0, 100.0), (1.7, 2.0, 4.0), (0.6, 5.0, -3.0)]:
        ref = 3.0 / 8.0 * g * g + s * s - g * o
        want = _objective_numpy(img, g, s, o)
        assert np.any(img <= -ref / g)
        assert O.gat_objective(img, g, s, o) == pytest.approx(want, rel=1e-12)


def test_objective_small_image_is_nan():
    assert np.isnan(O.gat_objective(np.zeros((5, 40), np.float32), 1, 10, 100))


def test_nelder_mead_quadratic_minimum():
    f = lambda p: (p[0] - 1.5) ** 2 + 2 * (p[1] + 0.5) ** 2 + 3 * (p[2] - 2) ** 2 + 0.1 * p[0] * p[1]
    x, fx = O.nelder_mead3(f, [0, 0, 0], [1, 1, 1], 2000)
    # gradient zero: 2(x-1.5) + 0.1y = 0, 4(y+0.5) + 0.1x = 0, z = 2
    A = np.array([[2.0, 0.1], [0.1, 4.0]])
    xy = np.linalg.solve(A, [3.0, -2.0])
    np.testing.assert_allclose(x, [xy[0], xy[1], 2.0], atol=1e-6)


def test_nelder_mead_stops_on_ties():
    calls = []
    x, fx = O.nelder_mead3(lambda p: calls.append(1) or 5.0, [1, 2, 3], [1, 1, 1], 100)
    assert len(calls) == 4 and fx == 5.0
    np.testing.assert_array_equal(x, [1, 2, 3])


def test_gat_fit_lowers_objective():
    rng = np.random.default_rng(2)
    img = (2.0 * rng.poisson(50.0, (96, 120)) + rng.normal(10.0, 3.0, (96, 120))).astype(np.float32)
    f0 = O.gat_objective(img, 1.0, 10.0, 100.0)
    p, f = O.gat_minimize(img, max_iter=400)
    assert f < 0.05 * f0 and f == O.gat_objective(img, *p)
