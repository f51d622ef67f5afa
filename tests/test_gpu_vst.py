"""GPU variance stabilisation against the oracle: getGAT bit-exact, the
GATMinimizer2D objective bit-exact (same summation order), and the simplex
fit step for step (identical parameters), then the stabilised input image
of Utils.getInputImage2D."""
import ctypes

import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from sred_amd import utils

pytestmark = pytest.mark.gpu


def poisson_gauss(h, w, seed, gain=2.0, mean=50.0, sigma=3.0, offset=10.0):
    rng = np.random.default_rng(seed)
    lam = mean * (1.0 + 0.5 * np.sin(np.arange(w)[None, :] / 17.0) * np.cos(np.arange(h)[:, None] / 23.0))
    return (gain * rng.poisson(lam) + rng.normal(offset, sigma, (h, w))).astype(np.float32)


@pytest.mark.parametrize("params", [(1.0, 10.0, 100.0), (2.5, 3.0, 40.0), (0.7, 0.0, -5.0)])
def test_gat_bitexact(ctx, params):
    img = poisson_gauss(123, 97, 1)
    img[:3] = -500.0  # special case rows
    got = utils.get_gat_2d(img, *params, resources=ctx)
    np.testing.assert_array_equal(got, O.gat(img, *params))


@pytest.mark.parametrize("shape", [(6, 6), (13, 29), (100, 77), (256, 256), (1024, 1024)])
def test_gat_objective_bitexact(ctx, shape):
    img = poisson_gauss(*shape, seed=2)
    m = utils.GATMinimizer2D(img, shape[1], shape[0], 1, 10, 100, 10, resources=ctx)
    for p in [(1.0, 10.0, 100.0), (2.0, 3.0, 10.0), (0.5, 20.0, -30.0)]:
        assert m.user_function(p) == O.gat_objective(img, *p)


@pytest.mark.parametrize("shape,iters", [((48, 60), 120), ((128, 96), 250), ((256, 256), 300)])
def test_gat_minimize_same_path_as_oracle(ctx, shape, iters):
    img = poisson_gauss(*shape, seed=3)
    m = utils.GATMinimizer2D(img, shape[1], shape[0], 1, 10, 100, iters, resources=ctx)
    m.run()
    p, f = O.gat_minimize(img, 1.0, 10.0, 100.0, iters)
    assert (m.gain, m.sigma, m.offset) == tuple(p)
    assert m.error == f


def test_stabilised_input_image(ctx):
    img = poisson_gauss(96, 120, 4)
    got = utils.get_input_image_2d(img, stabilise_noise_variance=True, max_iter=200, resources=ctx).image_array
    p, _ = O.gat_minimize(img, 1.0, 10.0, 100.0, 200)
    want = O.normalize_input(O.gat(img, *(float(np.float32(v)) for v in p)))
    np.testing.assert_array_equal(got, want)


def test_gat_minimize_device_path_and_errors(ctx):
    torch = pytest.importorskip("torch")
    img = poisson_gauss(64, 64, 5)
    d = torch.from_numpy(img).cuda()
    params = (ctypes.c_double * 3)()
    err = ctypes.c_double()
    L.check(L.load().sred_gat_minimize_2d_dev(ctx.handle, d.data_ptr(), 64, 64, 1.0, 10.0, 100.0, 80, params,
                                              ctypes.byref(err)))
    p, f = O.gat_minimize(img, 1.0, 10.0, 100.0, 80)
    assert tuple(params) == tuple(p) and err.value == f
    # measurement: the row kernel's time summed over the fit's launches (with
    # the timing option on), and its pixel evaluations (4 sets per candidate
    # launch, 3 per shrink); the fit itself is unchanged by the option
    ms, launches, evals = ctx.last_contraction()
    assert ms == 0.0 and 1 <= launches <= 81 + 80
    assert 3 * 64 * 64 * launches <= evals <= 4 * 64 * 64 * launches
    ctx.set_option(L.OPT_FIT_TIMING, 1)
    try:
        L.check(L.load().sred_gat_minimize_2d_dev(ctx.handle, d.data_ptr(), 64, 64, 1.0, 10.0, 100.0, 80, params,
                                                  ctypes.byref(err)))
    finally:
        ctx.set_option(L.OPT_FIT_TIMING, 0)
    assert tuple(params) == tuple(p) and err.value == f
    ms2, launches2, evals2 = ctx.last_contraction()
    assert ms2 > 0.0 and (launches2, evals2) == (launches, evals)
    with pytest.raises(sred_amd.SredError):
        utils.GATMinimizer2D(np.zeros((5, 40), np.float32), 40, 5, 1, 10, 100, 10, resources=ctx).run()


def test_vst2d_plugin(ctx):
    rng = np.random.default_rng(6)
    img = (2.0 * rng.poisson(np.linspace(5, 80, 96 * 80).reshape(80, 96)) + rng.normal(10, 3, (80, 96))).astype(
        np.uint16)
    out, (g, s, o) = sred_amd.variance_stabilising_transform_2d(img, max_iter=200, roi=(4, 6, 20, 10),
                                                                resources=ctx)
    from sred_amd.commands import _roi_mean_std  # the plugin's sequential double sums
    mean, std = _roi_mean_std(img[6:16, 4:24])
    assert mean == pytest.approx(img[6:16, 4:24].mean(), rel=1e-12)
    p, f = O.gat_minimize(img.astype(np.float32), 1.0, std, mean, 200)
    assert (g, s, o) == tuple(p)
    ref = (3.0 / 8.0) * g * g + s * s - g * o
    x = img.astype(np.float64)
    with np.errstate(invalid="ignore"):
        want = np.where(g * x + ref < 0.0, 0.0, (2.0 / g) * np.sqrt(g * x + ref)).astype(np.float32)
    np.testing.assert_array_equal(out, want)
    # a ROI partly outside the image: its clipped rectangle's statistics
    out2, (g2, s2, o2) = sred_amd.variance_stabilising_transform_2d(img, max_iter=200, roi=(-5, 70, 20, 20),
                                                                   resources=ctx)
    mean2, std2 = _roi_mean_std(img[70:80, 0:15])
    p2, _ = O.gat_minimize(img.astype(np.float32), 1.0, std2, mean2, 200)
    assert (g2, s2, o2) == tuple(p2)
