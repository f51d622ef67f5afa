"""GPU GATMinimizer3D objective and fit (GATMinimizer3D.java:40-157) and the
VarianceStabilisingTransform3D_ plugin against the oracle, through the C ABI:
the objective bit for bit (block grids with and without halving, NaN for
invalid parameters), the simplex on the same path as the restatement
(parameters and error bit for bit), the plugin's output volume, the 2D
objective unchanged by the volume generalisation, the errors."""
import ctypes

import numpy as np
import pytest

import oracle as O
import sred_amd
from sred_amd import _lib as L
from sred_amd import utils

pytestmark = pytest.mark.gpu


def _vol(d, h, w, seed, gain=2.0, sigma=3.0, offset=10.0):
    rng = np.random.default_rng(seed)
    lam = 20.0 + 30.0 * np.sin(np.arange(w) / 5.0)[None, None, :] ** 2 + np.zeros((d, h, w))
    return (gain * rng.poisson(lam) + rng.normal(offset, sigma, (d, h, w))).astype(np.float32)


@pytest.mark.parametrize("shape", [(16, 40, 36), (7, 13, 64), (3, 6, 6), (30, 5, 8), (64, 128, 128), (2, 50, 40)])
def test_objective_bitexact(ctx, shape):
    vol = _vol(*shape, seed=sum(shape))
    d, h, w = shape
    m = utils.GATMinimizer3D(vol, w, h, d, 1.0, 100.0, 10.0, 10, ctx)
    for p in [(2.0, 3.0, 10.0), (0.7, 40.0, 0.0), (5.0, 0.0, 250.0), (1.0, 100.0, 10.0)]:
        assert m.user_function(p) == O.gat_objective3d(vol, *p)
    for bad in [(0.0, 3.0, 10.0), (2.0, -0.1, 10.0), (2.0, 3.0, -5.0)]:
        assert np.isnan(m.user_function(bad))


@pytest.mark.parametrize("shape,iters", [((16, 40, 36), 300), ((12, 30, 30), 1000), ((5, 64, 48), 150)])
def test_minimize_same_path_as_oracle(ctx, shape, iters):
    vol = _vol(*shape, seed=3 + sum(shape))
    d, h, w = shape
    m = utils.GATMinimizer3D(vol, w, h, d, 1.0, 100.0, 10.0, iters, ctx)
    m.run()
    p, f = O.gat_minimize3d(vol, 1.0, 100.0, 10.0, iters)
    print(f"gat3d {shape} {iters} it: params {m.gain:.6g} {m.sigma:.6g} {m.offset:.6g} error {m.error:.6g}",
          flush=True)
    assert (m.gain, m.sigma, m.offset) == tuple(p)
    assert m.error == f


def test_plugin_and_device_entry(ctx):
    torch = pytest.importorskip("torch")
    vol = _vol(10, 36, 30, 8)
    out, (g, s, o) = sred_amd.variance_stabilising_transform_3d(vol, max_iter=400, resources=ctx)
    p, f = O.gat_minimize3d(vol, 1.0, 100.0, 10.0, 400)
    assert (g, s, o) == tuple(p)
    want = vol.copy()
    ref = (3.0 / 8.0) * g * g + s * s - g * o
    x = want.astype(np.float64)
    with np.errstate(invalid="ignore"):
        want = np.where(g * x + ref < 0.0, 0.0, (2.0 / g) * np.sqrt(g * x + ref)).astype(np.float32)
    np.testing.assert_array_equal(out, want)
    d, h, w = vol.shape
    dv = torch.from_numpy(vol).cuda()
    params = (ctypes.c_double * 3)()
    err = ctypes.c_double()
    L.check(L.load().sred_gat_minimize_3d_dev(ctx.handle, dv.data_ptr(), w, h, d, 1.0, 100.0, 10.0, 400, params,
                                              ctypes.byref(err)))
    assert tuple(params) == tuple(p) and err.value == f


def test_image_objective_unchanged(ctx):
    img = _vol(1, 64, 72, 5)[0]
    m = utils.GATMinimizer2D(img, 72, 64, 1.0, 10.0, 100.0, 10, ctx)
    for p in [(2.0, 3.0, 10.0), (0.5, 30.0, 5.0)]:
        assert m.user_function(p) == O.gat_objective(img, *p)


def test_errors(ctx):
    with pytest.raises(sred_amd.SredError):  # the reference's block search never ends
        utils.GATMinimizer3D(np.ones((20, 20, 1), np.float32), 1, 20, 20, 1, 100, 10, 5, ctx).run()
    with pytest.raises(ValueError):
        sred_amd.variance_stabilising_transform_3d(np.ones((1, 8, 8), np.float32), resources=ctx)
