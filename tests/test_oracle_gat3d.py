"""Pins the oracle's GATMinimizer3D objective and fit (GATMinimizer3D.java:
40-157, run by the VarianceStabilisingTransform3D_ plugin); no GPU needed.

* the block grid: sides max(n/6, 1) with the halving rule, against a direct
  Python statement of the reference's loop on many shapes; the shapes for
  which that loop never ends (a 1-voxel axis while another side stays above
  1) are refused;
* the objective against an independent FP64 numpy restatement in the
  reference's sequential order (transform with the v <= -ref/gain zero
  branch, per-block double sums of v and v*v, mean of (var - 1)^2) to
  1e-12 relative; NaN for gain <= 0, sigma < 0, offset < 0;
* on a Poisson-Gaussian volume (gain 2, read noise 3, offset 10) the fit
  lowers the objective from its start and does not increase it over the
  iterations (the simplex keeps its best vertex).
"""
import numpy as np
import pytest

import oracle as O


def _blocks_py(w, h, d):
    a, b, c = max(w // 6, 1), max(h // 6, 1), max(d // 6, 1)
    nx, ny, nz = w // a, h // b, d // c
    steps = 0
    while nx < 2 or ny < 2 or nz < 2:
        if nx < 2 and a > 1:
            a = max(a // 2, 1)
        if ny < 2 and b > 1:
            b = max(b // 2, 1)
        if nz < 2 and c > 1:
            c = max(c // 2, 1)
        nx, ny, nz = w // a, h // b, d // c
        if a == 1 and b == 1 and c == 1:
            break
        steps += 1
        if steps > 64:
            return None  # the reference's loop does not end
    return a, b, c, nx, ny, nz


def test_block_grid():
    for shape in [(128, 128, 64), (36, 40, 16), (5, 7, 3), (1, 1, 1), (12, 12, 2), (7, 100, 9), (2, 2, 2)]:
        assert O.gat_blocks3d(*shape) == _blocks_py(*shape)
    for shape in [(1, 20, 20), (13, 2, 1), (40, 1, 40)]:
        assert _blocks_py(*shape) is None
        with pytest.raises(ValueError):
            O.gat_blocks3d(*shape)


def _objective_numpy(vol, gain, sigma, offset):
    d, h, w = vol.shape
    bw, bh, bd, nbx, nby, nbz = _blocks_py(w, h, d)
    ref = (3.0 / 8.0) * gain * gain + sigma * sigma - gain * offset
    v = vol.astype(np.float64)
    with np.errstate(invalid="ignore"):
        t = np.where(v <= -ref / gain, 0.0, (2.0 / gain) * np.sqrt(gain * v + ref)).astype(np.float32)
    t = t.astype(np.float64)
    nb = nbx * nby * nbz
    err = 0.0
    for bz in range(nbz):
        for by in range(nby):
            for bx in range(nbx):
                blk = t[bz * bd:(bz + 1) * bd, by * bh:(by + 1) * bh, bx * bw:(bx + 1) * bw]
                mean = blk.sum() / blk.size
                var = (blk * blk).sum() / blk.size - mean * mean
                err += (var - 1.0) ** 2 / nb
    return err


def _vol(d, h, w, seed, gain=2.0, sigma=3.0, offset=10.0):
    rng = np.random.default_rng(seed)
    lam = 20.0 + 30.0 * np.sin(np.arange(w) / 5.0)[None, None, :] ** 2 + np.zeros((d, h, w))
    return (gain * rng.poisson(lam) + rng.normal(offset, sigma, (d, h, w))).astype(np.float32)


@pytest.mark.parametrize("shape", [(16, 40, 36), (7, 13, 64), (3, 6, 6), (30, 5, 8)])
def test_objective_against_numpy(shape):
    vol = _vol(*shape, seed=sum(shape))
    for p in [(2.0, 3.0, 10.0), (0.7, 40.0, 0.0), (5.0, 0.0, 250.0)]:
        got = O.gat_objective3d(vol, *p)
        assert got == pytest.approx(_objective_numpy(vol, *p), rel=1e-12)
    for bad in [(0.0, 3.0, 10.0), (-1.0, 3.0, 10.0), (2.0, -0.1, 10.0), (2.0, 3.0, -5.0)]:
        assert np.isnan(O.gat_objective3d(vol, *bad))


def test_fit_lowers_the_objective():
    vol = _vol(12, 30, 30, 3)
    start = O.gat_objective3d(vol, 1.0, 100.0, 10.0)
    errs = []
    for it in (0, 20, 100, 400):
        p, err = O.gat_minimize3d(vol, 1.0, 100.0, 10.0, it)
        assert err == pytest.approx(O.gat_objective3d(vol, *p), rel=1e-15)
        errs.append(err)
    assert errs[0] <= start and all(a >= b for a, b in zip(errs, errs[1:])) and errs[-1] < 0.5 * start
