"""Pins the FP64 autograd restatement used to check the adjoint: its forward
equals oracle64's map, and its gradient equals FP64 central differences."""
import numpy as np
import pytest

import oracle as O
from sred_amd import synthetic

torch = pytest.importorskip("torch")
from oracle import torch_ref as T  # noqa: E402


@pytest.mark.parametrize("shape,b,seed", [((26, 28), 5, 4), ((30, 27), 7, 9), ((34, 29), (5, 9), 11),
                                         ((29, 36), (11, 3), 12)])
def test_torch_forward_matches_oracle64(shape, b, seed):
    img = O.normalize_input(synthetic.spots(shape[1], shape[0], 8, seed=seed))
    m, _ = T.pearson_map_t(torch.tensor(img.astype(np.float64)), b)
    bw, bh = (b, b) if np.isscalar(b) else b
    ref = O.global_map(O.PEARSON, img, bw, bh, 0.0, True, fp64=True)
    assert np.abs(m.numpy() - ref).max() < 1e-12


@pytest.mark.parametrize("b", [5, (7, 3)])
def test_torch_gradient_matches_central_differences(b):
    img = O.normalize_input(synthetic.spots(28, 26, 8, seed=4))
    u = np.random.default_rng(40).random(img.shape)
    _, g, aux = T.pearson_map_grad(img, b, u)
    h = 1e-6   # FP64: below the spacing of the max(0, r) kinks, above rounding
    for y, x in [(3, 3), (7, 13), (11, 8), (15, 18), (22, 24), (12, 2), (0, 5)]:
        p = img.astype(np.float64).copy(); p[y, x] += h
        q = img.astype(np.float64).copy(); q[y, x] -= h
        fd = (T.loss_fixed(p, b, u, aux) - T.loss_fixed(q, b, u, aux)) / (2 * h)
        assert abs(fd - g[y, x]) <= 1e-6 * np.abs(g).max() + 1e-9
