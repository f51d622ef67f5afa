"""FP64 autograd restatement of the global Pearson map -- TEST INFRASTRUCTURE.

The checker for the hand-written adjoint: the same forward map as
``oracle.global_map(PEARSON, ..., fp64=True)`` written with torch float64
tensors on the CPU, so ``torch.autograd`` differentiates it mechanically
(independently of the adjoint's hand derivation).  It follows
kernelGetGlobalPearson2D.cl:67-166 and the finalize of CLUtils.java:1747-1770
(relevance constant 0, normalisation on).

Held fixed (stop-gradient), as the adjoint's contract states: the noise
variance h (filter parameter), nPixels, the relevance mask and the
normalisation min/max; blocks that are flat in float32 (std == 0) stay flat
(no gradient flows through their statistics).  Only small crops are
practical (N x N pair matrices).
"""
import numpy as np
import torch

import oracle as O


def _block(b):
    """(bw, bh) of a block given as one odd side or a (width, height) pair."""
    return (b, b) if np.isscalar(b) else (int(b[0]), int(b[1]))


def pearson_map_t(x, b, h_d=None, stds32=None, n_pixels=None, minmax=None):
    """Normalised global Pearson map of x (H, W) float64 -> (H, W) float64,
    with b x b blocks or, for b = (bw, bh), bw wide and bh high.

    Returns (map, aux) where aux holds the fixed quantities used.
    """
    H, W = x.shape
    bw, bh = _block(b)
    rw, rh = bw // 2, bh // 2
    k, dx, dy = O.footprint(bw, bh)
    x32 = x.detach().numpy().astype(np.float32)
    if stds32 is None:
        _, stds32 = O.local_stats_f32(x32, bw, bh)
    if h_d is None:
        h_d = O.java_float_literal(O.noise_var(x32))
    ys, xs = np.mgrid[rh:H - rh, rw:W - rw]
    ys, xs = ys.ravel(), xs.ravel()
    idx = torch.from_numpy((ys[:, None] + dy[None, :]) * W + (xs[:, None] + dx[None, :]))
    A = x.reshape(-1)[idx]                                   # N x K
    Ac = A - A.mean(1, keepdim=True)
    flat = torch.from_numpy(stds32[ys, xs] == 0)
    var = (Ac * Ac).sum(1) / (k - 1)
    sd = torch.where(flat, torch.zeros_like(var), torch.sqrt(torch.where(flat, torch.ones_like(var), var)))
    cov = Ac @ Ac.T / (k - 1)
    rr = cov / (sd[:, None] * sd[None, :] + 1e-7)
    w = torch.exp(-((sd[:, None] - sd[None, :]) ** 2) / (h_d + 1e-7))
    T = torch.clamp(rr, min=0.0) * w
    both = flat[:, None] & flat[None, :]
    one = flat[:, None] ^ flat[None, :]
    T = torch.where(both, w, torch.where(one, torch.zeros_like(T), T))
    S = T.sum(1)
    Wsum = w.sum(1)
    if n_pixels is None:
        n_pixels = float(np.count_nonzero(~flat.numpy()))
    rep = S / (Wsum * n_pixels + 1e-7)
    mask = ~flat
    if minmax is None:
        sel = rep.detach()[mask]
        minmax = (float(sel.min()), float(sel.max())) if sel.numel() else (np.finfo(np.float32).max,
                                                                          -np.finfo(np.float32).max)
    mn, mx = minmax
    val = torch.where(mask, (rep - mn) / (mx - mn + 1e-7), rep)
    out = torch.zeros(H * W, dtype=x.dtype)
    out = out.index_put((torch.from_numpy(ys * W + xs),), val)
    return out.reshape(H, W), dict(h_d=h_d, stds32=stds32, n_pixels=n_pixels, minmax=minmax)


def pearson_map_grad(img, b, weights):
    """(map, dL/dimg) of L = sum(weights * map) in FP64 autograd."""
    x = torch.tensor(np.asarray(img, np.float64), requires_grad=True)
    m, aux = pearson_map_t(x, b)
    L = (m * torch.from_numpy(np.asarray(weights, np.float64))).sum()
    L.backward()
    return m.detach().numpy(), x.grad.numpy(), aux


def loss_fixed(img, b, weights, aux):
    """L at img with the fixed quantities of a previous evaluation (for finite differences)."""
    x = torch.tensor(np.asarray(img, np.float64))
    m, _ = pearson_map_t(x, b, aux["h_d"], aux["stds32"], aux["n_pixels"], aux["minmax"])
    return float((m * torch.from_numpy(np.asarray(weights, np.float64))).sum())
