"""Shared test helpers: fixtures, java.util.Random, the KAT reference block."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def kats():
    with open(os.path.join(GOLDEN, "kats.json")) as f:
        return json.load(f)


class JavaRandom:
    """java.util.Random: the documented 48-bit LCG (seed ^ 0x5DEECE66D,
    seed = seed * 0x5DEECE66D + 0xB mod 2^48, nextFloat = next(24) / 2^24)."""

    MUL, ADD, MASK = 0x5DEECE66D, 0xB, (1 << 48) - 1

    def __init__(self, seed):
        self.seed = (seed ^ self.MUL) & self.MASK

    def next(self, bits):
        self.seed = (self.seed * self.MUL + self.ADD) & self.MASK
        r = self.seed >> (48 - bits)
        return r - (1 << 32) if r >= (1 << 31) else r

    def next_float(self):
        return np.float32(self.next(24) / float(1 << 24))


def random42_inputs(k):
    """UtilsTest.java:651-661 inputs from java.util.Random(42)."""
    w, h = k["width"], k["height"]
    rnd = JavaRandom(k["seed"])
    img = np.zeros(w * h, np.float32)
    stds = np.zeros(w * h, np.float32)
    for i in range(w * h):
        img[i] = np.float32(1.0) if i % 50 == 0 else np.float32(np.float32(1.0) + rnd.next_float() * np.float32(2.0))
        stds[i] = np.float32(0.0) if i % 50 == 0 else rnd.next_float()
    return img.reshape(h, w), stds.reshape(h, w)


def kat_reference_block(block, footprint_fn):
    """ReferenceBlock2D as CLUtilsTest.java:560-597 builds it: raw footprint
    pixels with their float mean and unbiased float std."""
    block = np.asarray(block, np.float32)
    bh, bw = block.shape
    k, dx, dy = footprint_fn(bw, bh)
    px = np.array([block[bh // 2 + y, bw // 2 + x] for x, y in zip(dx, dy)], np.float32)
    m = np.float32(0)
    for v in px:
        m = np.float32(m + v)
    m = np.float32(m / np.float32(k))
    s = np.float32(0)
    for v in px:
        s = np.float32(s + np.float32(v - m) * np.float32(v - m))
    s = np.float32(s / np.float32(np.float32(k) - np.float32(1)))
    s = np.float32(np.sqrt(np.float64(s)))
    return px, float(m), float(s)


def fp64_map_at(O, metric, img, bw, bh, c, pixels):
    """oracle64's normalised global map at the given pixels (flat indices)
    without computing every row: the FP64 accumulators of those rows only,
    finalised with the mask restricted to them, so the normalisation's
    min/max come from the given pixels.  Pass the map's extreme pixels among
    them and the values equal the full map's."""
    import numpy as np

    h, w = img.shape
    m32, s32 = O.local_stats_f32(img, bw, bh)
    nv = O.noise_var(img)
    mask, thr, npx = O.relevance_mask(s32, bw // 2, bh // 2, nv, c)
    rows = np.unique(np.asarray(pixels, np.int64))
    S, W = O.global_accum_f64(metric, img, s32, bw, bh, nv, thr, rows=rows)
    sub = np.zeros_like(mask)
    sub.ravel()[rows] = mask.ravel()[rows]
    out = O.global_finalize_f64(metric, S, W, sub, npx, c, True, bw, bh)
    return out.ravel()[np.asarray(pixels, np.int64)], mask


def sample_pixels(gpu_map, mask, n, seed):
    """n seeded interior pixels with mask > 0 plus the 4 smallest and 4
    largest values of gpu_map among them (flat indices)."""
    import numpy as np

    idx = np.flatnonzero(mask.ravel() > 0)
    rng = np.random.default_rng(seed)
    vals = gpu_map.ravel()[idx]
    order = np.argsort(vals, kind="stable")
    extremes = np.concatenate([idx[order[:4]], idx[order[-4:]]])
    return np.unique(np.concatenate([rng.choice(idx, min(n, idx.size), replace=False), extremes]))
