"""Seeded synthetic microscopy-like images for the benchmark workloads.

Every generator uses numpy.random.default_rng(seed) and returns a float32
array (height, width).  Shapes and contents:

* spots(256, 256, 200, seed=0)   Gaussian spots (sigma 2 px, amplitude U[0.5,1])
                                 on background 0.1 plus N(0, 0.05^2) noise;
* ring_template(15)              15x15 ring, radius 5 px, Gaussian profile width 1.5 px;
* rings_and_blobs(2048, ...)     3000 rings + 3000 Gaussian blobs (sigma 2.5),
                                 background 0.1, Poisson noise at 200 photons peak;
* lattice_and_blobs(1024, ...)   left half: ring lattice (pitch 24 px, rotated
                                 U[0,30] degrees); right half: ~900 blobs with
                                 sigma U[1,4]; background 0.1 plus N(0, 0.05^2);
* spot_ring_mixture(w, h, seed)  spots at the density of spots(256, 256, 200)
                                 and rings at that of rings_and_blobs(2048, 3000),
                                 background 0.1 plus N(0, 0.05^2) (any w x h).
"""
import numpy as np


def _splat(img, cx, cy, radius, profile):
    """Add profile(dx, dy) on the window of +-radius around (cx, cy)."""
    h, w = img.shape
    x0, x1 = max(int(np.floor(cx - radius)), 0), min(int(np.ceil(cx + radius)) + 1, w)
    y0, y1 = max(int(np.floor(cy - radius)), 0), min(int(np.ceil(cy + radius)) + 1, h)
    if x0 >= x1 or y0 >= y1:
        return
    xs = np.arange(x0, x1, dtype=np.float64) - cx
    ys = np.arange(y0, y1, dtype=np.float64) - cy
    img[y0:y1, x0:x1] += profile(xs[None, :], ys[:, None])


def _gauss(sigma, amp):
    return lambda dx, dy: amp * np.exp(-(dx * dx + dy * dy) / (2.0 * sigma * sigma))


def _ring(radius, width, amp):
    return lambda dx, dy: amp * np.exp(-((np.sqrt(dx * dx + dy * dy) - radius) ** 2) / (2.0 * width * width))


def spots(width=256, height=256, n_spots=200, seed=0, sigma=2.0, background=0.1, noise=0.05):
    rng = np.random.default_rng(seed)
    img = np.full((height, width), background, np.float64)
    xs = rng.uniform(0, width, n_spots)
    ys = rng.uniform(0, height, n_spots)
    amps = rng.uniform(0.5, 1.0, n_spots)
    for x, y, a in zip(xs, ys, amps):
        _splat(img, x, y, 4 * sigma, _gauss(sigma, a))
    img += rng.normal(0.0, noise, img.shape)
    return img.astype(np.float32)


def ring_template(size=15, radius=5.0, width=1.5):
    c = (size - 1) / 2.0
    img = np.zeros((size, size), np.float64)
    _splat(img, c, c, size, _ring(radius, width, 1.0))
    return img.astype(np.float32)


def rings_and_blobs(size=2048, n_rings=3000, n_blobs=3000, seed=1, background=0.1, photons=200.0):
    rng = np.random.default_rng(seed)
    img = np.full((size, size), background, np.float64)
    for _ in range(n_rings):
        x, y = rng.uniform(0, size, 2)
        rng.uniform(0, 2 * np.pi)  # rotation angle: a ring is rotation invariant up to sampling
        _splat(img, x, y, 10.0, _ring(5.0, 1.5, 1.0))
    for _ in range(n_blobs):
        x, y = rng.uniform(0, size, 2)
        _splat(img, x, y, 10.0, _gauss(2.5, rng.uniform(0.5, 1.0)))
    lam = photons * img
    return rng.poisson(lam).astype(np.float32)


def lattice_and_blobs(size=1024, pitch=24.0, n_blobs=900, seed=2, background=0.1, noise=0.05):
    rng = np.random.default_rng(seed)
    img = np.full((size, size), background, np.float64)
    theta = np.deg2rad(rng.uniform(0.0, 30.0))
    ct, st = np.cos(theta), np.sin(theta)
    half = size // 2
    n = int(2 * size / pitch) + 2
    for i in range(-n, n + 1):
        for j in range(-n, n + 1):
            u, v = i * pitch, j * pitch
            x = half / 2 + ct * u - st * v
            y = size / 2 + st * u + ct * v
            if -10 <= x < half + 10 and -10 <= y < size + 10:
                _splat(img[:, :half], x, y, 10.0, _ring(5.0, 1.5, 1.0))
    right = img[:, half:]
    for _ in range(n_blobs):
        x, y = rng.uniform(0, size - half), rng.uniform(0, size)
        s = rng.uniform(1.0, 4.0)
        _splat(right, x, y, 4 * s, _gauss(s, rng.uniform(0.5, 1.0)))
    img += rng.normal(0.0, noise, img.shape)
    return img.astype(np.float32)


def spot_ring_mixture(width, height, seed=3, background=0.1, noise=0.05):
    rng = np.random.default_rng(seed)
    img = np.full((height, width), background, np.float64)
    n_spots = int(round(200.0 * width * height / (256 * 256)))
    n_rings = int(round(3000.0 * width * height / (2048 * 2048)))
    for _ in range(n_spots):
        x, y = rng.uniform(0, width), rng.uniform(0, height)
        _splat(img, x, y, 8.0, _gauss(2.0, rng.uniform(0.5, 1.0)))
    for _ in range(n_rings):
        x, y = rng.uniform(0, width), rng.uniform(0, height)
        _splat(img, x, y, 10.0, _ring(5.0, 1.5, rng.uniform(0.5, 1.0)))
    img += rng.normal(0.0, noise, img.shape)
    return img.astype(np.float32)
