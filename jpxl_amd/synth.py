"""Seeded synthetic test images (the workload generator of the benchmark).

Per channel: a Gaussian field with a 1/f^2 power spectrum (FFT-shaped white
noise), channels correlated at rho = 0.8 (z_c = 0.8 z_0 + 0.6 e_c), normalised
to [0, 1] as linear light and sRGB-encoded; then 40 random axis-aligned
rectangles and disks with hard edges and uniform random colours; then
uniform noise of amplitude 0.1 over 10% of the area; quantised to 8 or 16
bits.  Deterministic for a given (seed, size, bits).
"""
from __future__ import annotations

import numpy as np


def _field(rng: np.random.Generator, h: int, w: int) -> np.ndarray:
    noise = rng.standard_normal((h, w))
    fy = np.fft.fftfreq(h)[:, None]
    fx = np.fft.rfftfreq(w)[None, :]
    f = np.sqrt(fx * fx + fy * fy)
    f[0, 0] = 1.0
    spec = np.fft.rfft2(noise) / f  # power ~ 1/f^2
    spec[0, 0] = 0.0
    z = np.fft.irfft2(spec, s=(h, w))
    return z / (z.std() + 1e-12)


def _srgb_encode(v: np.ndarray) -> np.ndarray:
    return np.where(v <= 0.0031308, 12.92 * v, 1.055 * np.power(np.maximum(v, 0.0), 1.0 / 2.4) - 0.055)


def synthetic_rgb(width: int, height: int, seed: int, bits: int = 8) -> np.ndarray:
    """(height, width, 3) uint8 (bits = 8) or uint16 (bits > 8) sRGB image."""
    rng = np.random.default_rng(seed)
    z0 = _field(rng, height, width)
    img = np.empty((height, width, 3), np.float64)
    for c in range(3):
        zc = 0.8 * z0 + 0.6 * _field(rng, height, width)
        lo, hi = zc.min(), zc.max()
        img[..., c] = _srgb_encode((zc - lo) / max(hi - lo, 1e-12))
    for _ in range(40):
        colour = rng.uniform(0.0, 1.0, 3)
        if rng.uniform() < 0.5:
            x0, x1 = np.sort(rng.integers(0, width, 2))
            y0, y1 = np.sort(rng.integers(0, height, 2))
            img[y0 : y1 + 1, x0 : x1 + 1] = colour
        else:
            cx, cy = rng.uniform(0, width), rng.uniform(0, height)
            r = rng.uniform(0.02, 0.15) * min(width, height)
            # the disk's mask, evaluated over its bounding box only
            bx0, bx1 = max(0, int(np.floor(cx - r)) - 1), min(width, int(np.ceil(cx + r)) + 2)
            by0, by1 = max(0, int(np.floor(cy - r)) - 1), min(height, int(np.ceil(cy + r)) + 2)
            yy, xx = np.mgrid[by0:by1, bx0:bx1]
            m = (xx - cx) ** 2 + (yy - cy) ** 2 <= r * r
            img[by0:by1, bx0:bx1][m] = colour
    covered = np.zeros((height, width), bool)
    target = 0.10 * width * height
    ncov = 0
    while ncov < target:
        pw = int(rng.integers(max(1, width // 32), max(2, width // 6)))
        ph = int(rng.integers(max(1, height // 32), max(2, height // 6)))
        x0 = int(rng.integers(0, max(1, width - pw)))
        y0 = int(rng.integers(0, max(1, height - ph)))
        patch = covered[y0 : y0 + ph, x0 : x0 + pw]
        ncov += int(patch.size - np.count_nonzero(patch))
        patch[...] = True
    img[covered] += rng.uniform(-0.1, 0.1, (int(covered.sum()), 3))
    img = np.clip(img, 0.0, 1.0)
    maxv = (1 << bits) - 1
    q = np.rint(img * maxv)
    return q.astype(np.uint8 if bits == 8 else np.uint16)


def uneven_batch(count: int, seed: int, lo: int = 256, hi: int = 3072):
    """Sizes and quality targets of the batched workload (uneven images)."""
    rng = np.random.default_rng(seed)
    sizes = [(int(rng.integers(lo, hi + 1)), int(rng.integers(lo, hi + 1))) for _ in range(count)]
    targets = [float(rng.choice([70.0, 85.0, 90.0])) for _ in range(count)]
    return sizes, targets
