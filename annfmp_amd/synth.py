"""Seeded synthetic image pairs for the NNF workloads.

Every choice the workload descriptions leave open is fixed here once:

* noise image: U{0..255} per channel, float, Gaussian blur per channel with
  ``scipy.ndimage.gaussian_filter`` (``mode='wrap'``), then rectangles;
* rectangles: width U[W/32, W/4], height U[H/32, H/4] (integers, at least 2
  px), top-left uniform so the rectangle lies inside the image, colour
  U{0..255}^3, painted in order over the blurred noise;
* shift pair (256x256 case): ``B = roll(A, (+17, -9), axis=(0, 1))``, i.e.
  17 px down and 9 px left, wrapping;
* warp pair (1080p / 4K cases): the image is cut into a ``gx x gy`` grid of
  blocks (block edges at ``round(i*W/gx)``); every block gets an integer
  offset ``(dy, dx)`` with each component U[-m, m]; ``B[y, x] = A[(y+dy) mod
  H, (x+dx) mod W]`` inside the block;
* sensor noise: ``N(0, 4)`` is a standard deviation of 4 grey levels, added
  to B in float, rounded half-to-even, clipped to [0, 255], cast to uint8.

All randomness comes from ``numpy.random.default_rng(seed)``, drawn in a
fixed order, so a (config, seed) pair names one image pair everywhere.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass(frozen=True)
class PairSpec:
    height: int
    width: int
    n_rects: int
    sigma: float
    mode: str = "warp"              # "shift" or "warp"
    shift: tuple = (17, -9)         # (dy, dx) for mode "shift"
    grid: tuple = (16, 9)           # (gx, gy) blocks for mode "warp"
    max_offset: int = 32
    noise_std: float = 4.0


# the workload shapes used by tests and bench.py
SPECS = {
    "256x256-shift": PairSpec(256, 256, 20, 2.0, mode="shift"),
    "1080p-warp": PairSpec(1080, 1920, 200, 3.0, grid=(16, 9), max_offset=32),
    "512x512-warp": PairSpec(512, 512, 200, 3.0, grid=(4, 4), max_offset=32),
    "4k-warp": PairSpec(2160, 3840, 200, 3.0, grid=(32, 18), max_offset=64),
}


def _blurred_noise(rng, H, W, sigma):
    from scipy.ndimage import gaussian_filter
    noise = rng.integers(0, 256, size=(H, W, 3)).astype(np.float32)
    out = np.empty_like(noise)
    for c in range(3):
        out[..., c] = gaussian_filter(noise[..., c], sigma=sigma, mode="wrap")
    return out


def make_image_a(spec: PairSpec, rng) -> np.ndarray:
    H, W = spec.height, spec.width
    img = _blurred_noise(rng, H, W, spec.sigma)
    wmin, wmax = max(2, W // 32), max(3, W // 4)
    hmin, hmax = max(2, H // 32), max(3, H // 4)
    for _ in range(spec.n_rects):
        w = int(rng.integers(wmin, wmax + 1))
        h = int(rng.integers(hmin, hmax + 1))
        x0 = int(rng.integers(0, W - w + 1))
        y0 = int(rng.integers(0, H - h + 1))
        col = rng.integers(0, 256, size=3).astype(np.float32)
        img[y0:y0 + h, x0:x0 + w, :] = col
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def _warp(a, spec: PairSpec, rng):
    H, W = spec.height, spec.width
    gx, gy = spec.grid
    xs = [round(i * W / gx) for i in range(gx + 1)]
    ys = [round(i * H / gy) for i in range(gy + 1)]
    offs = rng.integers(-spec.max_offset, spec.max_offset + 1, size=(gy, gx, 2))
    out = np.empty_like(a)
    for by in range(gy):
        for bx in range(gx):
            dy, dx = int(offs[by, bx, 0]), int(offs[by, bx, 1])
            yy = (np.arange(ys[by], ys[by + 1]) + dy) % H
            xx = (np.arange(xs[bx], xs[bx + 1]) + dx) % W
            out[ys[by]:ys[by + 1], xs[bx]:xs[bx + 1]] = a[yy[:, None], xx[None, :]]
    return out, offs


def make_pair(spec: PairSpec | str, seed: int = 0):
    """Returns (A, B) uint8 (H, W, 3) arrays."""
    if isinstance(spec, str):
        spec = SPECS[spec]
    rng = np.random.default_rng(seed)
    a = make_image_a(spec, rng)
    if spec.mode == "shift":
        b = np.roll(a, spec.shift, axis=(0, 1))
    else:
        b, _ = _warp(a, spec, rng)
    noisy = b.astype(np.float32) + rng.normal(0.0, spec.noise_std, size=b.shape).astype(np.float32)
    b = np.clip(np.rint(noisy), 0, 255).astype(np.uint8)
    return a, b


def small_spec(height, width, n_rects=6, sigma=1.5, mode="warp", grid=(4, 3), max_offset=6):
    return PairSpec(height, width, n_rects, sigma, mode=mode, grid=grid, max_offset=max_offset)
