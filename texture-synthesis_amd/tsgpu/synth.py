"""Seeded synthetic inputs for the benchmark and parity workloads.

All images are RGBA8 numpy arrays of shape (H, W, 4), row-major, alpha 255,
the layout the reference works in after ``to_rgba8``
(lib/src/utils.rs:62-76).

* ``example(seed, w, h)``: per RGB channel a sum of 8 random-phase 2-D
  sinusoids with periods log-uniform in [4, 64] px, plus 8-px-lattice value
  noise (smoothstep interpolation) at 0.25 of the sinusoid range, min-max
  normalised and quantised with round(255 v).
* ``blobs(seed, w, h)``: 8-16 random discs of 255 on 0, Gaussian-blurred with
  sigma 12 px, grey replicated to RGB.
* ``ellipse_mask(seed, w, h)``: 1-3 random ellipses of 0 (holes) on 255,
  redrawn until the hole fraction lies in [0.20, 0.40].
"""
from __future__ import annotations

import numpy as np


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(int(seed) & ((1 << 64) - 1)))


def _smoothstep(t: np.ndarray) -> np.ndarray:
    return t * t * (3.0 - 2.0 * t)


def example(seed: int, w: int, h: int) -> np.ndarray:
    rng = _rng(seed)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    out = np.empty((h, w, 4), dtype=np.uint8)
    out[..., 3] = 255
    for c in range(3):
        periods = np.exp(rng.uniform(np.log(4.0), np.log(64.0), size=8))
        thetas = rng.uniform(0.0, np.pi, size=8)
        phis = rng.uniform(0.0, 2.0 * np.pi, size=8)
        v = np.zeros((h, w), dtype=np.float64)
        for p, t, f in zip(periods, thetas, phis):
            v += np.sin(2.0 * np.pi * (xs * np.cos(t) + ys * np.sin(t)) / p + f)
        lat = rng.uniform(0.0, 1.0, size=(h // 8 + 2, w // 8 + 2))
        gx, gy = xs / 8.0, ys / 8.0
        ix, iy = np.floor(gx).astype(np.int64), np.floor(gy).astype(np.int64)
        fx, fy = _smoothstep(gx - ix), _smoothstep(gy - iy)
        n00, n10 = lat[iy, ix], lat[iy, ix + 1]
        n01, n11 = lat[iy + 1, ix], lat[iy + 1, ix + 1]
        noise = (n00 * (1 - fx) + n10 * fx) * (1 - fy) + (n01 * (1 - fx) + n11 * fx) * fy
        rng_v = float(v.max() - v.min()) or 1.0
        v = v + 0.25 * rng_v * noise
        v = (v - v.min()) / ((v.max() - v.min()) or 1.0)
        out[..., c] = np.round(255.0 * v).astype(np.uint8)
    return out


def _gauss_blur(img: np.ndarray, sigma: float) -> np.ndarray:
    r = int(np.ceil(3.0 * sigma))
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-(x * x) / (2.0 * sigma * sigma))
    k /= k.sum()
    pad = np.pad(img, r, mode="reflect")
    tmp = np.zeros((pad.shape[0], img.shape[1]), dtype=np.float64)
    for i, kv in enumerate(k):
        tmp += kv * pad[:, i:i + img.shape[1]]
    out = np.zeros(img.shape, dtype=np.float64)
    for i, kv in enumerate(k):
        out += kv * tmp[i:i + img.shape[0], :]
    return out


def blobs(seed: int, w: int, h: int) -> np.ndarray:
    rng = _rng(seed)
    n = int(rng.integers(8, 17))
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    g = np.zeros((h, w), dtype=np.float64)
    for _ in range(n):
        cx, cy = rng.uniform(0, w), rng.uniform(0, h)
        r = rng.uniform(0.04, 0.16) * min(w, h)
        g[(xs - cx) ** 2 + (ys - cy) ** 2 <= r * r] = 255.0
    g = _gauss_blur(g, 12.0)
    v = np.clip(np.round(g), 0, 255).astype(np.uint8)
    out = np.empty((h, w, 4), dtype=np.uint8)
    out[..., 0] = v
    out[..., 1] = v
    out[..., 2] = v
    out[..., 3] = 255
    return out


def ellipse_mask(seed: int, w: int, h: int) -> np.ndarray:
    rng = _rng(seed)
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    while True:
        hole = np.zeros((h, w), dtype=bool)
        for _ in range(int(rng.integers(1, 4))):
            cx, cy = rng.uniform(0.2 * w, 0.8 * w), rng.uniform(0.2 * h, 0.8 * h)
            a, b = rng.uniform(0.1, 0.4) * w, rng.uniform(0.1, 0.4) * h
            th = rng.uniform(0, np.pi)
            dx, dy = xs - cx, ys - cy
            u = dx * np.cos(th) + dy * np.sin(th)
            v = -dx * np.sin(th) + dy * np.cos(th)
            hole |= (u / a) ** 2 + (v / b) ** 2 <= 1.0
        frac = hole.mean()
        if 0.20 <= frac <= 0.40:
            break
    out = np.full((h, w, 4), 255, dtype=np.uint8)
    out[hole, 0:3] = 0
    return out
