"""Seeded synthetic target images for the benchmark configurations, and the
reference's image ingest (gray conversion, nearest resize).

Config names follow the workloads they describe:

* ``gauss5``   256x256 u8: 0.05 floor + 5 isotropic Gaussians, scaled to 255
* ``rects12``  512x512 u8: 0.1 background + 12 axis-aligned rectangles at 1.0
* ``noise3``   1024x1024 f32: 3-octave value noise x (1 + 0.3 checker 16x16), unit mass
* ``mixture``  Gaussians (1-8) + rectangles (0-6) over a floor U[0.02, 0.2]
* ``lognormal`` 2048x2048 f32: log-normal noise (sigma 1) smoothed at 8 px, 100:1 range

Ingest follows main.cpp: gray = 0.299 r + 0.587 g + 0.114 b with channels in
[0,1] (main.cpp:91-99) and nearest resize src = dst * old / new (main.cpp:13-27).
"""
from __future__ import annotations

import numpy as np


def gray_from_rgb8(r, g, b) -> np.ndarray:
    """main.cpp:95-98 on 8-bit channels."""
    r = np.asarray(r, dtype=np.float64) / 255.0
    g = np.asarray(g, dtype=np.float64) / 255.0
    b = np.asarray(b, dtype=np.float64) / 255.0
    return (0.299 * r) + (0.587 * g) + (0.114 * b)


def gray_from_u8(img: np.ndarray) -> np.ndarray:
    """A single-channel 8-bit image read as RGBA with equal channels."""
    return gray_from_rgb8(img, img, img)


def resize_nearest(img: np.ndarray, new_w: int, new_h: int) -> np.ndarray:
    """main.cpp:13-27: src_x = x * old_w / new_w (integer division)."""
    old_h, old_w = img.shape
    ys = (np.arange(new_h, dtype=np.int64) * old_h) // new_h
    xs = (np.arange(new_w, dtype=np.int64) * old_w) // new_w
    return np.ascontiguousarray(img[ys][:, xs], dtype=np.float64)


def gauss5(seed: int, n: int = 256) -> np.ndarray:
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:n, 0:n] / n
    img = np.full((n, n), 0.05)
    for _ in range(5):
        c = rng.uniform(0, 1, 2)
        s = rng.uniform(0.05, 0.2)
        a = rng.uniform(0.3, 1.0)
        img += a * np.exp(-((xx - c[0]) ** 2 + (yy - c[1]) ** 2) / (2 * s * s))
    return np.round(255 * img / img.max()).astype(np.uint8)


def rects12(seed: int, n: int = 512) -> np.ndarray:
    rng = np.random.default_rng(seed)
    img = np.full((n, n), 0.1)
    for _ in range(12):
        w, h = rng.uniform(0.05, 0.3, 2) * n
        x0, y0 = rng.uniform(0, n - w), rng.uniform(0, n - h)
        img[int(y0):int(y0 + h), int(x0):int(x0 + w)] = 1.0
    return np.round(255 * img).astype(np.uint8)


def _value_noise(rng, n: int, cells: int) -> np.ndarray:
    lat = rng.uniform(0, 1, (cells + 1, cells + 1))
    t = (np.arange(n) + 0.5) / n * cells
    i0 = np.minimum(t.astype(np.int64), cells - 1)
    f = t - i0
    f = f * f * (3 - 2 * f)  # smoothstep
    a = lat[i0][:, i0]
    b = lat[i0][:, i0 + 1]
    c = lat[i0 + 1][:, i0]
    d = lat[i0 + 1][:, i0 + 1]
    fx, fy = f[None, :], f[:, None]
    return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy


def noise3(seed: int, n: int = 1024) -> np.ndarray:
    rng = np.random.default_rng(seed)
    img = np.zeros((n, n))
    for o in range(3):
        img += 0.5 ** o * _value_noise(rng, n, 4 * 2 ** o)
    k = np.arange(n) * 16 // n
    checker = np.where((k[:, None] + k[None, :]) % 2 == 0, 1.0, -1.0)
    img = img * (1.0 + 0.3 * checker)
    return (img / img.sum()).astype(np.float32)


def mixture(seed: int, n: int = 256) -> np.ndarray:
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:n, 0:n] / n
    img = np.full((n, n), rng.uniform(0.02, 0.2))
    for _ in range(rng.integers(1, 9)):
        c = rng.uniform(0, 1, 2)
        s = rng.uniform(0.03, 0.2)
        a = rng.uniform(0.3, 1.0)
        img += a * np.exp(-((xx - c[0]) ** 2 + (yy - c[1]) ** 2) / (2 * s * s))
    for _ in range(rng.integers(0, 7)):
        w, h = rng.uniform(0.05, 0.3, 2) * n
        x0, y0 = rng.uniform(0, n - w), rng.uniform(0, n - h)
        img[int(y0):int(y0 + h), int(x0):int(x0 + w)] += rng.uniform(0.3, 1.0)
    return (img / img.max()).astype(np.float32)


def _gauss_smooth(a: np.ndarray, sigma: float) -> np.ndarray:
    n = a.shape[0]
    k = np.fft.fftfreq(n)
    g = np.exp(-2 * (np.pi * sigma) ** 2 * (k[:, None] ** 2 + k[None, :] ** 2))
    return np.real(np.fft.ifft2(np.fft.fft2(a) * g))


def lognormal(seed: int, n: int = 2048) -> np.ndarray:
    rng = np.random.default_rng(seed)
    g = _gauss_smooth(rng.standard_normal((n, n)), 8.0)
    g = (g - g.mean()) / g.std()
    img = np.exp(1.0 * g)
    img = np.clip(img, img.max() / 100.0, None)
    return img.astype(np.float32)


CONFIGS = {
    # name: (generator, source size, mesh M, W, focal)
    "gauss5": (gauss5, 256, 100, 0.5, 1.5),
    "rects12": (rects12, 512, 256, 0.5, 1.5),
    "noise3": (noise3, 1024, 512, 0.5, 1.5),
    "mixture": (mixture, 256, 256, 0.5, 1.5),
    "lognormal": (lognormal, 2048, 1024, 0.5, 1.5),
}


def target_for(name: str, seed: int, mesh_res: int | None = None) -> tuple[np.ndarray, dict]:
    """Image resized to the Poisson grid N = 4*M (main.cpp:222) and the design
    parameters (main.cpp:226-235; square targets so aspect = 1)."""
    gen, size, M, W, f = CONFIGS[name]
    if mesh_res is not None:
        M = mesh_res
    src = gen(seed, size)
    src = gray_from_u8(src) if src.dtype == np.uint8 else src.astype(np.float64)
    N = 4 * M
    img = resize_nearest(src, N, N)
    H = np.floor(M / 1.0) * (W / M)  # main.cpp:229
    return img, dict(mesh_res=M, res=N, width=W, height=float(H), focal_l=f)
