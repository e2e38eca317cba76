"""Seeded synthetic inputs for the GGDT workloads.

The reference draws its test images from ``std::random_device``
(dt/include/dt/random_image2d.hpp:13-14), which is not reproducible, and its
benchmark uses an uninitialised mask (benchmarks/distance_transform.cpp:66-67).
This module replaces both with SplitMix64 (Steele, Lea & Flood 2014; the
public ``splitmix64`` mixer of Vigna), evaluated counter-style so that whole
blocks of draws vectorise in numpy:  draw i of a stream with seed s is
``mix(s + (i + 1) * 0x9E3779B97F4A7C15)``.

Workloads (seed = 1000 + workload index):
  blob512    512x512 u8 blob field + 8 seed discs (r=5); C4, lambda=1
  rgb4096    4096x4096 RGB8 (3 blob fields) + soft mask (64 discs, r 3..20,
             nu=1e3); C8, lambda in {0, 0.5, 1}
  maze8192   8192x8192 u8 square spiral maze (wall 3 px = 255, corridor
             9 px = 0, pitch 12 px), seed at (4096, 4096); C8, lambda=1;
             known answer D = image
  batch256   256 u8 images, sides log-uniform in [64, 2048], aspect in
             [0.5, 2], blob field + 1..16 discs (r=5); C8, lambda=0.8
  diff32     32 float32 1024x1024 images (blob/255 + N(0, 1e-3)), soft masks
             as rgb4096, loss weights U[0,1]; C8, lambda=1
"""
from __future__ import annotations

import math

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based SplitMix64 stream."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def next_u64(self, n: int) -> np.ndarray:
        i = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
        self.count += n
        with np.errstate(over="ignore"):
            z = self.seed + i * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def scalar(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return float(self.uniform(1, lo, hi)[0])

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller on pairs of uniforms."""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


def blob_field(rng: SplitMix64, W: int, H: int, n_blobs: int = 16, noise: float = 8.0) -> np.ndarray:
    """Sum of Gaussians (centre U[0,W)xU[0,H), sigma U[10,60], amplitude
    U[0,255]) plus U(-noise, noise) per pixel, as float64 (unclipped)."""
    params = rng.uniform(4 * n_blobs).reshape(n_blobs, 4)
    x = np.arange(W, dtype=np.float64)
    y = np.arange(H, dtype=np.float64)
    acc = np.zeros((H, W), np.float64)
    for cx, cy, sg, amp in params:
        cx, cy = cx * W, cy * H
        sigma = 10.0 + 50.0 * sg
        a = 255.0 * amp
        gx = np.exp(-((x - cx) ** 2) / (2 * sigma * sigma))
        gy = np.exp(-((y - cy) ** 2) / (2 * sigma * sigma))
        acc += a * np.outer(gy, gx)
    if noise > 0:
        acc += rng.uniform(W * H, -noise, noise).reshape(H, W)
    return acc


def to_u8(a: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(a), 0, 255).astype(np.uint8)


def binary_discs(rng: SplitMix64, W: int, H: int, n: int, radius: float = 5.0) -> np.ndarray:
    """uint8 mask, 1 inside n discs of the given radius (centre U[0,W)xU[0,H))."""
    m = np.zeros((H, W), np.uint8)
    c = rng.uniform(2 * n).reshape(n, 2)
    r = int(math.ceil(radius))
    for ux, uy in c:
        cx, cy = ux * W, uy * H
        x0, x1 = max(0, int(cx) - r - 1), min(W, int(cx) + r + 2)
        y0, y1 = max(0, int(cy) - r - 1), min(H, int(cy) + r + 2)
        if x0 >= x1 or y0 >= y1:
            continue
        yy, xx = np.mgrid[y0:y1, x0:x1]
        inside = (xx - cx) ** 2 + (yy - cy) ** 2 <= radius * radius
        m[y0:y1, x0:x1][inside] = 1
    return m


def soft_discs(rng: SplitMix64, W: int, H: int, n: int = 64, rmin: float = 3.0, rmax: float = 20.0) -> np.ndarray:
    """float32 soft mask: s = max over discs of clamp(r + 0.5 - |p - c|, 0, 1)."""
    s = np.zeros((H, W), np.float32)
    params = rng.uniform(3 * n).reshape(n, 3)
    for ux, uy, ur in params:
        cx, cy, rad = ux * W, uy * H, rmin + (rmax - rmin) * ur
        r = int(math.ceil(rad)) + 2
        x0, x1 = max(0, int(cx) - r), min(W, int(cx) + r + 1)
        y0, y1 = max(0, int(cy) - r), min(H, int(cy) + r + 1)
        if x0 >= x1 or y0 >= y1:
            continue
        yy, xx = np.mgrid[y0:y1, x0:x1]
        d = np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2)
        v = np.clip(rad + 0.5 - d, 0.0, 1.0).astype(np.float32)
        s[y0:y1, x0:x1] = np.maximum(s[y0:y1, x0:x1], v)
    return s


def blob512(size: int = 512, seed: int = 1000):
    """Grey blob image + 8 binary seed discs; run with C4, lambda=1, v=1e10."""
    rng = SplitMix64(seed)
    img = to_u8(blob_field(rng, size, size))
    mask = binary_discs(rng, size, size, 8, 5.0)
    return img, mask


def rgb4096(size: int = 4096, seed: int = 1001, n_discs: int = 64):
    """RGB8 image (H, W, 3) + soft float mask; run with C8, nu=1e3."""
    rng = SplitMix64(seed)
    chans = [to_u8(blob_field(rng, size, size)) for _ in range(3)]
    img = np.ascontiguousarray(np.stack(chans, axis=-1))
    mask = soft_discs(rng, size, size, n_discs)
    return img, mask


def spiral_maze(size: int = 8192, wall: int = 3, pitch: int = 12):
    """Square spiral maze: walls (value 255, `wall` px thick) on a 0 background,
    successive turns `pitch` px apart, so corridors are pitch - wall px wide.
    The wall is one non-self-intersecting polyline that starts beside the
    centre and stops before any segment would leave the image, so the
    corridor (its complement) is one connected region that winds from the
    centre to the outer border.  Returns (image, seed (x, y))."""
    img = np.zeros((size, size), np.uint8)
    c = size // 2
    half = wall // 2
    x, y = c + pitch // 2, c
    dirs = [(1, 0), (0, 1), (-1, 0), (0, -1)]
    k = 0
    while True:
        length = pitch * (k // 2 + 1)
        dx, dy = dirs[k % 4]
        nx, ny = x + dx * length, y + dy * length
        if min(nx, ny) - half < 1 or max(nx, ny) + half > size - 2:
            break
        xa, xb = sorted((x, nx))
        ya, yb = sorted((y, ny))
        img[ya - half:yb + half + 1, xa - half:xb + half + 1] = 255
        x, y = nx, ny
        k += 1
    return img, (c, c)


def maze8192(size: int = 8192):
    img, (sx, sy) = spiral_maze(size)
    mask = np.zeros_like(img)
    mask[sy, sx] = 1
    return img, mask


def batch256(count: int = 256, seed: int = 1003):
    """List of (img u8, mask u8) with uneven sizes; run with C8, lambda=0.8."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        s = math.exp(rng.scalar(math.log(64), math.log(2048)))
        a = math.exp(rng.scalar(math.log(0.5), math.log(2.0)))
        W = int(min(4096, max(32, round(s))))
        H = int(min(4096, max(32, round(s * a))))
        img = to_u8(blob_field(rng, W, H))
        nd = 1 + int(rng.scalar(0, 16))
        nd = min(nd, 16)
        mask = binary_discs(rng, W, H, nd, 5.0)
        out.append((img, mask))
    return out


def diff32(count: int = 32, size: int = 1024, seed: int = 1004, n_discs: int = 64):
    """float32 images, soft masks and loss weights; run with C8, lambda=1, nu=1e3."""
    rng = SplitMix64(seed)
    imgs, masks, wts = [], [], []
    for _ in range(count):
        f = blob_field(rng, size, size, noise=0.0) / 255.0
        f = f + 1e-3 * rng.normal(size * size).reshape(size, size)
        imgs.append(f.astype(np.float32))
        masks.append(soft_discs(rng, size, size, n_discs))
        wts.append(rng.uniform(size * size).reshape(size, size).astype(np.float32))
    return np.stack(imgs), np.stack(masks), np.stack(wts)
