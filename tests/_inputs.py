"""Seeded small test inputs shared by the CPU and GPU test files."""
from __future__ import annotations

import numpy as np


def image(kind: str, W: int, H: int, seed: int, smooth: bool = False) -> np.ndarray:
    rng = np.random.default_rng(seed)
    if kind == "u8":
        a = rng.integers(0, 256, (H, W), dtype=np.uint8)
        if smooth:
            yy, xx = np.mgrid[0:H, 0:W]
            a = np.clip(128 + 100 * np.sin(xx / 7.0) * np.cos(yy / 5.0) + rng.integers(-4, 5, (H, W)), 0, 255)
            a = a.astype(np.uint8)
        return a
    if kind == "rgb8":
        return rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    if kind == "f32":
        yy, xx = np.mgrid[0:H, 0:W]
        f = 0.5 + 0.4 * np.sin(xx / 9.0 + seed) * np.cos(yy / 6.0) + 1e-3 * rng.standard_normal((H, W))
        return f.astype(np.float32)
    raise ValueError(kind)


def mask(kind: str, W: int, H: int, seed: int, nseeds: int = 3) -> np.ndarray:
    rng = np.random.default_rng(seed + 7919)
    if kind == "u8":
        m = np.zeros((H, W), np.uint8)
        for _ in range(nseeds):
            m[rng.integers(0, H), rng.integers(0, W)] = rng.integers(1, 256)
        return m
    if kind == "f32":
        s = np.zeros((H, W), np.float32)
        yy, xx = np.mgrid[0:H, 0:W]
        for _ in range(nseeds):
            cx, cy, r = rng.uniform(0, W), rng.uniform(0, H), rng.uniform(1.5, 6.0)
            d = np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2)
            s = np.maximum(s, np.clip(r + 0.5 - d, 0, 1).astype(np.float32))
        return s
    raise ValueError(kind)


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_bit_exact(got: np.ndarray, want: np.ndarray, what: str = ""):
    g, w = bits(got), bits(want)
    if g.shape != w.shape:
        raise AssertionError(f"{what}: shape {g.shape} != {w.shape}")
    bad = np.argwhere(g != w)
    if len(bad):
        y, x = bad[0]
        raise AssertionError(
            f"{what}: {len(bad)} / {g.size} pixels differ; first at (x={x}, y={y}): got {got[y, x]!r} want {want[y, x]!r}")
