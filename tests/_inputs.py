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


def diff_crops(n_crops: int = 2, size: int = 64):
    """64x64 crops of the differentiable-GGDT workload (float32 blob fields
    plus N(0, 1e-3) noise, soft masks, loss weights; 1024x1024 images, seed
    1004): per image, the crop on a 96-pixel grid with the most fractional
    mask pixels among those holding some, but few, full seeds."""
    from distance_transform_benchmarks_amd import synth

    imgs, masks, wts = synth.diff32(n_crops, 1024)
    out = []
    for k in range(n_crops):
        best = None
        for y in range(0, 1024 - size, 96):
            for x in range(0, 1024 - size, 96):
                sm = masks[k, y:y + size, x:x + size]
                frac = int(((sm > 0) & (sm < 1)).sum())
                full = int((sm >= 1).sum())
                if 0 < full < size * size // 8 and (best is None or frac > best[0]):
                    best = (frac, y, x)
        _, y, x = best
        out.append(tuple(np.ascontiguousarray(a[k, y:y + size, x:x + size]) for a in (imgs, masks, wts)))
    return out


def fd_indices(gi: np.ndarray, gm: np.ndarray, seed: int, n_random: int = 200, n_top: int = 28) -> np.ndarray:
    """Random pixels plus those with the largest reference gradients."""
    rng = np.random.default_rng(seed)
    n = gi.size
    return np.unique(np.concatenate([rng.choice(n, min(n, n_random), replace=False),
                                     np.argsort(-np.abs(gi.ravel()))[:n_top],
                                     np.argsort(-np.abs(gm.ravel()))[:n_top]]))


def assert_rel(got: np.ndarray, want: np.ndarray, rtol: float, what: str = ""):
    """|got - want| <= rtol * max(|want|, 1e-3 * max|want|), elementwise."""
    got, want = np.asarray(got, np.float64), np.asarray(want, np.float64)
    floor = 1e-3 * max(np.abs(want).max(), 1e-30)
    rel = np.abs(got - want) / np.maximum(np.abs(want), floor)
    if rel.max() > rtol:
        i = int(rel.argmax())
        raise AssertionError(f"{what}: max relative error {rel.max():.3e} > {rtol} at {i} "
                             f"(got {got.ravel()[i]!r}, want {want.ravel()[i]!r}; {int((rel > rtol).sum())} bad)")


def nonfinite_image(W, H, seed):
    """float32 image with NaN, +inf and -inf pixels scattered over a smooth field."""
    rng = np.random.default_rng(seed)
    img = image("f32", W, H, seed)
    flat = img.reshape(-1)
    idx = rng.choice(flat.size, size=3 * max(1, flat.size // 40), replace=False)
    k = len(idx) // 3
    flat[idx[:k]] = np.nan
    flat[idx[k:2 * k]] = np.inf
    flat[idx[2 * k:]] = -np.inf
    return img
