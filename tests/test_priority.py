"""CPU checks of the oracle's block priorities (dt::compute_block_priorities,
dt/src/block_priority.cu:74-155), pinned by computations that do not share
its steps: the distance by brute force over the seed tiles, the cdf by
sorting, the offsets by counting, and closed forms.  The reference holds no
test or fixture for this function.
"""
from __future__ import annotations

import numpy as np
import pytest


def _seed_tiles(m):
    H, W = m.shape
    gh, gw = (H + 31) // 32, (W + 31) // 32
    s = np.zeros((gh, gw), bool)
    ys, xs = np.nonzero(m)
    s[ys // 32, xs // 32] = True
    return s


def _brute(m):
    s = _seed_tiles(m)
    gh, gw = s.shape
    sy, sx = np.nonzero(s)
    yy, xx = np.mgrid[0:gh, 0:gw]
    d = np.abs(yy[..., None] - sy) + np.abs(xx[..., None] - sx)
    return d.min(axis=-1), s


def _masks():
    rng = np.random.default_rng(11)
    for W, H, k in ((250, 230, 3), (97, 203, 1), (1100, 70, 5), (64, 64, 2), (33, 1, 1), (1, 65, 1)):
        m = np.zeros((H, W), np.uint8)
        for _ in range(k):
            m[rng.integers(0, H), rng.integers(0, W)] = 1
        yield f"{W}x{H}-{k}", m
    m = (rng.random((300, 400)) < 0.002).astype(np.float32) * rng.random((300, 400)).astype(np.float32)
    yield "soft", m


@pytest.mark.parametrize("name", [n for n, _ in _masks()])
def test_oracle_priorities_vs_brute_force(oracle, name):
    m = dict(_masks())[name]
    prio, offsets, level0, d = oracle.block_priorities(m, return_distance=True)
    want_d, seed = _brute(m)
    n = want_d.size
    assert np.array_equal(d, want_d)
    flat = np.sort(want_d.ravel())
    cdf_of = np.searchsorted(flat, want_d, side="right")
    assert np.array_equal(prio, (cdf_of * 64 // (n + 1)).astype(np.uint8))
    assert prio.max() < 64
    assert np.array_equal(offsets, np.array([(prio < k).sum() for k in range(64)], np.uint32))
    assert level0 == int(seed.sum())


def test_oracle_priorities_closed_forms(oracle):
    m = np.zeros((100, 70), np.uint8)
    m[40, 50] = 1  # tile (1, 1) of a 3 x 4 grid
    prio, offsets, level0, d = oracle.block_priorities(m, return_distance=True)
    yy, xx = np.mgrid[0:4, 0:3]
    assert np.array_equal(d, np.abs(yy - 1) + np.abs(xx - 1))
    # distances 0,1,2,3 hold 1,4,5,2 tiles: cdf 1,5,10,12, priority cdf*64//13
    assert np.array_equal(prio, np.array([[49, 24, 49], [24, 4, 24], [49, 24, 49], [59, 49, 59]], np.uint8))
    assert level0 == 1
    # no seed: every cdf is n; every tile seeded: every distance 0
    for fill, l0 in ((0, 0), (1, 12)):
        prio, offsets, level0 = oracle.block_priorities(np.full((100, 70), fill, np.uint8))
        assert (prio == 12 * 64 // 13).all() and level0 == l0
        assert offsets[12 * 64 // 13] == 0 and offsets[63] == 12
    prio, offsets, level0 = oracle.block_priorities(np.zeros((0, 5), np.uint8))
    assert prio.size == 0 and level0 == 0 and not offsets.any()

