"""GPU parity of the block priorities (dtk_block_priorities through the C ABI;
dt::compute_block_priorities, dt/src/block_priority.cu:74-155) against the
oracle, bit for bit: priorities, bucket offsets and level0."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def dt(dev):
    import distance_transform_benchmarks_amd as m

    return m


def _cases():
    rng = np.random.default_rng(23)
    for W, H, k in ((250, 230, 3), (97, 203, 1), (1100, 70, 5), (64, 64, 2), (33, 1, 1), (1, 65, 1), (1, 1, 1)):
        m = np.zeros((H, W), np.uint8)
        for _ in range(k):
            m[rng.integers(0, H), rng.integers(0, W)] = 7
        yield f"u8-{W}x{H}-{k}", m
    yield "zeros", np.zeros((100, 70), np.uint8)
    yield "ones", np.ones((100, 70), np.uint8)
    soft = (rng.random((300, 400)) < 0.002).astype(np.float32) * rng.random((300, 400)).astype(np.float32)
    yield "f32-soft", soft
    yield "f32-zeros", np.zeros((65, 97), np.float32)


@pytest.mark.parametrize("name", [n for n, _ in _cases()])
def test_block_priorities_vs_oracle(dt, dev, oracle, name):
    import torch

    m = dict(_cases())[name]
    _say("block priorities", name, m.shape, m.dtype)
    prio, offsets, level0 = dt.block_priorities(torch.from_numpy(m).to(dev))
    want_p, want_o, want_l0 = oracle.block_priorities(m)
    assert np.array_equal(prio.cpu().numpy(), want_p)
    assert np.array_equal(offsets.cpu().numpy().astype(np.uint32), want_o)
    assert level0 == want_l0


def test_block_priorities_padded_rows_and_full_size(dt, dev, oracle):
    import torch

    from distance_transform_benchmarks_amd import synth

    _say("block priorities: padded u8 rows")
    rng = np.random.default_rng(3)
    wide = np.zeros((150, 300), np.uint8)
    wide[rng.integers(0, 150, 4), rng.integers(0, 300, 4)] = 1
    wide[:, 257:] = 1  # outside the view: must not count
    view = torch.from_numpy(wide).to(dev)[:, :257]
    prio, offsets, level0 = dt.block_priorities(view)
    want = oracle.block_priorities(wide[:, :257])
    assert np.array_equal(prio.cpu().numpy(), want[0]) and level0 == want[2]
    assert np.array_equal(offsets.cpu().numpy().astype(np.uint32), want[1])

    _, m = synth.rgb4096()
    _say("block priorities: 4096^2 soft mask")
    prio, offsets, level0 = dt.block_priorities(torch.from_numpy(m).to(dev))
    want = oracle.block_priorities(m)
    assert np.array_equal(prio.cpu().numpy(), want[0]) and level0 == want[2]
    assert np.array_equal(offsets.cpu().numpy().astype(np.uint32), want[1])


def test_block_priorities_errors(dt, dev):
    import torch

    with pytest.raises(Exception, match="dtype"):
        dt.block_priorities(torch.zeros((10, 10), dtype=torch.int32, device=dev))
