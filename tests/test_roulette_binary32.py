"""The transport kernel's binary32 roulette equals the reference's binary64 arithmetic (CPU).

The reference survives a packet when a draw e satisfies e < CHANCE and then sets W = W / CHANCE,
both in binary64 with CHANCE = 0.1 (src/mc3d_gpu.cu:275-277, src/mc3d_types.h:11-12); the oracle
keeps that arithmetic.  The kernel uses e < 0.1f and W * 10.0f.  Checked here over every float of
the ranges that occur: every positive W below the roulette threshold 1e-4 (including subnormals),
and every e in [0, 0.2).
"""
import numpy as np


def test_survivor_weight_equals_binary64_division_for_every_float_below_the_threshold():
    top = int(np.float32(1e-4).view(np.uint32))
    chunk = 1 << 24
    for start in range(1, top + 1, chunk):
        W = np.arange(start, min(start + chunk, top + 1), dtype=np.uint32).view(np.float32)
        ref = (W.astype(np.float64) / 0.1).astype(np.float32)
        got = W * np.float32(10.0)
        bad = np.flatnonzero(ref.view(np.uint32) != got.view(np.uint32))
        assert bad.size == 0, W[bad[:5]]


def test_survival_compare_equals_binary64_compare():
    e = np.arange(0, int(np.float32(0.2).view(np.uint32)), dtype=np.uint32).view(np.float32)
    assert np.array_equal(e.astype(np.float64) < 0.1, e < np.float32(0.1))
    # the two floats around the binary64 constant: 0.1f lies above it, its predecessor below
    below = np.nextafter(np.float32(0.1), np.float32(0))
    assert float(np.float32(0.1)) > 0.1 > float(below)
