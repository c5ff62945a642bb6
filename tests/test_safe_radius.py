"""Safe radius of the voxel words (engine-internal acceleration structure, no reference counterpart).

Pinned by brute force on its definition: r(v) is the largest r <= 31 such that every voxel within
Chebyshev distance r of v has v's row, lies inside the lateral grid and not above the top layer,
with the bottom layer repeating below the grid (the reference's clamp, src/mc3d_gpu.cu:416-418).
"""
import ctypes as C

import numpy as np

import ymc3d_amd as y
from ymc3d_amd._abi import ptr


def brute(rows):
    nz, ny, nx = rows.shape
    out = np.zeros_like(rows)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                t = rows[k, j, i]
                r = 0
                while r < 31:
                    q = r + 1
                    if i - q < 0 or i + q > nx - 1 or j - q < 0 or j + q > ny - 1 or k - q < 0:
                        break
                    ks = np.minimum(np.arange(k - q, k + q + 1), nz - 1)
                    cube = rows[ks][:, j - q:j + q + 1, i - q:i + q + 1]
                    if not np.all(cube == t):
                        break
                    r = q
                out[k, j, i] = r
    return out


def lib_radius(rows):
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    nz, ny, nx = rows.shape
    out = np.zeros_like(rows)
    assert y.load_library().ymc_safe_radius(ptr(rows, C.c_uint8), nx, ny, nz, ptr(out, C.c_uint8)) == 0
    return out


def test_safe_radius_matches_definition():
    rng = np.random.default_rng(0)
    for shape in ((9, 11, 13), (14, 12, 10), (20, 7, 8)):
        rows = np.ones(shape, dtype=np.uint8)
        for _ in range(3):  # a few random boxes of other rows
            lo = [rng.integers(0, s) for s in shape]
            hi = [min(s, l + rng.integers(1, 5)) for s, l in zip(shape, lo)]
            rows[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = rng.integers(0, 4)
        assert np.array_equal(lib_radius(rows), brute(rows)), shape


def test_safe_radius_uniform_block_and_cap():
    rows = np.full((40, 70, 70), 2, dtype=np.uint8)
    r = lib_radius(rows)
    # top layer and lateral walls are 0; deep interior reaches the cap
    assert r[0].max() == 0 and r[:, 0].max() == 0 and r[:, :, -1].max() == 0
    assert r[35, 35, 35] == 31
    assert r[5, 35, 35] == 5  # distance to the top layer
    assert r[39, 35, 35] == 31  # bottom repeats below the grid
