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


def lib_box(rows, zcap=255, xycap=255):
    rows = np.ascontiguousarray(rows, dtype=np.uint8)
    nz, ny, nx = rows.shape
    zup, zdn, rxy = np.zeros_like(rows), np.zeros_like(rows), np.zeros_like(rows)
    assert y.load_library().ymc_safe_box(ptr(rows, C.c_uint8), nx, ny, nz, zcap, xycap, ptr(zup, C.c_uint8),
                                         ptr(zdn, C.c_uint8), ptr(rxy, C.c_uint8)) == 0
    return zup, zdn, rxy


def _box_ok(rows, k, j, i, zu, zd, s):
    """Box [i-s, i+s] x [j-s, j+s] x [k-zu, k+zd]: voxel's row everywhere, inside the lateral grid,
    not above the top layer, bottom layer repeating below the grid."""
    nz, ny, nx = rows.shape
    if i - s < 0 or i + s > nx - 1 or j - s < 0 or j + s > ny - 1 or k - zu < 0:
        return False
    ks = np.minimum(np.arange(k - zu, k + zd + 1), nz - 1)
    return bool(np.all(rows[ks][:, j - s:j + s + 1, i - s:i + s + 1] == rows[k, j, i]))


def _runs(rows, k, j, i, cap):
    nz = rows.shape[0]
    up = 0
    while up < cap and k - up - 1 >= 0 and rows[k - up - 1, j, i] == rows[k, j, i]:
        up += 1
    dn = 0
    while dn < cap and rows[min(k + dn + 1, nz - 1), j, i] == rows[k, j, i]:
        dn += 1
    return up, dn


def test_safe_box_matches_definition():
    """Brute force on the definition: each voxel's box holds its row only and stays in bounds; it
    contains the safe-radius cube; its lateral extent is the widest square for its vertical
    extents; and it is the best, by the harmonic mean of rxy+1, rxy+1 and (zup+zdn)/2+1, of the
    candidates (r, r), (up, dn), (up, r), (r, dn) that contain the cube (r the safe radius, up/dn
    the vertical runs of the row), with both storage caps."""
    rng = np.random.default_rng(1)
    for shape, (zcap, xycap) in (((9, 11, 13), (255, 255)), ((16, 12, 10), (15, 31)), ((12, 9, 21), (2, 3))):
        rows = np.ones(shape, dtype=np.uint8)
        nz = shape[0]
        rows[: rng.integers(1, 4)] = 3  # a top layer
        for _ in range(2):
            lo = [rng.integers(0, s) for s in shape]
            hi = [min(s, l + rng.integers(1, 4)) for s, l in zip(shape, lo)]
            rows[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = rng.integers(0, 4)
        zup, zdn, rxy = lib_box(rows, zcap, xycap)
        cube = brute(rows)
        for k in range(shape[0]):
            for j in range(shape[1]):
                for i in range(shape[2]):
                    zu, zd, s = int(zup[k, j, i]), int(zdn[k, j, i]), int(rxy[k, j, i])
                    r = min(int(cube[k, j, i]), zcap)
                    assert _box_ok(rows, k, j, i, zu, zd, s), (shape, k, j, i, zu, zd, s)
                    assert zu >= r and zd >= r and s >= r and zu <= zcap and zd <= zcap and s <= xycap
                    up, dn = _runs(rows, k, j, i, zcap)

                    def widest(a, b):
                        w = 0
                        while w < xycap and _box_ok(rows, k, j, i, a, b, w + 1):
                            w += 1
                        return w

                    score = lambda a, b, w: 1.0 / (2.0 / (w + 1) + 2.0 / (a + b + 2))
                    want = (r, r, widest(r, r))
                    best = score(*want)
                    for a, b in ((up, dn), (up, r), (r, dn)):
                        if (a, b) == want[:2]:
                            continue
                        w = widest(a, b)
                        if w >= r and score(a, b, w) > best:
                            want, best = (a, b, w), score(a, b, w)
                    assert (zu, zd, s) == want, (shape, k, j, i, (zu, zd, s), want)


def test_safe_box_layers():
    """Flat layers: the box spans the layer's full height above and below the voxel and reaches the
    lateral walls; the 16-bit storage caps hold."""
    rows = np.full((30, 40, 50), 1, dtype=np.uint8)
    rows[:5] = 2
    zup, zdn, rxy = lib_box(rows)
    assert (zup[2, 20, 25], zdn[2, 20, 25], rxy[2, 20, 25]) == (2, 2, 19)  # epidermis middle
    assert (zup[0, 20, 25], zdn[0, 20, 25], rxy[0, 20, 25]) == (0, 4, 19)  # top voxel: whole layer below
    assert (zup[4, 20, 25], zdn[4, 20, 25]) == (4, 0)  # bottom voxel of the layer
    assert rxy[:, 0].max() == 0 and rxy[:, :, -1].max() == 0
    zup, zdn, rxy = lib_box(rows, 15, 31)
    assert zup.max() <= 15 and zdn.max() <= 15 and rxy.max() <= 31
