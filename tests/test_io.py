"""Binary formats of the reference (src/mc3d_io.cu) through the engine's C ABI (CPU).

- geometry.bin: int32 3, int32 nx, ny, nz, f64 dx, dy, dz, int32 values x fastest (the MATLAB
  writer, matlab/geometry/write_geometry_file.m:16-19); the reader adds 1 and maps 255 to the
  ambient row 0 (src/mc3d_io.cu:156-163).
- optprops: int32 T then T records (mua, mus, g, n) as f64; the reader prepends the ambient row
  (src/mc3d_io.cu:38-52).
- save_property: int32 dims, int32 nx, ny[, nz], f64 dx, dy[, dz], f64 data (src/mc3d_io.cu:175-216),
  read back by the layout of matlab/read/read_general_mcfile.m.
"""
import struct

import numpy as np
import pytest

import ymc3d_amd as y
from ymc3d_amd import io


def test_geometry_round_trip_and_remap(tmp_path):
    rng = np.random.default_rng(0)
    vals = rng.integers(0, 4, (5, 6, 7)).astype(np.int32)
    vals[0, :, :] = 255
    p = str(tmp_path / "g.bin")
    io.write_geometry(p, vals, 0.01, 0.02, 0.005)
    raw = open(p, "rb").read()
    assert struct.unpack("<4i", raw[:16]) == (3, 7, 6, 5)
    assert struct.unpack("<3d", raw[16:40]) == (0.01, 0.02, 0.005)
    assert np.array_equal(np.frombuffer(raw[40:], "<i4").reshape(5, 6, 7), vals)  # x fastest
    rows, d = io.read_geometry(p)
    assert d == (0.01, 0.02, 0.005)
    assert np.array_equal(rows, np.where(vals == 255, 0, vals + 1))
    assert np.array_equal(io.rows_to_file_values(rows), vals)


def test_geometry_errors(tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(struct.pack("<4i3d", 2, 2, 2, 2, 1.0, 1.0, 1.0) + b"\0" * 32)
    with pytest.raises(y.YmcError) as ei:
        io.read_geometry(str(p))
    assert ei.value.status == y._abi.ERR_FORMAT  # not 3-D (src/mc3d_io.cu:120-124)
    p.write_bytes(struct.pack("<4i3d", 3, 2, 2, 2, 1.0, 1.0, 1.0) + b"\0" * 16)
    with pytest.raises(y.YmcError) as ei:
        io.read_geometry(str(p))
    assert ei.value.status == y._abi.ERR_IO  # short voxel data
    p.write_bytes(struct.pack("<4i3d", 3, 1, 1, 2, 1.0, 1.0, 1.0) + struct.pack("<2i", 0, 300))
    with pytest.raises(y.YmcError) as ei:
        io.read_geometry(str(p))
    assert ei.value.status == y._abi.ERR_FORMAT  # value outside 0..255
    with pytest.raises(y.YmcError) as ei:
        io.read_geometry(str(tmp_path / "missing.bin"))
    assert ei.value.status == y._abi.ERR_IO


def test_optprops_round_trip(tmp_path):
    rec = np.array([[1.0, 100.0, 0.9, 1.4], [20.0, 90.0, 0.8, 1.33]])
    p = str(tmp_path / "o.bid")
    io.write_optprops(p, rec)
    raw = open(p, "rb").read()
    assert struct.unpack("<i", raw[:4]) == (2,)
    assert np.array_equal(np.frombuffer(raw[4:], "<f8").reshape(2, 4), rec)  # record by record
    mua, mus, g, n = io.read_optprops(p)
    assert mua.tolist() == [0, 1, 20] and mus.tolist() == [0, 100, 90]
    assert g.tolist() == pytest.approx([0, 0.9, 0.8]) and n.tolist() == pytest.approx([1.0, 1.4, 1.33])
    (tmp_path / "short.bid").write_bytes(struct.pack("<i", 3) + b"\0" * 40)
    with pytest.raises(y.YmcError):
        io.read_optprops(str(tmp_path / "short.bid"))


def test_property_layout(tmp_path):
    data = np.arange(12, dtype=np.float64).reshape(3, 4) / 7
    p = str(tmp_path / "r_drs.bin")
    io.write_property(p, data, (0.01, 0.02))
    raw = open(p, "rb").read()
    assert len(raw) == 28 + 12 * 8  # 28-byte 2-D header
    assert struct.unpack("<3i2d", raw[:28]) == (2, 4, 3, 0.01, 0.02)
    back, d = io.read_property(p)
    assert np.array_equal(back, data) and d == (0.01, 0.02)
    vol = np.arange(24, dtype=np.float64).reshape(2, 3, 4)
    io.write_property(p, vol, (0.1, 0.2, 0.3))
    back, d = io.read_property(p)
    assert np.array_equal(back, vol) and d == (0.1, 0.2, 0.3)
    assert struct.unpack("<4i", open(p, "rb").read()[:16]) == (3, 4, 3, 2)


def test_scene_files_load_back(tmp_path):
    from ymc3d_amd import configs

    s = configs.skin_vessels(photons=1)
    io.write_geometry(str(tmp_path / "g.bin"), s.file_values, *s.voxel)
    io.write_optprops(str(tmp_path / "o.bid"), s.records[0])
    rows, d = io.read_geometry(str(tmp_path / "g.bin"))
    assert np.array_equal(rows, s.rows) and d == pytest.approx(s.voxel)
    mua, mus, g, n = io.read_optprops(str(tmp_path / "o.bid"))
    ref = s.optprops()
    for got, want in zip((mua, mus, g, n), ref):
        assert np.array_equal(got, want[0])
