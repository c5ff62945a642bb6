"""Host counterparts of the reference's MATLAB model tools (CPU): ymc3d_amd/matlab_tools.py.

What pins each function (values derived by hand, not by re-running the module's expressions):
- prepare_geometry (matlab/geometry/prepare_geometry.m): layer depths from exact decimal arithmetic
  (0.1/0.02 = 5, 0.15/0.02 = 7.5, 1.35/0.02 = 67.5 -> 5, 12, 79 voxels; 1.7/0.02 = 85).  The vessel
  test at 1-based (i, k) reduces, in units of 0.02 mm, to (i - 25)^2 + (k - 12.5)^2 <= 9: a lattice
  count of 26 (i, k) pairs (5 + 5 + 3 per side of the axis), with no pair within 0.25 units of the
  boundary, so rounding cannot move one.
- melanin / musr / g (prepare_optprops.m:1-14): closed-form points (lambda = 694, 347, 500, 1000).
- optprops_table (prepare_optprops.m:27-51): special cases of the blood terms (equal oxy and deoxy
  spectra make every blood mix 25; oxy-only spectra give the 0.5 and 0.8 saturations apart).  The
  blood spectra themselves (muab.mat) are not shipped by the reference: parity unpinned for them.
- write_optical_props_file / read_optical_props_file: byte layout of MATLAB's column-major fwrite
  (write_optical_props_file.m:6-8) and the C reader's records (src/mc3d_io.cu:38-52), including the
  transposition a MATLAB-written file suffers when read as records.
- read_general_mcfile (matlab/read/read_general_mcfile.m): against files written by the engine's
  save_property writer and by hand.
"""
import os
import struct

import numpy as np
import pytest

from ymc3d_amd import io
from ymc3d_amd import matlab_tools as mt


def test_geometry_layers_and_size():
    voi, d = mt.prepare_geometry()
    assert d == (0.02, 0.02, 0.02)
    assert voi.shape == (85, 50, 50) and voi.dtype == np.int32
    no_vessel = voi != mt.VESSEL
    for z0, z1, code in ((0, 5, mt.EPI), (5, 12, mt.PAPDER), (12, 79, mt.RETDER), (79, 85, mt.SUBC)):
        layer = voi[z0:z1]
        assert np.all(layer[no_vessel[z0:z1]] == code), (z0, z1, code)
    assert set(np.unique(voi).tolist()) == {mt.EPI, mt.PAPDER, mt.RETDER, mt.SUBC, mt.VESSEL}


def test_geometry_vessel_lattice_count():
    voi, _ = mt.prepare_geometry()
    vessel = voi == mt.VESSEL
    assert np.all(vessel == vessel[:, :1, :])  # a cylinder along y
    cross = vessel[:, 0, :]  # [z, x], 0-based
    expected = set()
    for k1, i1s in ((11, range(23, 28)), (12, range(23, 28)), (13, range(23, 28)), (14, range(23, 28)),
                    (10, range(24, 27)), (15, range(24, 27))):
        expected |= {(k1 - 1, i1 - 1) for i1 in i1s}
    assert len(expected) == 26
    assert set(map(tuple, np.argwhere(cross).tolist())) == expected
    assert vessel.sum() == 26 * 50


def test_geometry_file_matches_the_matlab_writer(tmp_path):
    base = str(tmp_path / "examplemodel_geometry")
    voi, _ = mt.prepare_geometry(base)
    path = base + ".bin"
    raw = open(path, "rb").read()
    assert struct.unpack("<4i", raw[:16]) == (3, 50, 50, 85)
    assert struct.unpack("<3d", raw[16:40]) == (0.02, 0.02, 0.02)
    assert len(raw) == 40 + 4 * 50 * 50 * 85
    # MATLAB writes VOI(Nx, Ny, Nz) column-major: x fastest, then y, then z
    flat = np.frombuffer(raw[40:], "<i4")
    i, j, k = 24, 7, 11  # 0-based voxel inside the vessel
    assert flat[i + 50 * (j + 50 * k)] == mt.VESSEL
    data, dx, dy, dz = mt.read_general_mcfile(path, "int32")
    assert (dx, dy, dz) == (0.02, 0.02, 0.02)
    assert data.shape == (50, 50, 85)
    assert np.array_equal(data, voi.transpose(2, 1, 0))
    rows, dvox = io.read_geometry(path)  # the C reader: value v -> row v + 1
    assert dvox == (0.02, 0.02, 0.02)
    assert np.array_equal(rows, voi + 1)


def test_spectral_models_closed_forms():
    assert mt.melanin(694.0) == 1.0
    assert mt.melanin(347.0) == pytest.approx(2.0 ** 3.46, rel=1e-15)
    assert mt.musr(500.0) == pytest.approx(100.0 * (18.780 + 17.6), rel=1e-15)
    assert mt.musr(1000.0) == pytest.approx(100.0 * (18.780 * 2.0 ** -0.22 + 17.6 / 16.0), rel=1e-15)
    assert mt.g(0.0) == 0.62
    assert mt.g(1000.0) == pytest.approx(0.91, abs=1e-15)
    lam = np.array([400.0, 600.0, 850.0])
    assert np.all(np.diff(mt.musr(lam)) < 0) and np.all(np.diff(mt.melanin(lam)) < 0) and np.all(np.diff(mt.g(lam)) > 0)


def test_optprops_table_blood_special_cases():
    lam = 694.0
    # equal oxy and deoxy spectra of 25: every blood mix is 25, so the dermis rows are 25 * 0.01
    t = mt.optprops_table(lam, lambda x: 25.0, lambda x: 25.0)
    assert t.shape == (3, 4)
    assert t[1, 0] == pytest.approx(0.25, rel=1e-14) and t[2, 0] == pytest.approx(0.25, rel=1e-14)
    assert t[0, 0] == pytest.approx((227.0 + 24.75) * 0.01, rel=1e-14)  # melanin(694) = 1
    # oxy-only spectrum of 1000: the saturations 0.5 and 0.8 give blood mixes of 500 and 800
    t = mt.optprops_table(lam, lambda x: 1000.0, lambda x: 0.0)
    assert t[0, 0] == pytest.approx((227.0 + 1.0 + 0.988 * 25) * 0.01, rel=1e-14)
    assert t[1, 0] == pytest.approx((5.0 + 24.75) * 0.01, rel=1e-14)
    assert t[2, 0] == pytest.approx((8.0 + 24.75) * 0.01, rel=1e-14)
    # deoxy-only: 1 - saturation
    t = mt.optprops_table(lam, lambda x: 0.0, lambda x: 1000.0)
    assert t[2, 0] == pytest.approx((2.0 + 24.75) * 0.01, rel=1e-14)
    # scattering, anisotropy and index are the same for the three tissues
    t = mt.optprops_table(500.0, lambda x: 0.0, lambda x: 0.0)
    assert np.all(t[:, 1] == pytest.approx(3638.0 / (1.0 - 0.765) * 0.01, rel=1e-14))
    assert np.all(t[:, 2] == pytest.approx(0.765, rel=1e-15))
    assert np.all(t[:, 3] == 1.4)


def test_optprops_file_layouts_and_the_transposition_pitfall(tmp_path):
    t = np.array([[1.0, 100.0, 0.9, 1.4], [2.0, 200.0, 0.8, 1.37], [3.0, 300.0, 0.7, 1.33]])
    pm = mt.write_optical_props_file(str(tmp_path / "model"), t, layout="matlab")
    assert pm == str(tmp_path / "modelbid")  # no dot, as [fname, 'bid']
    raw = open(pm, "rb").read()
    assert struct.unpack("<i", raw[:4]) == (3,)
    assert struct.unpack("<12d", raw[4:]) == (1.0, 2.0, 3.0, 100.0, 200.0, 300.0, 0.9, 0.8, 0.7, 1.4, 1.37, 1.33)
    mua, mus, g, n = mt.read_optical_props_file(pm, layout="matlab")
    assert np.array_equal(mua, np.float32([0, 1, 2, 3])) and np.array_equal(mus, np.float32([0, 100, 200, 300]))
    assert np.array_equal(g, np.float32([0, 0.9, 0.8, 0.7])) and np.array_equal(n, np.float32([1, 1.4, 1.37, 1.33]))
    # the same file through the C reader's record layout: tissue 1 reads as (mua1, mua2, mua3, mus1)
    tmua, tmus, tg, tn = mt.read_optical_props_file(pm, layout="records")
    assert (tmua[1], tmus[1], tg[1], tn[1]) == (1.0, 2.0, 3.0, 100.0)
    assert (tmua[3], tmus[3], tg[3], tn[3]) == (np.float32(0.7), np.float32(1.4), np.float32(1.37), np.float32(1.33))
    pr = mt.write_optical_props_file(str(tmp_path / "rec"), t, layout="records")
    raw = open(pr, "rb").read()
    assert struct.unpack("<12d", raw[4:]) == tuple(t.ravel())
    got = mt.read_optical_props_file(pr, layout="records")
    for c in range(4):
        assert np.array_equal(got[c][1:], t[:, c].astype(np.float32))
    assert got[3][0] == 1.0 and got[0][0] == 0.0
    # one tissue: both layouts are the same bytes
    a = mt.write_optical_props_file(str(tmp_path / "one_m"), t[:1], layout="matlab")
    b = mt.write_optical_props_file(str(tmp_path / "one_r"), t[:1], layout="records")
    assert open(a, "rb").read() == open(b, "rb").read()
    with pytest.raises(ValueError):
        mt.write_optical_props_file(str(tmp_path / "x"), t, layout="rows")
    with pytest.raises(ValueError):
        mt.read_optical_props_file(pr + "_missing", layout="columns")


def test_matlab_layout_reader_rejects_truncated_file(tmp_path):
    p = tmp_path / "shortbid"
    p.write_bytes(struct.pack("<i", 3) + struct.pack("<11d", *range(11)))
    with pytest.raises(ValueError):
        mt.read_optical_props_file(str(p), layout="matlab")


def test_prepare_optprops_names_and_contents(tmp_path):
    name = str(tmp_path / "examplemodel_optprops")
    paths = mt.prepare_optprops([400.7, 500.0], name, lambda x: 10.0, lambda x: 30.0)
    assert [os.path.basename(p) for p in paths] == ["examplemodel_optprops400bid", "examplemodel_optprops500bid"]
    mua, mus, g, n = mt.read_optical_props_file(paths[1], layout="matlab")
    t = mt.optprops_table(500.0, lambda x: 10.0, lambda x: 30.0)
    assert np.array_equal(mua[1:], t[:, 0].astype(np.float32)) and np.array_equal(g[1:], t[:, 2].astype(np.float32))


def test_prepare_model_reports_the_tissue_mismatch(tmp_path):
    with pytest.raises(ValueError, match=r"tissue values \[3, 5\]"):
        mt.prepare_model([600.0], str(tmp_path / "op"), str(tmp_path / "geo"), lambda x: 1.0, lambda x: 1.0)
    assert (tmp_path / "op600bid").exists() and (tmp_path / "geo.bin").exists()
    mt.check_model(np.array([0, 1, 2, 255]), 3)
    with pytest.raises(ValueError):
        mt.check_model(np.array([0, -1]), 3)


def test_read_general_mcfile_2d_and_3d(tmp_path):
    rng = np.random.default_rng(3)
    a2 = rng.random((4, 6))  # [ny, nx]
    p2 = str(tmp_path / "p2_drs.bin")
    io.write_property(p2, a2, (0.01, 0.02))
    data, dx, dy, dz = mt.read_ymc3d_outputfile(p2)
    assert data.shape == (6, 4, 1) and (dx, dy, dz) == (0.01, 0.02, 0.0)
    assert np.array_equal(data[:, :, 0], a2.T)
    a3 = rng.random((3, 4, 5))  # [nz, ny, nx]
    p3 = str(tmp_path / "p3_abs.bin")
    io.write_property(p3, a3, (0.1, 0.2, 0.3))
    data, dx, dy, dz = mt.read_ymc3d_outputfile(p3)
    assert data.shape == (5, 4, 3) and (dx, dy, dz) == (0.1, 0.2, 0.3)
    assert data[4, 1, 2] == a3[2, 1, 4]
    p = tmp_path / "short.bin"
    p.write_bytes(struct.pack("<3i2d", 2, 3, 3, 1.0, 1.0) + b"\0" * 8)
    with pytest.raises(ValueError):
        mt.read_general_mcfile(str(p))
