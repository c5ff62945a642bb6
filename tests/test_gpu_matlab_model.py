"""The reference's MATLAB example model through the files its scripts write, on the GPU (C ABI).

prepare_geometry.m builds a 50x50x85 layered skin model of 0.02 mm voxels with one vessel; the
optical properties come from the spectral models of prepare_optprops.m at 600 nm, extended to the
six tissue values the geometry uses (the reference's example stops at three, which its own comment
says makes it non-working).  The optical-property file is written in MATLAB's column-major layout
and read back with layout="matlab"; the geometry file goes through the C reader.  The device run is
then compared with the oracle on the same streams: total reflectance within 0.5% and the map within
3 combined standard errors (at most 1% of compared bins outside).  Blood absorption spectra are
stand-in constants (muab.mat is not shipped by the reference).
"""
import numpy as np
import pytest

import ymc3d_amd as y
from oracle import oracle as O
from ymc3d_amd import configs, io
from ymc3d_amd import matlab_tools as mt

pytestmark = pytest.mark.gpu


def _model(tmp_path):
    lam = 600.0
    base = mt.optprops_table(lam, lambda x: 30.0, lambda x: 20.0)
    mus, g = base[0, 1], base[0, 2]
    extra = np.array([[0.12, mus * 0.8, g, 1.44],           # subcutis
                      [(227 * 40 * float(mt.melanin(lam)) + 25) * 0.01, mus, g, 1.4],  # mole
                      [0.285, mus * 0.6, 0.98, 1.4]])       # vessel (blood, oxy 0.75)
    table = np.vstack([base, extra])
    props = mt.write_optical_props_file(str(tmp_path / "examplemodel_optprops600"), table, layout="matlab")
    mt.prepare_geometry(str(tmp_path / "examplemodel_geometry"))
    rows, voxel = io.read_geometry(str(tmp_path / "examplemodel_geometry.bin"))
    mua, mus_, g_, n = mt.read_optical_props_file(props, layout="matlab")
    mt.check_model(io.rows_to_file_values(rows), mua.size - 1)
    records = np.stack([mua[1:], mus_[1:], g_[1:], n[1:]], axis=1).astype(np.float64)[None]
    params = dict(num_photons=200_000, seed=11, illumination=1, pencil_x=0.5, pencil_y=0.5)
    return configs.Scene("matlab_example_50x50x85", io.rows_to_file_values(rows), voxel, records, params)


def test_matlab_example_model_parity(engine, multipliers, tmp_path):
    scene = _model(tmp_path)
    assert scene.shape == (50, 50, 85)
    scene.load(engine)
    gres = engine.simulate(**scene.params)
    rres = O.run_scene(scene, multipliers)
    assert gres.totals[0]["photons"] == rres.totals[0]["photons"] == 200_000
    rel = abs(gres.total_R()[0] / rres.total_R()[0] - 1)
    print(f"MATLAB model total R gpu {gres.total_R()[0]:.6f} oracle {rres.total_R()[0]:.6f} rel {rel:.2e}")
    assert rel < 0.005
    Rg, Rc = gres.drs()[0], rres.drs()[0]
    se = np.hypot(gres.drs_se()[0], rres.drs_se()[0])
    mask = (gres.R_n[0] >= 50) & (rres.R_n[0] >= 50) & (se > 0)
    z = (Rg - Rc)[mask] / se[mask]
    print(f"bins compared {mask.sum()}, |z|>3 fraction {np.mean(np.abs(z) > 3):.4f}")
    assert mask.sum() > 100
    assert np.mean(np.abs(z) > 3) <= 0.01
