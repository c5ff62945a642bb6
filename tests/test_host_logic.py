"""Host-side logic on the CPU: the parametric (layered) geometry against a numpy restatement of its
rule, the surrogate's parameter sampling and encodings, and the autotuner's parameter store."""
import json

import numpy as np
import pytest

import ymc3d_amd as y
from ymc3d_amd import _abi, autotune, configs
from ymc3d_amd import surrogate as S
from ymc3d_amd import surrogate_data as D


def _voxelize_np(g, shape):
    """Row of voxel (i, j, k): the top layer above top_voxels, the bottom layer below, the vessel
    where the voxel centre lies within the cylinder of radius r about (vessel_x, vessel_z) in the
    x-z plane; every operation in binary32 as the engine performs it."""
    nx, ny, nz = shape
    f32 = np.float32
    i = np.arange(nx, dtype=f32)[None, :]
    k = np.arange(nz, dtype=f32)[:, None]
    fx = (i + f32(0.5)) - f32(g["vessel_x"])
    fz = (k + f32(0.5)) - f32(g["vessel_z"])
    d2 = fx * fx + fz * fz
    r2 = f32(g["vessel_r"]) * f32(g["vessel_r"])
    row = np.where(np.arange(nz)[:, None] < g["top_voxels"], g["row_top"], g["row_bottom"]) * np.ones((1, nx), int)
    if g["row_vessel"] >= 0:
        row = np.where(d2 <= r2, g["row_vessel"], row)
    return np.repeat(row[:, None, :], ny, axis=1).astype(np.uint8)


@pytest.mark.parametrize("seed", range(6))
def test_layered_voxelize_matches_rule(seed):
    rng = np.random.default_rng(seed)
    g = np.zeros(1, dtype=_abi.LAYERED_DTYPE)[0]
    shape = (int(rng.integers(8, 40)), int(rng.integers(2, 9)), int(rng.integers(8, 40)))
    g["top_voxels"] = rng.integers(0, shape[2])
    g["row_top"], g["row_bottom"], g["row_vessel"] = 1, 2, (3 if seed % 3 else -1)
    g["vessel_x"] = rng.uniform(0, shape[0])
    g["vessel_z"] = rng.uniform(0, shape[2])
    # radii that put voxel centres exactly on the circle in some cases
    g["vessel_r"] = rng.choice([rng.uniform(0.5, 8), 2.5, 3.0])
    got = y.layered_voxelize(np.array([g]), shape)
    assert got.shape == (shape[2], shape[1], shape[0])
    np.testing.assert_array_equal(got, _voxelize_np(g, shape))


def _edge_geometry(case):
    g = np.zeros(1, dtype=_abi.LAYERED_DTYPE)[0]
    g["top_voxels"] = 5
    g["row_top"], g["row_bottom"], g["row_vessel"] = 1, 2, 3
    g["vessel_x"], g["vessel_z"], g["vessel_r"] = case["vessel_x"], case["vessel_z"], case["vessel_r"]
    return g


def test_layered_voxelize_rounds_each_operation():
    """Voxel centres whose cylinder test flips when a product is fused into an fma with the sum
    (tests/golden/cylinder_edges.json): the voxeliser gives the answer of separately rounded
    operations, as does the numpy restatement."""
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "cylinder_edges.json")) as f:
        cases = json.load(f)
    assert len(cases) >= 4 and {c["inside"] for c in cases} == {True, False}
    shape = (128, 2, 64)
    for c in cases:
        g = _edge_geometry(c)
        got = y.layered_voxelize(np.array([g]), shape)
        assert (got[c["kc"], 0, c["ic"]] == 3) == c["inside"], c
        np.testing.assert_array_equal(got, _voxelize_np(g, shape))


def test_cylinder_edge_fixture_is_current_and_flips_under_fusion():
    """The fixture equals its generator's output, and in every case the cylinder test with each
    operation rounded on its own and the test with either product fused into an fma (exact
    rational arithmetic, rounded once) give opposite answers."""
    import importlib.util
    import os

    here = os.path.dirname(__file__)
    spec = importlib.util.spec_from_file_location("make_cylinder_edges", os.path.join(here, "golden", "make_cylinder_edges.py"))
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    with open(os.path.join(here, "golden", "cylinder_edges.json")) as f:
        fixture = json.load(f)
    assert gen.cases() == fixture
    f32 = np.float32
    for c in fixture:
        unfused, fused_x, fused_z = gen.forms(c["ic"], c["kc"], f32(c["vessel_x"]), f32(c["vessel_z"]))
        r2 = f32(f32(c["vessel_r"]) * f32(c["vessel_r"]))
        assert (unfused <= r2) == c["inside"]
        assert (fused_x <= r2) != c["inside"] and (fused_z <= r2) != c["inside"], c


def test_skin2_tables_follow_the_parameter_definitions():
    P = configs.sample_skin2_params(64, 3)
    (mua, mus, g, n), geo = configs.skin2_tables(P)
    assert mua.shape == (64, 4) and np.all(mua[:, 0] == 0) and np.all(n[:, 0] == 1)
    np.testing.assert_allclose(mua[:, 1:], P[:, 0:3].astype(np.float32))
    # mus = mus' / (1 - g) (matlab/geometry/prepare_optprops.m:35)
    np.testing.assert_allclose(mus[:, 1:], (P[:, 3:6] / (1 - P[:, 6:9])).astype(np.float32), rtol=1e-6)
    np.testing.assert_array_equal(geo["top_voxels"], np.rint(P[:, 9] / configs.SKIN2_VOXEL))
    np.testing.assert_allclose(geo["vessel_r"], P[:, 11] / 2 / configs.SKIN2_VOXEL, rtol=1e-6)


def test_parameter_sampling_and_encoding():
    P = configs.sample_skin2_params(20_000, 1)
    assert S.in_range(P).all()
    x = S.encode_params(P)
    assert x.min() >= -1 - 1e-6 and x.max() <= 1 + 1e-6
    # log-sampled parameters are uniform in the encoding: mean near 0
    assert np.all(np.abs(x.mean(0)) < 0.03)
    np.testing.assert_allclose(S.encode_params(configs.PARAM_LO[None]), -1, atol=1e-6)
    np.testing.assert_allclose(S.encode_params(configs.PARAM_HI[None]), 1, atol=1e-6)
    Q = configs.sample_skin2_params(20_000, 2, outlier_frac=1.0)
    out = ~S.in_range(Q)
    assert 0.5 < out.mean() < 1.0  # extended range: most rows fall outside the training ranges
    t = S.encode_targets(np.full((2, 32), 1e-9), np.array([0.5, 0.0]))
    assert np.allclose(t[:, :32], np.log10(S.LOG_FLOOR)) and np.isclose(t[0, 32], np.log10(0.5))


def test_hyperspectral_cube():
    Q, outl = D.hyperspectral_cube(16, 16, 64, outlier_frac=0.25, seed=9)
    assert Q.shape == (16 * 16 * 64, 12) and outl.shape == (16 * 16 * 64,)
    assert S.in_range(Q[~outl]).all()
    assert 0.1 < outl.mean() < 0.4


def test_autotune_store(tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "device_name", lambda: "dev")
    a, b = configs.skin_slab_sphere(photons=10), configs.skin_slab_sphere(photons=10, seed=5)
    assert autotune.signature(a, a.params) == autotune.signature(b, b.params)  # seed is not part of it
    c = configs.van_de_hulst_slab(photons=10)
    assert autotune.signature(a, a.params) != autotune.signature(c, c.params)
    assert autotune.tuned(a, a.params) == {}
    best = {"pool_lanes": 1024, "jump_max_crossings": 64, "chain_lanes": 4096}
    (tmp_path / "tune.json").write_text(json.dumps({"dev": {autotune.signature(a, a.params): {"params": best}}}))
    assert autotune.tuned(b, b.params) == best
    assert autotune.tuned(c, c.params) == {}
    # the budget's decade is part of the key: 12 packets share it with 10, 1000 do not
    assert autotune.tuned(configs.skin_slab_sphere(photons=12), a.params | {"num_photons": 12}) == best
    big = configs.skin_slab_sphere(photons=1000)
    assert autotune.signature(big, big.params) != autotune.signature(a, a.params)
    assert autotune.tuned(big, big.params) == {}


def test_autotune_warm_start_coordinates():
    """Stored launch parameters map back to the search coordinates the descent starts from: exact
    for a setting on the grid, snapped to the nearest candidate otherwise, defaults for missing keys."""
    full = 524288
    stored = {"pool_lanes": 458752, "jump_max_crossings": 63, "relaunch_lanes": 10, "relaunch_wait": 32,
              "chain_lanes": 1048576}
    s = autotune._setting(stored, full)
    assert s == {"pool_frac": 0.875, "jump_max_crossings": 63, "relaunch_lanes": 10, "relaunch_wait": 32,
                 "chain_factor": 2}
    assert autotune._candidate(s, full) == stored
    off = autotune._setting({"pool_lanes": 300000, "chain_lanes": 3 * full, "relaunch_lanes": 11}, full)
    assert off["pool_frac"] == 0.625 and off["chain_factor"] in (2, 4) and off["relaunch_lanes"] in (10, 12)
    assert off["jump_max_crossings"] == autotune.DEFAULTS["jump_max_crossings"]
    assert all(off[k] in autotune.SEARCH[k] for k in autotune.SEARCH)
