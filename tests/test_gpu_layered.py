"""Parametric (per-table) geometry on the GPU.

The kernel evaluates the layered geometry of each table voxel by voxel.  Pinned by:
- bit-identity with the voxel grid the same parameters describe (ymc_layered_voxelize), uploaded as
  an ordinary geometry, with flight jumps off: the rows agree exactly, so step-by-step transport
  gives identical tallies and counters.  With jumps on, the safe radii of the two forms differ (the
  parametric bound is conservative), so different crossings are jumped; the draws are the same and
  positions differ by rounding only, so the totals agree far inside one standard error;
- the oracle on that voxel grid (same chains and streams), per table.
"""
import numpy as np
import pytest

import ymc3d_amd as y
from oracle import oracle as O
from ymc3d_amd import configs

pytestmark = pytest.mark.gpu


def test_layered_equals_materialised_grid(engine, multipliers):
    P = configs.sample_skin2_params(3, seed=21)
    for w in range(3):
        props, geo = configs.skin2_tables(P[w:w + 1])
        params = dict(configs.skin2_params(100_000), seed=3, flags=y.FLAG_COUNTERS)
        rows = y.layered_voxelize(geo[0], configs.SKIN2_SHAPE)
        res = {}
        for mode in (2, 1):
            engine.set_layered(geo, configs.SKIN2_SHAPE, (configs.SKIN2_VOXEL,) * 3)
            engine.set_optprops(*props)
            a = engine.simulate(**params, jump_mode=mode)
            engine.set_geometry(rows, *(configs.SKIN2_VOXEL,) * 3)
            engine.set_optprops(*props)
            b = engine.simulate(**params, jump_mode=mode)
            res[mode] = (a, b)
        a, b = res[2]
        assert np.array_equal(a.Rr_w, b.Rr_w) and np.array_equal(a.Rr_w2, b.Rr_w2)
        assert a.totals.tobytes() == b.totals.tobytes()
        a, b = res[1]
        assert np.array_equal(a.totals["photons"], b.totals["photons"])
        assert abs(a.total_R()[0] - b.total_R()[0]) < 0.3 * a.total_R_se()[0], w


def test_layered_safe_boxes(engine):
    """Device voxel words of parametric geometries against the voxel grid the host voxeliser
    writes: rows identical, and every voxel's safe box (extents zup, zdn, rxy) holds that row only,
    lies inside the lateral walls and below the top surface (brute force on a sample of voxels)."""
    rng = np.random.default_rng(5)
    P = configs.sample_skin2_params(4, seed=31)
    shape = (48, 40, 36)  # (nx, ny, nz): small enough for brute force, with walls nearby
    nx, ny, nz = shape
    for w in range(4):
        _, geo = configs.skin2_tables(P[w:w + 1])
        g = geo[0].copy()
        g["vessel_x"] = g["vessel_x"] * nx / configs.SKIN2_SHAPE[0]
        g["vessel_z"] = min(float(g["vessel_z"]), nz - 4.0)
        g["top_voxels"] = min(int(g["top_voxels"]), nz // 3)
        rows = y.layered_voxelize(g, shape)
        words = engine.layered_words(g, shape)
        assert np.array_equal((words & 0xFF).astype(np.uint8), rows), w
        zup, zdn, rxy = (words >> 8) & 0xFF, (words >> 16) & 0xFF, words >> 24
        for _ in range(3000):
            k, j, i = rng.integers(0, nz), rng.integers(0, ny), rng.integers(0, nx)
            zu, zd, s_ = int(zup[k, j, i]), int(zdn[k, j, i]), int(rxy[k, j, i])
            assert i - s_ >= 0 and i + s_ <= nx - 1 and j - s_ >= 0 and j + s_ <= ny - 1 and k - zu >= 0
            ks = np.minimum(np.arange(k - zu, k + zd + 1), nz - 1)
            assert np.all(rows[ks][:, j - s_:j + s_ + 1, i - s_:i + s_ + 1] == rows[k, j, i]), (w, k, j, i, zu, zd, s_)


def test_cylinder_edges_match_host(engine):
    """Voxel centres whose cylinder test flips when a product is fused into an fma
    (tests/golden/cylinder_edges.json): the device's rows equal the host voxeliser's, both in the
    word hook and inside transport (bit-identical tallies against the materialised grid, jumps off,
    pencil beam aimed at the edge voxel's column)."""
    import json
    import os

    from ymc3d_amd import _abi

    with open(os.path.join(os.path.dirname(__file__), "golden", "cylinder_edges.json")) as f:
        cases = json.load(f)
    shape = configs.SKIN2_SHAPE
    P = configs.sample_skin2_params(1, seed=41)
    props, _ = configs.skin2_tables(P)
    for n, c in enumerate(cases):
        g = np.zeros(1, dtype=_abi.LAYERED_DTYPE)
        g["top_voxels"] = 5
        g["row_top"], g["row_bottom"], g["row_vessel"] = 1, 2, 3
        g["vessel_x"], g["vessel_z"], g["vessel_r"] = c["vessel_x"], c["vessel_z"], c["vessel_r"]
        rows = y.layered_voxelize(g[0], shape)
        assert (rows[c["kc"], 0, c["ic"]] == 3) == c["inside"], c
        words = engine.layered_words(g[0], shape)
        assert np.array_equal((words & 0xFF).astype(np.uint8), rows), c
        if n < 2:
            v = configs.SKIN2_VOXEL
            params = dict(configs.skin2_params(50_000), seed=5, flags=y.FLAG_COUNTERS, jump_mode=2,
                          pencil_x=(c["ic"] + 0.5) * v)
            engine.set_layered(g, shape, (v,) * 3)
            engine.set_optprops(*props)
            a = engine.simulate(**params)
            engine.set_geometry(rows, v, v, v)
            engine.set_optprops(*props)
            b = engine.simulate(**params)
            assert a.totals.tobytes() == b.totals.tobytes(), c
            assert np.array_equal(a.Rr_w, b.Rr_w), c


def test_layered_batch_matches_oracle(engine, multipliers):
    P = configs.sample_skin2_params(6, seed=22)
    props, geo = configs.skin2_tables(P)
    budgets = np.full(6, 40_000, np.uint64)
    params = dict(configs.skin2_params(0), seed=9)
    params.pop("num_photons")
    engine.set_layered(geo, configs.SKIN2_SHAPE, (configs.SKIN2_VOXEL,) * 3)
    engine.set_optprops(*props)
    chains = y.build_chains(budgets, 8192)
    x, a = y.seed_streams(multipliers, 9, chains.size)
    engine.prepare_chains(chains, x, a, **params, flags=y.FLAG_COUNTERS)
    engine.run()
    g = engine.read_tally()
    assert np.array_equal(g.totals["photons"], budgets)
    for w in range(6):
        sc = configs.skin2_scene(P[w], photons=40_000, seed=9)
        sel = chains["table"] == w
        ch = chains[sel].copy()
        ch["table"] = 0
        r = O.run(sc.rows, sc.voxel, sc.optprops(), ch, x[sel], a[sel], **{k: v for k, v in sc.params.items()
                                                                            if k != "num_photons"})
        d = g.total_R()[w] - r.total_R()[0]
        se = np.hypot(g.total_R_se()[w], r.total_R_se()[0])
        assert abs(d) < 3 * se, (w, d, se)
        zr = (g.radial()[w] - r.radial()[0]) / np.maximum(np.hypot(g.radial_se()[w], r.radial_se()[0]), 1e-12)
        assert np.mean(np.abs(zr) > 3) <= 2 / 32, (w, zr)


# A weakly absorbing query (dermis mu_a 0.035/cm): packets walk 1.7e4 reference steps on average and
# up to 5e5, against the mirror walls and through the cylinder many times.
LONG_WALK = [42.57030400217538, 0.03539621026722863, 37.85999861225982, 18.55017331336466, 46.95608120610629,
             54.30326081246652, 0.8404832991257235, 0.9298052533655712, 0.8632422053734459, 0.002,
             0.1094988628650844, 0.002]


@pytest.mark.parametrize("jump_mode", [1, 2])
def test_long_walks_end_as_in_the_oracle(engine, multipliers, jump_mode):
    """How long-walking packets end (escape, roulette, kill) matches the oracle on the same streams.
    The kill rules fire on non-finite state, a negative step, the step cap and zero-length steps at
    the reference's launch boundaries only; a build that killed 1.7% of these packets (seven times
    the oracle's rate, mostly away from launch boundaries) shortened their walks by 12%."""
    n = 20_000
    P = np.array([LONG_WALK])
    props, geo = configs.skin2_tables(P)
    engine.set_layered(geo, configs.SKIN2_SHAPE, (configs.SKIN2_VOXEL,) * 3)
    engine.set_optprops(*props)
    chains = np.zeros(n, dtype=y._abi.CHAIN_DTYPE)
    chains["count"] = 1
    x, a = y.seed_streams(multipliers, 31, n)
    params = configs.skin2_params(0)
    params.pop("num_photons")
    engine.prepare_chains(chains, x, a, flags=y.FLAG_RECORD | y.FLAG_COUNTERS, jump_mode=jump_mode, **params)
    engine.run()
    g = engine.read_records(n)
    sc = configs.skin2_scene(P[0], photons=n, seed=9)
    r = O.run(sc.rows, sc.voxel, sc.optprops(), chains, x, a, records=n,
              **{k: v for k, v in sc.params.items() if k not in ("num_photons", "budgets")}).records
    gf, rf = np.bincount(g["fate"], minlength=3), np.bincount(r["fate"], minlength=3)
    print(f"jump_mode {jump_mode}: fates (escaped, frozen, killed) gpu {gf.tolist()} oracle {rf.tolist()}")
    # killed and frozen counts within 4 Poisson standard deviations of the difference
    for f in (1, 2):
        assert abs(int(gf[f]) - int(rf[f])) <= 4 * np.sqrt(gf[f] + rf[f] + 1), (f, gf, rf)
    # most kills are zero-length steps at the launch boundaries (steps 45, 90, ...)
    dead = g["fate"] == 2
    assert np.mean(g["steps"][dead] % 45 == 0) > 0.6
    # escaped weight within 4 standard errors of the packet weights' difference
    ew_g = np.where(g["fate"] == 0, g["w"], 0.0).astype(np.float64)
    ew_r = np.where(r["fate"] == 0, r["w"], 0.0).astype(np.float64)
    assert abs(ew_g.mean() - ew_r.mean()) < 4 * np.std(ew_g - ew_r) / np.sqrt(n) + 1e-6
