"""Oracle transport (CPU): known answer, invariants, determinism, schedule.

- Literature known answer: an index-matched slab (n = 1, so neither Fresnel nor the reference's
  R2 term matters) with mua=10, mus=90 /cm, g=0.75, d=0.02 cm: Rd = 0.09739, Tt = 0.66096
  (van de Hulst 1980, as tabulated by Wang, Jacques and Zheng, Comput. Methods Programs Biomed.
  47:131, 1995).  The slab is tissue voxels over ambient voxels; the lateral mirror walls of
  src/mc3d_gpu.cu:113-126 make it laterally infinite.
- Energy: with every weight change booked in fixed point, N*W0 = R + T + absorbed - roulette
  gain + killed holds exactly (a dropped or double-counted term breaks it).
- Symmetry: a pencil beam at the centre of a mirror-symmetric geometry gives a reflectance map
  symmetric under x and y reflection and transposition.
- Determinism: fixed-point tallies and chain-bound streams make results independent of the thread
  count.
- Chain schedule: the product's guided schedule equals the oracle's restatement and conserves the
  budget.
"""
import json
import os

import numpy as np
import pytest

import ymc3d_amd as y
from oracle import oracle as O
from ymc3d_amd import configs

FX = 2.0 ** 32
GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def test_van_de_hulst_slab(multipliers):
    s = configs.van_de_hulst_slab(photons=1_000_000)
    r = O.run_scene(s, multipliers, chain_lanes=8192)
    rd, tt = r.total_R()[0], r.total_T()[0]
    assert abs(rd - 0.09739) < 4 * r.total_R_se()[0] + 2e-4, rd
    assert abs(tt - 0.66096) < 2e-3, tt


def _energy(res, w0):
    t = res.totals[0]
    fx0 = int(np.rint(np.float64(w0) * FX))
    lhs = int(t["photons"]) * fx0
    rhs = int(t["R_w"]) + int(t["T_w"]) + int(t["A_w"]) - int(t["roulette_w"]) + int(t["killed_w"])
    return lhs, rhs


@pytest.mark.parametrize("fn,n", [(configs.skin_slab_sphere, 20000), (configs.skin_vessels, 5000),
                                  (configs.van_de_hulst_slab, 20000)])
def test_energy_identity(multipliers, fn, n):
    s = fn(photons=n)
    r = O.run_scene(s, multipliers, chain_lanes=1024)
    nrow = s.optprops()[3][0]
    rsp = np.float32(np.float32(nrow[0] - nrow[1]) / np.float32(nrow[0] + nrow[1])) ** 2
    lhs, rhs = _energy(r, np.float32(1) - np.float32(rsp))
    assert lhs == rhs
    t = r.totals[0]
    assert t["photons"] == n and t["killed"] < 1e-3 * n


def test_pencil_beam_symmetry(multipliers):
    s = configs.skin_slab_sphere(photons=300_000)
    r = O.run_scene(s, multipliers, chain_lanes=4096)
    R, se = r.drs()[0], r.drs_se()[0]
    n = r.R_n[0]
    for other, se2, n2 in ((R[:, ::-1], se[:, ::-1], n[:, ::-1]), (R[::-1, :], se[::-1, :], n[::-1, :]),
                           (R.T, se.T, n.T)):
        m = (n >= 100) & (n2 >= 100)
        z = (R - other)[m] / np.hypot(se, se2)[m]
        assert m.sum() > 100
        assert np.mean(np.abs(z) > 3.5) < 0.01
    # the reflectance peaks at the beam: the four centre bins hold the maximum
    assert R[31:33, 31:33].min() >= 0.5 * R.max()


def test_thread_count_does_not_change_results(multipliers):
    s = configs.skin_vessels(photons=4000)
    a = O.run_scene(s, multipliers, chain_lanes=512, threads=1)
    b = O.run_scene(s, multipliers, chain_lanes=512, threads=8)
    assert np.array_equal(a.R_w, b.R_w) and np.array_equal(a.T_w2, b.T_w2)
    assert a.totals.tobytes() == b.totals.tobytes()


def test_tiny_run_golden(multipliers):
    g = GOLDEN["tiny_run"]
    vals = np.array(g["file_values"], np.int32)
    rows = np.where(vals == 255, 0, vals + 1).astype(np.uint8)
    rec = np.array(g["records"])
    props = []
    for q, amb in enumerate((0.0, 0.0, 0.0, 1.0)):
        arr = np.empty((1, rec.shape[1] + 1), dtype=np.float32)
        arr[:, 0] = amb
        arr[:, 1:] = rec[:, :, q]
        props.append(arr)
    table = np.array(GOLDEN["multipliers_first_65"], np.uint32)
    chains = O.build_chains([g["photons"]], g["chain_lanes"])
    st, x, a = O.init_rng(table, g["seed"], chains.size)
    res = O.run(rows, (g["voxel"],) * 3, props, chains, x, a, illumination=1, num_photons=g["photons"])
    assert res.R_w[0].tolist() == g["R_w"] and res.R_n[0].tolist() == g["R_n"]
    assert {k: int(res.totals[0][k]) for k in res.totals.dtype.names} == g["totals"]


@pytest.mark.parametrize("budgets,lanes", [([100_000_000], 524288), ([1_000_000], 524288), ([12345, 1, 0, 999999], 64),
                                           ([7] * 64, 3)])
def test_chain_schedule(budgets, lanes):
    a = y.build_chains(budgets, lanes)
    b = O.build_chains(budgets, lanes)
    assert np.array_equal(a, b)
    for w, bud in enumerate(budgets):
        assert a["count"][a["table"] == w].sum() == bud
    assert a["count"].min() >= 1
    assert a["count"][-1] == 1 or sum(budgets) < 2 * lanes
    # lengths never grow along the dispatch order (guided schedule)
    assert np.all(np.diff(a["count"].astype(np.int64)) <= 0) or len(budgets) > 1


def test_oracle_cancel_flag(multipliers):
    import ctypes as C

    flag = np.ones(1, dtype=np.int32)
    s = configs.skin_slab_sphere(photons=1000)
    r = O.run_scene(s, multipliers, cancel=flag.ctypes.data_as(C.c_void_p))
    assert r.status == y._abi.ERR_CANCELLED and r.totals[0]["photons"] == 0


def test_kill_rules_are_rare(multipliers):
    s = configs.skin_slab_sphere(photons=50_000)
    r = O.run_scene(s, multipliers)
    t = r.totals[0]
    assert t["killed"] <= 1e-3 * t["photons"]
    assert t["killed_w"] / FX / t["photons"] < 1e-4
    # the reference's per-launch zero-step check kills far fewer packets than a per-step check,
    # which would kill every packet whose first scatter turns it back across the beam's face
    r2 = O.run_scene(s, multipliers, kill_zero_step=2)
    assert r2.totals[0]["killed"] > 100 * max(1, int(t["killed"]))


def test_roulette_is_unbiased(multipliers):
    """Russian roulette (src/mc3d_gpu.cu:275-277) must not change the expected weight: a packet
    below 1e-4 survives with p = 0.1 at ten times its weight, so each event changes the tallied
    weight by +9W or -W (W < 1e-4, mean 0, standard deviation <= 3e-4).  The net roulette weight
    of 50 000 packets (about 1.1 events each in this absorbing slab) stays within 5 of those
    standard deviations of zero; a survivor left at W, or a survival chance off by 2x, would shift
    it by more than 2."""
    s = configs.skin_slab_sphere(photons=50_000)
    r = O.run_scene(s, multipliers)
    t = r.totals[0]
    net = int(t["roulette_w"]) / FX
    assert abs(net) < 5 * 3e-4 * np.sqrt(2.0 * t["photons"]), net
    assert t["A_w"] > 0


def test_absorption_grid(multipliers):
    """Absorption grid (the deposit src/mc3d_gpu.cu:265-272 has disabled): every absorbed weight
    lands in a voxel, so the grid sums exactly to the totals' A_w; nothing is absorbed in ambient
    voxels (mu_a = 0); and on the index-matched van de Hulst slab the absorbed fraction is
    1 - Rd - Tt = 1 - 0.09739 - 0.66096 = 0.24165 of the launched weight."""
    n = 100_000
    s = configs.van_de_hulst_slab(photons=n)
    r = O.run_scene(s, multipliers, absorption=True)
    t = r.totals[0]
    grid = r.absorption
    assert grid.shape == (1, 14, 8, 8)
    assert int(grid.sum()) == int(t["A_w"])
    assert not grid[0, 10:].any()
    assert abs(int(t["A_w"]) / FX / n - 0.24165) < 0.006
    # absorption falls with depth in a slab lit from above
    layers = grid[0, :10].sum(axis=(1, 2)).astype(float)
    assert layers[0] > layers[-1]
