"""GPU parity of the HIP transport path against the CPU oracle (through the C ABI).

Bars (from the reference's path, src/mc3d_gpu.cu:24-376 and gpumcml_rng/gpumcml_rng.cu:38-54):
- MWC draws and stream states: bit-exact.
- Per-packet lockstep: with the same stream per packet the device and the oracle take the same
  draws in the same order, so most packets finish after the same number of steps with the same
  weight; the exceptions are packets where a float-rounding difference flips a comparison.
- Maps: per bin within 3 combined Monte Carlo standard errors (bins with >= 50 packets on both
  sides; the fraction outside 3 SE may not exceed 1%), total diffuse reflectance within 0.5%.
- Tallies are bit-identical for any pool size / poll interval (fixed-point, chain-bound streams).
"""
import ctypes as C

import numpy as np
import pytest

import ymc3d_amd as y
from oracle import oracle as O
from ymc3d_amd import configs

pytestmark = pytest.mark.gpu

FX = 2.0 ** 32


def test_rng_bit_exact(engine, multipliers):
    x, a = y.seed_streams(multipliers, 987654321, 4096)
    # extreme but valid states: carry a-2, value 2^32-2
    x[:4] = [(int(a[0]) - 2) << 32 | 0xFFFFFFFE, 1, 1 << 32, (int(a[3]) - 2) << 32]
    for kind in (0, 1):
        got, gx = engine.rng_fill(x, a, 200, kind)
        want, wx = O.mwc_fill(x, a, 200, kind)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert np.array_equal(gx, wx)
    assert got.min() > 0.0 and got.max() <= 1.0  # oc in (0, 1]


def test_draw_conversion_edges(engine, multipliers):
    """rand_MWC_co / rand_MWC_oc at the edges of the round-toward-zero conversion: zero, values
    below 2^24 (exact), just above 2^24 and 2^31 (low bits truncated), and values that round to
    nearest up to 2^32.  A state with low word 0 and carry u draws u first."""
    us = [0, 1, 2, 255, 0xFFFFFF, 0x1000000, 0x1000001, 0x1FFFFFF, 0x7FFFFFFF, 0x80000000, 0x80000081,
          0xFFFFFF7F, 0xFFFFFF80, 0xFFFFFFFE, 0xFFFFFFFF]
    x = np.array([u << 32 for u in us], dtype=np.uint64)
    a = multipliers[:len(us)].astype(np.uint32)

    def rz(u):  # keep the 24 leading significant bits, then scale by 2^-32 (both exact in f32)
        sh = max(0, u.bit_length() - 24)
        return np.float32((u >> sh) << sh) * np.float32(2.0 ** -32)

    co = np.array([rz(u) for u in us], dtype=np.float32)
    for kind, want in ((0, co), (1, np.float32(1.0) - co)):
        got, _ = engine.rng_fill(x, a, 1, kind)
        assert np.array_equal(got[0].view(np.uint32), want.view(np.uint32)), kind
        ora, _ = O.mwc_fill(x, a, 1, kind)
        assert np.array_equal(ora[0].view(np.uint32), want.view(np.uint32)), kind


def test_stream_skip_bit_exact(engine, multipliers):
    """A box flight skips the draws of n crossings by stepping n mod 4 crossings and taking up to two
    modular products (Montgomery, for the multiples of 4 below 32 and for 32); the states must equal
    those of 4n draws taken one by one, for every n a flight can take and for extreme states (the
    largest states p - 1 and p - 2, and the state 1).  States that are not MWC states of their
    multiplier are refused."""
    x, a = y.seed_streams(multipliers, 4711, 1024)
    p = (a.astype(np.uint64) << np.uint64(32)) - np.uint64(1)
    x[:4] = [1, (int(a[1]) - 2) << 32 | 0xFFFFFFFE, int(p[2]) - 1, (int(a[3]) - 1) << 32 | 0xFFFFFFFD]
    for n in range(64):
        got = engine.rng_skip(x, a, n)
        _, want = O.mwc_fill(x, a, 4 * n, 0)
        assert np.array_equal(got, want), n
    for bad_x, bad_a in ((p[:1], a[:1]), (x[:1], np.ones(1, np.uint32))):
        with pytest.raises(y.YmcError):
            engine.rng_skip(bad_x, bad_a, 1)


def _chains_one_per_packet(multipliers, n, seed):
    chains = np.zeros(n, dtype=y._abi.CHAIN_DTYPE)
    chains["count"] = 1
    x, a = y.seed_streams(multipliers, seed, n)
    return chains, x, a


@pytest.mark.parametrize("scene_fn,n", [(configs.skin_slab_sphere, 20000), (configs.skin_vessels, 5000)])
def test_lockstep_packets(engine, multipliers, scene_fn, n):
    scene = scene_fn(photons=n)
    scene.load(engine)
    chains, x, a = _chains_one_per_packet(multipliers, n, 777)
    params = {k: v for k, v in scene.params.items() if k not in ("num_photons", "budgets")}
    engine.prepare_chains(chains, x, a, flags=y.FLAG_RECORD | y.FLAG_COUNTERS, **params)
    engine.run()
    g = engine.read_records(n)
    r = O.run(scene.rows, scene.voxel, scene.optprops(), chains, x, a, records=n, **params).records
    same = (g["steps"] == r["steps"]) & (g["fate"] == r["fate"])
    frac = same.mean()
    print(f"{scene.name}: identical step count and fate for {frac:.4f} of {n} packets")
    # Divergence comes from float rounding and grows with path length: the oracle against itself
    # built with FMA contraction agrees on 0.82 (slab-sphere) / 0.96 (vessels) of packets, 0.999
    # below 100 steps.  The device must do as well, stratified by length.
    st = r["steps"]
    for lo, hi, need in ((0, 100, 0.99), (100, 500, 0.97), (500, 1000, 0.85)):
        m = (st >= lo) & (st < hi)
        if m.sum() >= 200:
            assert same[m].mean() >= need, (lo, hi, same[m].mean())
    assert frac > 0.75
    w_ok = np.abs(g["w"][same] - r["w"][same]) <= 1e-4 * np.maximum(r["w"][same], 1e-3)
    assert w_ok.mean() > 0.99
    # End positions of the packets that carry weight (those the tallies bin).  A packet frozen by
    # roulette has weight zero and the device keeps the start of its last flight as its position
    # where the oracle keeps the last face the flight crossed.
    carried = same & (r["w"] > 0)
    pos = np.abs(g["x"][carried] - r["x"][carried]) + np.abs(g["y"][carried] - r["y"][carried])
    assert carried.sum() > n // 20
    assert np.quantile(pos, 0.99) < 1e-3


@pytest.mark.parametrize("flags", [y.FLAG_RECORD, y.FLAG_RECORD | y.FLAG_COUNTERS])
def test_totals_equal_the_records(engine, multipliers, flags):
    """The reflectance and transmittance totals of a run (fixed-point sums of W and W^2, counts)
    equal their recomputation in exact integers from the same run's per-packet records: f =
    RNE(W 2^32), f^2 term (f^2 + 2^31) >> 32, reflected when z <= 0, nothing for zero weight.  Pins
    the device's fixed-point conversion and its square, which the statistical comparisons cannot."""
    n = 20000
    scene = configs.skin_vessels(photons=n)
    scene.load(engine)
    chains, x, a = _chains_one_per_packet(multipliers, n, 4242)
    params = {k: v for k, v in scene.params.items() if k not in ("num_photons", "budgets")}
    engine.prepare_chains(chains, x, a, flags=flags, **params)
    engine.run()
    rec = engine.read_records(n)
    t = engine.read_tally().totals[0]
    f = np.rint(rec["w"].astype(np.float64) * 2.0 ** 32).astype(np.uint64)
    assert f.max() <= 2 ** 32 and (rec["w"] > 0).sum() > n // 2
    want = {"R_w": 0, "R_w2": 0, "R_n": 0, "T_w": 0, "T_w2": 0, "T_n": 0}
    for fi, z in zip(f.tolist(), rec["z"].tolist()):
        if fi == 0:
            continue
        side = "R" if z <= 0.0 else "T"
        want[side + "_w"] += fi
        want[side + "_w2"] += (fi * fi + (1 << 31)) >> 32
        want[side + "_n"] += 1
    for k, v in want.items():
        assert int(t[k]) == v, (k, int(t[k]), v)
    assert int(t["photons"]) == n


def _compare_maps(gres, rres, min_count=50):
    Rg, Rc = gres.drs()[0], rres.drs()[0]
    se = np.hypot(gres.drs_se()[0], rres.drs_se()[0])
    mask = (gres.R_n[0] >= min_count) & (rres.R_n[0] >= min_count) & (se > 0)
    z = (Rg - Rc)[mask] / se[mask]
    return z, mask.sum()


def test_c1_parity_same_streams(engine, multipliers):
    """Slab with inclusion, pencil beam, 1e6 packets, identical streams on both sides."""
    scene = configs.skin_slab_sphere(photons=1_000_000)
    scene.load(engine)
    g = engine.simulate(**scene.params, flags=y.FLAG_COUNTERS)
    r = O.run_scene(scene, multipliers)
    assert g.totals[0]["photons"] == r.totals[0]["photons"] == 1_000_000
    rel = abs(g.total_R()[0] / r.total_R()[0] - 1)
    print(f"total R gpu {g.total_R()[0]:.6f} oracle {r.total_R()[0]:.6f} rel diff {rel:.2e}")
    assert rel < 0.005
    z, nb = _compare_maps(g, r)
    print(f"bins compared {nb}, |z|>3 fraction {np.mean(np.abs(z) > 3):.4f}")
    assert nb > 500
    assert np.mean(np.abs(z) > 3) <= 0.01


def test_c1_parity_independent_streams(engine, multipliers):
    """Independent Monte Carlo realisations: a proper statistical comparison."""
    scene = configs.skin_slab_sphere(photons=1_000_000)
    scene.load(engine)
    g = engine.simulate(**{**scene.params, "seed": 2024})
    r = O.run_scene(scene, multipliers, seed=99)
    d = g.total_R()[0] - r.total_R()[0]
    se = np.hypot(g.total_R_se()[0], r.total_R_se()[0])
    print(f"total R gpu {g.total_R()[0]:.6f} oracle {r.total_R()[0]:.6f} diff {d:.2e} se {se:.2e}")
    assert abs(d) < 3 * se
    assert abs(d) / r.total_R()[0] < 0.005
    z, nb = _compare_maps(g, r)
    assert nb > 500
    assert np.mean(np.abs(z) > 3) <= 0.01
    # chi-square of z over bins stays near its expectation
    assert abs(np.mean(z ** 2) - 1.0) < 0.15


def test_c2_sample_parity(engine, multipliers):
    """Layered skin with vessels (256x256x128), uniform illumination, 1.6e7-packet sample (about
    29 reflected packets per bin): total R within 0.5%; the full-resolution 256x256 map within 3
    combined SE in bins with at least 20 packets each side (at most 1% of them outside); 16x16
    coarse bins likewise (at most 2% outside)."""
    scene = configs.skin_vessels(photons=16_000_000)
    scene.load(engine)
    g = engine.simulate(**scene.params, flags=y.FLAG_COUNTERS)
    r = O.run_scene(scene, multipliers)
    rel = abs(g.total_R()[0] / r.total_R()[0] - 1)
    assert rel < 0.005, rel
    z, nb = _compare_maps(g, r, min_count=20)
    print(f"skin_vessels full-resolution bins compared {nb}, |z|>3 fraction {np.mean(np.abs(z) > 3):.4f}")
    assert nb > 20000 and np.mean(np.abs(z) > 3) <= 0.01
    # 16x16 coarse bins
    cg = g.R_w[0].reshape(16, 16, 16, 16).sum((1, 3)).astype(float)
    cr = r.R_w[0].reshape(16, 16, 16, 16).sum((1, 3)).astype(float)
    c2g = g.R_w2[0].reshape(16, 16, 16, 16).sum((1, 3)).astype(float)
    c2r = r.R_w2[0].reshape(16, 16, 16, 16).sum((1, 3)).astype(float)
    N = 16_000_000
    se = np.sqrt((c2g / FX / N - (cg / FX / N) ** 2) / N + (c2r / FX / N - (cr / FX / N) ** 2) / N)
    z = (cg - cr) / FX / N / se
    assert np.mean(np.abs(z) > 3) <= 0.02
    assert g.totals[0]["fresnel"] > 0


def test_c2_parity_independent_streams(engine, multipliers):
    """Layered skin with vessels (256x256x128), uniform illumination: 1.6e7 packets per side on
    independent streams (seeds 31 and 4099), so the combined standard error is the noise of the
    difference.  Total R within 3 SE and within 0.5%; the full-resolution map with at most 1% of
    bins (>= 20 packets each side) beyond 3 SE; on 64x64 coarse bins (about 470 packets each) the
    mean of z^2 within 0.1 of 1 (its standard deviation for 4096 normal z is 0.022) and at most 1%
    of bins beyond 3 SE."""
    n = 16_000_000
    scene = configs.skin_vessels(photons=n)
    scene.load(engine)
    g = engine.simulate(**{**scene.params, "seed": 31})
    r = O.run_scene(scene, multipliers, seed=4099)
    d = g.total_R()[0] - r.total_R()[0]
    se = np.hypot(g.total_R_se()[0], r.total_R_se()[0])
    print(f"skin_vessels total R gpu {g.total_R()[0]:.6f} oracle {r.total_R()[0]:.6f} diff {d:.2e} se {se:.2e}", flush=True)
    assert abs(d) < 3 * se
    assert abs(d) / r.total_R()[0] < 0.005
    z, nb = _compare_maps(g, r, min_count=20)
    print(f"skin_vessels independent: bins {nb}, |z|>3 fraction {np.mean(np.abs(z) > 3):.4f}", flush=True)
    assert nb > 20000 and np.mean(np.abs(z) > 3) <= 0.01

    def coarse(res):
        w = res.R_w[0].reshape(64, 4, 64, 4).sum((1, 3)).astype(float) / FX
        w2 = res.R_w2[0].reshape(64, 4, 64, 4).sum((1, 3)).astype(float) / FX
        return w / n, np.sqrt(np.maximum(w2 / n - (w / n) ** 2, 0) / n)

    mg, sg = coarse(g)
    mr, sr = coarse(r)
    zc = (mg - mr) / np.hypot(sg, sr)
    print(f"skin_vessels independent 64x64: mean z^2 {np.mean(zc ** 2):.4f}, |z|>3 {np.mean(np.abs(zc) > 3):.4f}", flush=True)
    assert abs(np.mean(zc ** 2) - 1.0) < 0.1
    assert np.mean(np.abs(zc) > 3) <= 0.01


def test_pool_size_does_not_change_results(engine, multipliers):
    """Packets draw from their chains' streams: the pool size, the cancel-poll interval and when
    a lane relaunches (at once, or batched with others of its wave) change no tally bit."""
    scene = configs.skin_slab_sphere(photons=100_000)
    scene.load(engine)
    base = dict(scene.params, flags=y.FLAG_COUNTERS, chain_lanes=8192)
    a = engine.simulate(**base)
    for over in (dict(pool_lanes=256 * 37, compact_interval=97), dict(relaunch_lanes=1),
                 dict(relaunch_lanes=64, relaunch_wait=3)):
        b = engine.simulate(**base, **over)
        for f in ("R_w", "R_w2", "R_n", "T_w", "T_n"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), (over, f)
        assert a.totals.tobytes() == b.totals.tobytes(), over


def test_energy_identity(engine, multipliers):
    """N*W0 = R + T + absorbed - roulette gain + killed, exactly in fixed point."""
    scene = configs.skin_vessels(photons=50_000)
    scene.load(engine)
    g = engine.simulate(**scene.params, flags=y.FLAG_COUNTERS)
    t = g.totals[0]
    n = scene.optprops()[3][0]
    rsp = np.float32((n[0] - n[1]) / (n[0] + n[1])) ** 2
    w0 = np.float32(1.0) - np.float32(rsp)
    fx0 = int(np.rint(np.float64(w0) * FX))
    lhs = int(t["photons"]) * fx0
    rhs = int(t["R_w"]) + int(t["T_w"]) + int(t["A_w"]) - int(t["roulette_w"]) + int(t["killed_w"])
    assert lhs == rhs


def test_van_de_hulst_slab_gpu(engine, multipliers):
    """Index-matched slab mua=10 mus=90 g=0.75 d=0.02: Rd 0.09739, Tt 0.66096 (van de Hulst 1980,
    as tabulated by Wang, Jacques and Zheng, 1995)."""
    scene = configs.van_de_hulst_slab(photons=2_000_000)
    scene.load(engine)
    g = engine.simulate(**scene.params)
    rd, tt = g.total_R()[0], g.total_T()[0]
    print(f"Rd {rd:.5f} Tt {tt:.5f}")
    assert abs(rd - 0.09739) < 4 * g.total_R_se()[0] + 2e-4
    assert abs(tt - 0.66096) < 2e-3


def test_cancel_flag(engine, multipliers):
    scene = configs.skin_slab_sphere(photons=50_000)
    scene.load(engine)
    flag = np.zeros(1, dtype=np.int32)
    flag[0] = 1
    cancel = flag.ctypes.data_as(C.c_void_p)
    engine.prepare(**scene.params, cancel=cancel)
    with pytest.raises(y.YmcError) as ei:
        engine.run()
    assert ei.value.status == y._abi.ERR_CANCELLED
    t = engine.read_tally().totals[0]
    assert t["photons"] == 0


def test_run_3dmc_gpu_dropin(multipliers):
    """The drop-in (the arguments of run_3dmc_gpu, src/mc3d_gpu.h:16; uniform illumination as the
    reference hard-wires, :87-95) against the oracle on the same seed and multiplier table: raw
    R_tot and T_tot within 3 combined SE (same streams, so this bar is loose), maps that sum to the
    totals, and the finished count."""
    n_ph = 200_000
    scene = configs.skin_slab_sphere(photons=n_ph)
    rows = scene.rows.astype(np.int32)
    mua, mus, g, n = (v[0] for v in scene.optprops())
    R, Rtot, T, Ttot, nf = y.run_3dmc_gpu(rows, *scene.voxel, mua, mus, g, n, n_ph, seed=5, multipliers=multipliers)
    assert nf == n_ph
    assert R.shape == (64, 64) and abs(R.sum() - Rtot) < 1e-6 * max(Rtot, 1)
    assert T.shape == (64, 64) and abs(T.sum() - Ttot) <= 1e-6 * max(Ttot, 1)
    r = O.run_scene(scene, multipliers, seed=5, illumination=y._abi.ILLUM_UNIFORM)
    t = r.totals[0]
    for got, w, w2 in ((Rtot, t["R_w"], t["R_w2"]), (Ttot, t["T_w"], t["T_w2"])):
        mean = float(w) / FX / n_ph
        se = np.sqrt(max(float(w2) / FX / n_ph - mean ** 2, 0.0) / n_ph)
        print(f"drop-in {got / n_ph:.6f} oracle {mean:.6f} se {se:.2e}")
        assert abs(got / n_ph - mean) < 3 * np.sqrt(2) * se + 1e-12


@pytest.mark.parametrize("scene_fn,n", [(configs.skin_slab_sphere, 20000), (configs.skin_vessels, 5000)])
def test_flight_jumps_keep_lockstep(engine, multipliers, scene_fn, n):
    """A box flight takes the plain crossings inside a safe box in one step and advances the stream
    by their draws: with the same stream per packet, transport with and without box flights takes the
    same draws, and packets differ only where a rounding difference in the flight's end flips a
    later comparison.  The bars are those of the device-against-oracle lockstep test."""
    scene = scene_fn(photons=n)
    scene.load(engine)
    chains, x, a = _chains_one_per_packet(multipliers, n, 4242)
    params = {k: v for k, v in scene.params.items() if k not in ("num_photons", "budgets")}
    rec = {}
    for mode in (2, 1):
        engine.prepare_chains(chains, x, a, flags=y.FLAG_RECORD | y.FLAG_COUNTERS, jump_mode=mode, **params)
        engine.run()
        rec[mode] = engine.read_records(n)
    off, on = rec[2], rec[1]
    same = (off["steps"] == on["steps"]) & (off["fate"] == on["fate"])
    print(f"{scene.name}: identical step count and fate with and without jumps for {same.mean():.4f} of {n}")
    st = off["steps"]
    for lo, hi, need in ((0, 100, 0.99), (100, 500, 0.97), (500, 1000, 0.85)):
        m = (st >= lo) & (st < hi)
        if m.sum() >= 200:
            assert same[m].mean() >= need, (lo, hi, same[m].mean())
    assert same.mean() > 0.75
    w_ok = np.abs(off["w"][same] - on["w"][same]) <= 1e-4 * np.maximum(off["w"][same], 1e-3)
    assert w_ok.mean() > 0.99


@pytest.mark.parametrize("scene_fn,n", [(configs.skin_vessels, 200_000), (configs.skin_slab_sphere, 100_000)])
def test_flight_jumps_tallies(engine, multipliers, scene_fn, n):
    """Tallies with and without jumps: the same packets, exact energy identity on both, total
    reflectance far inside one standard error, and the same reference step count to 0.5%."""
    scene = scene_fn(photons=n)
    scene.load(engine)
    base = dict(scene.params, flags=y.FLAG_COUNTERS)
    slow = engine.simulate(**base, jump_mode=2)
    fast = engine.simulate(**base)
    nr = scene.optprops()[3][0]
    rsp = np.float32((nr[0] - nr[1]) / (nr[0] + nr[1])) ** 2
    fx0 = int(np.rint(np.float64(np.float32(1.0) - np.float32(rsp)) * FX))
    for res in (slow, fast):
        t = res.totals[0]
        assert int(t["photons"]) == n
        rhs = int(t["R_w"]) + int(t["T_w"]) + int(t["A_w"]) - int(t["roulette_w"]) + int(t["killed_w"])
        assert n * fx0 == rhs
    assert abs(slow.total_R()[0] - fast.total_R()[0]) < 0.3 * slow.total_R_se()[0]
    s0, s1 = int(slow.totals[0]["steps"]), int(fast.totals[0]["steps"])
    assert abs(s1 - s0) < 0.005 * s0, (s0, s1)
    print(f"{scene.name}: kernel {slow.kernel_ms:.1f} ms without, {fast.kernel_ms:.1f} ms with jumps")


def test_cli_end_to_end(tmp_path, multipliers):
    """ymc3d geometry.bin optprops.bid definition outname, as the reference's gpu-mc."""
    import os
    import subprocess

    from ymc3d_amd import io

    scene = configs.skin_slab_sphere()
    geo, props = tmp_path / "geometry.bin", tmp_path / "optprops.bid"
    io.write_geometry(str(geo), scene.file_values, *scene.voxel)
    io.write_optprops(str(props), scene.records[0])
    (tmp_path / "def.txt").write_text("photons 50000\nseed 7\nillumination pencil\n")
    exe = os.path.join(os.path.dirname(y._abi.LIB_PATH), "bin", "ymc3d")
    out = subprocess.run([exe, str(geo), str(props), str(tmp_path / "def.txt"), str(tmp_path / "run")],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    drs, d = io.read_property(str(tmp_path / "run_drs.bin"))
    assert drs.shape == (64, 64) and d == pytest.approx((0.01, 0.01))
    assert (tmp_path / "run_simparam.dat").read_text().startswith("50000 #num_photons")
    assert 0.1 < drs.sum() < 0.3


def test_one_shot_run_matches_engine(engine, multipliers):
    """ymc_run (one temporary engine, every parameter explicit) equals the engine API bit for bit,
    absorption grid included, for a batched two-table run with uneven budgets."""
    base = configs.skin_slab_sphere(photons=1)
    mua, mus, g, n = base.optprops()
    props = tuple(np.concatenate([v, v * np.float32(1.5) if q == 0 else v]) for q, v in enumerate((mua, mus, g, n)))
    budgets = np.array([30_000, 7_000], dtype=np.uint64)
    params = {k: v for k, v in base.params.items() if k != "num_photons"}
    params.update(budgets=budgets, flags=y.FLAG_COUNTERS | y.FLAG_ABSORPTION)
    engine.set_geometry(base.rows, *base.voxel)
    engine.set_optprops(*props)
    engine.prepare(**params)
    engine.run()
    want = engine.read_tally()
    want_abs = engine.read_absorption()
    got = y.run_once(base.rows, base.voxel, props, multipliers, **params)
    assert got.status == 0
    for f in ("R_w", "R_w2", "R_n", "T_w", "T_w2", "T_n"):
        assert np.array_equal(getattr(got, f), getattr(want, f)), f
    assert got.totals.tobytes() == want.totals.tobytes()
    assert np.array_equal(got.absorption, want_abs)
    assert list(got.totals["photons"]) == [30_000, 7_000]


def test_16_and_32_bit_voxel_words_agree(engine):
    """Grids whose rows fit 3 bits are stored as 16-bit voxel words (row and safe box packed),
    others as 32-bit words.  The slab-with-inclusion grid with half its slab voxels relabelled to
    an extra row 8 (same optical properties as row 1, so the same physics) takes the 32-bit path.
    With flight jumps off its tallies must equal the 16-bit path's bit for bit.  With jumps on the
    relabelling splits safe boxes (other crossings are jumped, positions differ by rounding, draws
    are the same), so the totals agree far inside one standard error."""
    s = configs.skin_slab_sphere(photons=100_000)
    params = dict(s.params)
    rows = s.rows.copy()
    mua, mus, g, n = s.optprops()
    wide = rows.copy()
    wide[:, :, : wide.shape[2] // 2][wide[:, :, : wide.shape[2] // 2] == 1] = 8
    pad = lambda v: np.concatenate([v, np.repeat(v[:, 1:2], 7, axis=1)], axis=1)  # rows 3..9 copy row 1
    for mode in (2, 1):
        engine.set_geometry(rows, *s.voxel)
        engine.set_optprops(mua, mus, g, n)
        engine.prepare(**params, flags=y.FLAG_COUNTERS, jump_mode=mode)
        engine.run()
        byte = engine.read_tally()
        engine.set_geometry(wide, *s.voxel)
        engine.set_optprops(pad(mua), pad(mus), pad(g), pad(n))
        engine.prepare(**params, flags=y.FLAG_COUNTERS, jump_mode=mode)
        engine.run()
        word = engine.read_tally()
        t0, t1 = byte.totals[0], word.totals[0]
        if mode == 2:
            for f in ("R_w", "R_w2", "R_n", "T_w", "T_w2", "T_n"):
                assert np.array_equal(getattr(byte, f), getattr(word, f)), f
            for f in ("photons", "R_w", "T_w", "A_w", "steps", "fresnel"):
                assert t0[f] == t1[f], f
        else:
            assert t0["photons"] == t1["photons"]
            assert abs(byte.total_R()[0] - word.total_R()[0]) < 0.3 * byte.total_R_se()[0]


def test_fma_probe_stays_below_the_fp32_peak(engine):
    """The FMA-rate probe's figure is physical for every operand form: positive, not falling from 1 to
    8 waves per SIMD (5% slack), and below MI355X's 157.3 TFLOP/s FP32 peak (plus 5% for clock boost)."""
    lib = y.load_library()
    rates = []
    for operands in (1, 2, 3):
        for waves in (1, 8):
            tf, ms = C.c_double(), C.c_double()
            y.engine.check(lib.ymc_probe_fma_rate(0, waves, operands, 2000, C.byref(tf), C.byref(ms)),
                           "ymc_probe_fma_rate")
            assert ms.value > 0
            rates.append(tf.value)
        assert 0 < rates[-2] <= rates[-1] * 1.05
        assert rates[-1] < 157.3 * 1.05
