#!/usr/bin/env python3
"""Benchmark: photon packets/s of the persistent HIP transport kernel on one MI355X.

Workload (default): layered skin with eight vessels, 256x256x128 voxels of 0.002 cm, uniform
illumination, 1e8 packets per step, 256x256 reflectance grid (configs.skin_vessels).  A step is
one full run of the budget: tally reset, transport kernel, tally read-back to the host.  Inputs
(voxel grid, optical-property tables, stream seeds and chain schedule) are resident on the device
before timing starts; stream seeding is host work done once per workload and is not timed.

Prints one JSON line with `roofline` (FP32 VALU work of the transport kernel over the MI355X FP32
peak, kernel time from HIP events on the engine's stream) and `cpu_baseline` (the CPU oracle on a
bounded sample of the same workload, all host threads).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# FP32 operations per photon step and per Fresnel event on the reference's common path
# (src/mc3d_gpu.cu:181-339 and :226-257): add/sub/mul/div/sqrt/rsqrt/log/sin/cos count 1,
# a*b+c counts 2, compares and selects 0.
FLOPS_PER_STEP = 108
FLOPS_PER_FRESNEL = 31
PEAK_FP32_TFLOPS = 157.3

# HBM traffic per launch of the timed kernel, from rocprofv3 FETCH_SIZE and WRITE_SIZE runs of this
# script's default workload, scaled by the same counters read on a 1 GiB device copy (x2.000 for
# FETCH_SIZE, x1.000 for WRITE_SIZE); profiles/transport_skin_vessels_256x256x128.md.  The launch
# is deterministic (same chains and streams every step), so the figure holds for every default run.
MEASURED_TRAFFIC = {("skin_vessels", 100_000_000, 458752): 7.385e11}

WORKLOADS = {
    "skin_vessels": ("skin_vessels", 100_000_000),
    "skin_slab_sphere": ("skin_slab_sphere", 1_000_000),
    "batched_spectra": ("batched_spectra", 0),
    "surrogate": ("surrogate", 0),
    "skin2_batch": ("skin2_batch", 0),
}
PEAK_BF16_TFLOPS = 2500.0  # dense MFMA


def main() -> int:
    # The surrogate-serving measurement streams Monte Carlo launches through 8 concurrent engines;
    # each needs its own hardware queue (HIP's default is 4).  Read when HIP starts, so set first.
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
        os.environ["GPU_MAX_HW_QUEUES"] = "16"
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="skin_vessels", choices=sorted(WORKLOADS))
    ap.add_argument("--photons", type=int, default=0, help="override the workload's packet budget")
    ap.add_argument("--pool-lanes", type=int, default=0)
    ap.add_argument("--chain-lanes", type=int, default=0)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="default workload only: skip the secondary workloads (slab with inclusion, batched "
                         "spectra) and the surrogate serving measurement")
    ap.add_argument("--absorption", action="store_true", help="also tally the per-voxel absorption grid")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="instead of the workload, time a plain device-to-device copy of this many bytes (the "
                         "reference point that scales FETCH_SIZE / WRITE_SIZE counter readings to bytes)")
    ap.add_argument("--atomic-rate", action="store_true",
                    help="instead of the workload, time 2^27 scattered 64-bit atomic adds into 2^23 counters (the "
                         "ceiling of the absorption grid's deposits)")
    ap.add_argument("--fma-rate", action="store_true",
                    help="instead of the workload, time plain FP32 FMA chains at 1..8 waves per SIMD reading 1..3 "
                         "VGPRs (what unpacked v_fma_f32 reaches against the 157.3 TFLOP/s peak)")
    args = ap.parse_args()
    if args.copy_bytes:
        return copy_reference(args.copy_bytes, max(args.steps, 1))
    if args.atomic_rate:
        return atomic_reference(max(args.steps, 1))
    if args.fma_rate:
        return fma_reference(max(args.steps, 1))

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the photon budget shards across GPUs with one integer all-reduce of "
              "the fixed-point tallies, which this engine does not run; use --gpus 1", file=sys.stderr)
        return 2

    if args.workload == "surrogate":
        return surrogate_bench(args)
    if args.workload == "skin2_batch":
        return parametric_bench(args)
    r = transport(args.workload, args.photons, args, args.steps, args.warmup, absorption=args.absorption,
                  pool_lanes=args.pool_lanes, chain_lanes=args.chain_lanes)
    out = {
        "metric": "photon packets/s (1 GPU)",
        "value": r["value"],
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": r["config"],
        "roofline": r["roofline"],
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(r["scene"], r["tab"], r["params"], r["photons"], args.cpu_seconds)
    default_run = args.workload == "skin_vessels" and not (args.photons or args.pool_lanes or args.chain_lanes or
                                                           args.absorption)
    if default_run and not args.no_secondary:
        # the other transport workloads, each with its CPU baseline on a bounded sample
        out["secondary"] = []
        for name in ("skin_slab_sphere", "batched_spectra"):
            s = transport(name, 0, args, args.steps, args.warmup)
            entry = {"workload": s["config"]["workload"], "value": s["value"], "unit": "photons/s",
                     "ms_per_step": s["ms_per_step"], "kernel_ms": s["config"]["kernel_ms"],
                     "photons_per_step": s["photons"], "tables": s["config"]["tables"],
                     "steps_per_photon": s["config"]["steps_per_photon"], "roofline_frac": s["roofline"]["frac"],
                     "spectra_per_s": s["config"]["spectra_per_s"], "launch_params": s["config"]["launch_params"],
                     "cpu_baseline": None}
            if not args.no_cpu:
                cb = cpu_baseline(s["scene"], s["tab"], s["params"], s["photons"], args.cpu_seconds / 2)
                entry["cpu_baseline"] = cb
                entry["vs_cpu"] = s["value"] / cb["value"]
            out["secondary"].append(entry)
        # the parametric kernel (the surrogate's Monte Carlo) alone; no CPU baseline (the oracle runs
        # voxel grids, and the serving line below compares with plain Monte Carlo instead)
        pr = parametric(args.steps, args.warmup)
        out["secondary"].append({"workload": pr["config"]["workload"], "value": pr["value"], "unit": "photons/s",
                                 "ms_per_step": pr["ms_per_step"], "kernel_ms": pr["config"]["kernel_ms"],
                                 "photons_per_step": pr["config"]["photons_per_step"],
                                 "tables": pr["config"]["tables"],
                                 "steps_per_photon": pr["config"]["steps_per_photon"],
                                 "roofline_frac": pr["roofline"]["frac"], "cpu_baseline": None})
        out["surrogate_serving"] = surrogate_serving()
    print(json.dumps(out), flush=True)
    return 0


def transport(workload, photons, args, steps, warmup, absorption=False, pool_lanes=0, chain_lanes=0) -> dict:
    """One transport workload: a counter pass (steps and Fresnel events for the roofline numerator),
    `warmup` untimed and `steps` timed runs of the whole budget (tally reset, persistent kernel,
    tally read-back), inputs resident on the device."""
    import numpy as np

    import ymc3d_amd as y
    from ymc3d_amd import autotune, configs

    fn, default_n = WORKLOADS[workload]
    if fn == "batched_spectra":
        scene = configs.batched_spectra()
        n_photons = int(np.sum(scene.params["budgets"]))
    else:
        n_photons = photons or default_n
        scene = getattr(configs, fn)(photons=n_photons)
    params = dict(scene.params)
    params.update(autotune.tuned(scene, scene.params))
    if pool_lanes:
        params["pool_lanes"] = pool_lanes
    if chain_lanes:
        params["chain_lanes"] = chain_lanes
    tab = y.default_multipliers()
    eng = y.Engine(0, tab)
    scene.load(eng)

    # Counter pass (deterministic: the same chains and streams as every timed step).
    extra = y.FLAG_ABSORPTION if absorption else 0
    eng.prepare(**params, flags=y.FLAG_COUNTERS | extra)
    eng.run()
    counted = eng.read_tally()
    n_steps = int(counted.totals["steps"].sum())
    fresnel = int(counted.totals["fresnel"].sum())
    tables = len(counted.totals)

    eng.prepare(**params, flags=extra)
    for _ in range(warmup):
        eng.run()
        eng.read_tally()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.run()
        res = eng.read_tally()
        kernel_ms.append(res.kernel_ms)
    elapsed = time.perf_counter() - t0
    assert int(res.totals["photons"].sum()) == n_photons
    # every step runs the same chains and streams: its maps equal the counter pass's bit for bit
    assert np.array_equal(res.R_w, counted.R_w) and np.array_equal(res.T_w, counted.T_w)

    kms = float(np.mean(kernel_ms))
    achieved = (n_steps * FLOPS_PER_STEP + fresnel * FLOPS_PER_FRESNEL) / (kms * 1e-3) / 1e12
    pool = eng.pool_lanes()
    eng.close()
    return {
        "value": steps * n_photons / elapsed,
        "ms_per_step": elapsed / steps * 1e3,
        "photons": n_photons,
        "scene": scene,
        "params": params,
        "tab": tab,
        "config": {
            "workload": scene.name,
            "photons_per_step": n_photons,
            "illumination": "uniform" if params.get("illumination", 0) == 0 else "pencil",
            "voxels": list(scene.shape),
            "voxel_cm": scene.voxel[0],
            "tally": [params.get("tally_nx", 0) or scene.shape[0], params.get("tally_ny", 0) or scene.shape[1]],
            "pool_lanes": pool,
            "tables": tables,
            "spectra_per_s": steps / elapsed if tables > 1 else None,
            "absorption_grid": bool(absorption),
            "launch_params": {k: params[k] for k in ("pool_lanes", "jump_max_crossings", "relaunch_lanes",
                                                      "relaunch_wait", "chain_lanes") if k in params},
            "steps_per_photon": n_steps / n_photons,
            "fresnel_per_photon": fresnel / n_photons,
            "kernel_ms": kms,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": None if absorption else MEASURED_TRAFFIC.get((workload, n_photons, pool)),
        },
    }


def parametric_bench(args) -> int:
    print(json.dumps(parametric(args.steps, args.warmup)), flush=True)
    return 0


def parametric(steps: int, warmup: int) -> dict:
    """The surrogate's Monte Carlo: the parametric-geometry kernel (two layers and a vessel computed per
    voxel on the device, pencil beam, 32 radial bins) on 512 in-range parameter rows of the
    128x128x64 skin model, one table each, 2e5 packets per table (1.02e8 per step) in one persistent
    launch.  A counter launch gives steps and Fresnel events for the roofline numerator; `warmup`
    untimed and `steps` timed launches follow, inputs resident."""
    import numpy as np

    import ymc3d_amd as y
    from ymc3d_amd import configs, surrogate_data as D

    tables, per_table = 512, 200_000
    P = configs.sample_skin2_params(tables, seed=7)
    props, geo = configs.skin2_tables(P)
    params = configs.skin2_params(0)
    params.pop("num_photons")
    params.update(budgets=np.full(tables, per_table, np.uint64), seed=11, chain_lanes=D.CHAIN_LANES)
    eng = y.Engine(0, y.default_multipliers())
    eng.set_layered(geo, configs.SKIN2_SHAPE, (configs.SKIN2_VOXEL,) * 3)
    eng.set_optprops(*props)
    eng.prepare(**params, flags=y.FLAG_COUNTERS)
    eng.run()
    counted = eng.read_tally()
    n_steps, fresnel = int(counted.totals["steps"].sum()), int(counted.totals["fresnel"].sum())
    n_photons = tables * per_table
    eng.prepare(**params)
    for _ in range(warmup):
        eng.run()
        eng.read_tally()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.run()
        res = eng.read_tally()
        kernel_ms.append(res.kernel_ms)
    elapsed = time.perf_counter() - t0
    assert int(res.totals["photons"].sum()) == n_photons
    assert np.array_equal(res.Rr_w, counted.Rr_w)  # same chains and streams: the counter pass's tallies
    kms = float(np.mean(kernel_ms))
    achieved = (n_steps * FLOPS_PER_STEP + fresnel * FLOPS_PER_FRESNEL) / (kms * 1e-3) / 1e12
    out = {
        "metric": "photon packets/s (1 GPU)",
        "value": steps * n_photons / elapsed,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"skin2_batch_{tables}x{per_table}", "tables": tables,
                   "photons_per_step": n_photons, "voxels": list(configs.SKIN2_SHAPE),
                   "voxel_cm": configs.SKIN2_VOXEL, "pool_lanes": eng.pool_lanes(),
                   "steps_per_photon": n_steps / n_photons, "fresnel_per_photon": fresnel / n_photons,
                   "kernel_ms": kms},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None},
        "cpu_baseline": None,
    }
    eng.close()
    return out


def surrogate_serving() -> dict:
    """Surrogate serving of the 256 x 256-pixel x 64-wavelength hyperspectral cube with the shipped
    ensemble: every query through the ensemble (fused passes of 4096), 1e6-packet Monte Carlo
    fallback for flagged queries and a 1% audit; the Monte Carlo runs on seeded samples of 4096 flagged
    and 1024 audit queries, each streamed through 8 concurrent engines (16 hardware queues) in launches
    of 512 queries, and its time for the whole cube is scaled from them (surrogate_pipeline.measure_cube).  Plain Monte Carlo at matched accuracy: the audit sample's
    packet rate over 64 x n_star packets per spectrum."""
    import ymc3d_amd as y
    from ymc3d_amd import surrogate as S, surrogate_pipeline as SP

    if not os.path.exists(S.DEFAULT_ENSEMBLE):
        return {"error": f"ensemble file {S.DEFAULT_ENSEMBLE} missing"}
    ens, extra = S.Ensemble.load(S.DEFAULT_ENSEMBLE)
    eng = y.Engine(0)
    try:
        r = SP.measure_cube(ens, eng, float(extra["threshold"]), float(extra["n_star"]), cube=(256, 256),
                            fallback_sample=4096, audit_sample=1024, seed=0, tables_per_launch=512, engines=8)
    finally:
        ens.close()
        eng.close()
    keep = ("cube", "queries", "outlier_pixels_frac", "flagged_frac", "out_of_range", "audit", "fallback_photons",
            "fallback_sample", "audit_sample", "surrogate_s", "surrogate_queries_per_s", "fallback_s_per_query",
            "audit_s_per_query",
            "mc_s_whole_cube", "total_s", "spectra_per_s_with_fallback", "surrogate_only_spectra_per_s",
            "audit_median_rel_err", "audit_mc_median_rel_se", "audit_rel_err_less_mc_noise", "mc_packets_per_s",
            "mc_spectra_per_s_at_matched_error", "speedup_with_fallback")
    out = {"workload": "surrogate_serving_256x256x64", "unit": "spectra/s", "ensemble": os.path.basename(S.DEFAULT_ENSEMBLE)}
    out.update({k: (r[k] if not hasattr(r[k], "tolist") else r[k].tolist()) for k in keep if k in r})
    out["value"] = out.get("spectra_per_s_with_fallback")
    for k in ("median_rel_err", "median_rel_err_total", "test_photons", "n_star"):
        if k in extra:
            out["eval_" + k] = float(extra[k])
    out["hw_queues"] = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
    out["note"] = ("Monte Carlo time of the whole cube scaled from the per-query times of the seeded fallback "
                   "(4096 queries) and audit (1024) samples, each streamed through 8 concurrent engines in launches "
                   "of 512; the fallback time is an upper bound (the sample ends on the 4e6-step walks of some "
                   "out-of-range queries); eval_* from the held-out evaluation stored with the ensemble")
    return out


def surrogate_bench(args) -> int:
    """Reflectance-surrogate serving: a 4-member ensemble of 12-256-256-256-33 GELU networks
    (random initialisation), batches of 4096 parameter vectors resident on the device, one fused
    launch per batch for all members; a spectrum is 64 wavelengths.  Also one training epoch (Adam, batch 512) over
    65536 synthetic samples.  Roofline: BF16 MFMA FLOPs of the products over the kernel time."""
    import numpy as np

    from ymc3d_amd import surrogate as S

    members, rows, reps = 4, 4096, max(50, args.steps * 50)
    nets = [S.MLP(seed=m) for m in range(members)]
    flops_row = 2 * sum(a * b for a, b in zip(S.WIDTHS[:-1], S.WIDTHS[1:]))
    for _ in range(args.warmup):
        S.time_forward(nets, rows, 10)
    t0 = time.perf_counter()
    ms_pass = S.time_forward(nets, rows, reps)  # one fused launch: every member, all rows
    wall = time.perf_counter() - t0
    queries_per_s = rows / (ms_pass * 1e-3)
    achieved = members * flops_row * rows / (ms_pass * 1e-3) / 1e12
    # training epoch
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (65536, S.WIDTHS[0])).astype(np.float32)
    yv = rng.normal(size=(65536, S.WIDTHS[-1])).astype(np.float32)
    nets[0].fit(x[:4096], yv[:4096], 1, 512)
    t0 = time.perf_counter()
    nets[0].fit(x, yv, 1, 512)
    t_train = time.perf_counter() - t0
    for n in nets:
        n.close()
    out = {
        "metric": "surrogate reflectance spectra/s (1 GPU)",
        "value": queries_per_s / 64,
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": args.warmup,
        "ms_per_step": ms_pass,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": "surrogate_ensemble_4x_12-256-256-256-33", "rows_per_pass": rows,
                   "members": members, "queries_per_s": queries_per_s, "ms_per_ensemble_pass": ms_pass,
                   "timed_wall_s": wall, "training_samples_per_s": 65536 / t_train,
                   "training_tflops": 3 * flops_row * 65536 / t_train / 1e12},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_BF16_TFLOPS, "traffic": None},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        from oracle import mlp_oracle as R

        p = S.MLP(seed=0)
        params = p.get_params().astype(np.float64)
        p.close()
        xq = rng.uniform(-1, 1, (4096, S.WIDTHS[0]))
        t0 = time.perf_counter()
        done = 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10.0):
            R.forward(params, S.WIDTHS, xq)
            done += 4096
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done / members / dt / 64, "unit": "spectra/s",
                               "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count())),
                               "kind": "port", "sample": f"{done} rows through one float64 numpy network, {dt:.1f} s, "
                                                         f"scaled to {members} members", **host_cpu()}
    print(json.dumps(out), flush=True)
    return 0


def copy_reference(nbytes: int, reps: int) -> int:
    """Device-to-device copy of nbytes (one read + one write of nbytes per launch), reps launches."""
    import torch

    src = torch.empty(nbytes, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes": nbytes, "launches": reps + 2, "ms_per_copy": dt * 1e3,
                      "GBps_read_plus_write": 2 * nbytes / dt / 1e9}), flush=True)
    return 0


def atomic_reference(reps: int) -> int:
    """Random 64-bit atomic adds: torch's index_add_ on int64 issues one device atomic per index.  A
    64 MiB grid of counters, like the absorption grid of the 256x256x128 vessel workload, and uniform
    random indices."""
    import torch

    nvox, n = 1 << 23, 1 << 27
    gen = torch.Generator(device="cuda").manual_seed(1)
    idx = torch.randint(0, nvox, (n,), device="cuda", dtype=torch.int64, generator=gen)
    ones = torch.ones(n, dtype=torch.int64, device="cuda")
    grid = torch.zeros(nvox, dtype=torch.int64, device="cuda")
    grid.index_add_(0, idx, ones)  # warm-up
    best = float("inf")
    for _ in range(reps):
        grid.zero_()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        grid.index_add_(0, idx, ones)
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b))
    if int(grid.sum()) != n:
        print("atomic reference: counter sum mismatch", file=sys.stderr)
        return 1
    print(json.dumps({"atomic_adds": n, "counters": nvox, "ms": best, "atomics_per_s": n / (best * 1e-3)}), flush=True)
    return 0


def fma_reference(reps: int) -> int:
    """Independent FP32 FMA chains, eight per lane, on every SIMD (library probe, HIP events), for
    FMAs reading 1, 2 and 3 distinct VGPRs."""
    import ctypes as C

    from ymc3d_amd.engine import check, load_library

    lib = load_library()
    rows = []
    for operands in (1, 2, 3):
        for waves in (1, 2, 4, 8):
            best_t, best_ms = 0.0, 0.0
            for _ in range(reps):
                tf, ms = C.c_double(), C.c_double()
                check(lib.ymc_probe_fma_rate(0, waves, operands, 20000, C.byref(tf), C.byref(ms)),
                      "ymc_probe_fma_rate")
                if tf.value > best_t:
                    best_t, best_ms = tf.value, ms.value
            rows.append({"vgpr_operands": operands, "waves_per_simd": waves, "tflops": best_t, "ms": best_ms,
                         "frac_of_peak": best_t / PEAK_FP32_TFLOPS})
    print(json.dumps({"fma_rate": rows}), flush=True)
    return 0


def cpu_baseline(scene, tab, params, n_photons, seconds):
    """CPU oracle (all host threads) on a bounded sample of the same workload."""
    from oracle import oracle as O

    import numpy as np

    threads = O.lib().ymc_oracle_max_threads()

    def run(count):
        if "budgets" in scene.params:  # batched: every table scaled to the sample size
            b = np.asarray(scene.params["budgets"], dtype=np.float64)
            budgets = np.maximum(1, np.round(b * count / b.sum())).astype(np.uint64)
            O.run_scene(scene, tab, budgets=budgets)
            return int(budgets.sum())
        O.run_scene(scene, tab, num_photons=count)
        return count

    probe = 20_000
    t0 = time.perf_counter()
    probe = run(probe)
    rate = probe / (time.perf_counter() - t0)
    sample = int(min(n_photons, max(probe, rate * seconds)))
    t0 = time.perf_counter()
    sample = run(sample)
    dt = time.perf_counter() - t0
    return {"value": sample / dt, "unit": "photons/s", "cores": threads, "kind": "port",
            "sample": f"{sample} packets of {scene.name} (same geometry, tables, schedule rule), {dt:.1f} s",
            **host_cpu()}


def host_cpu() -> dict:
    """The host the CPU baseline ran on: /proc/cpuinfo model name, logical CPUs online (nproc) and the
    CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))}


if __name__ == "__main__":
    sys.exit(main())
