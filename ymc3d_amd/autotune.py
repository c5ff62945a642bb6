"""Autotuner for the persistent transport launch.

Searches, on a sample budget of the actual workload:
  pool_lanes        resident lanes (workgroups x 256), up to full occupancy
  jump_max_crossings  most plain crossings one safe-box flight jump may take
  relaunch_lanes    a lane whose packet ended waits until this many lanes of its wave wait, then
                    the wave tallies and relaunches them together (the phase costs the wave as much
                    for one lane as for many; waiting lanes do nothing)
  chain_lanes       the compaction interval of the persistent pool.  A lane runs a chain's packets
                    back to back, then returns to the work queue for the next chain; the guided
                    schedule gives the first chains ceil(N / (2 chain_lanes)) packets and shrinks
                    them to single packets towards the end.  Relative to the per-lane share of the
                    budget, N / pool, the first chains hold pool / (2 chain_lanes) shares whatever
                    N is, so chain_lanes is searched as a multiple of the full-occupancy pool: short
                    chains keep lanes busy when packet costs vary between tables, long chains cost
                    fewer queue visits and stream seedings.
pool_lanes and relaunch_lanes do not change results (tallies are bit-identical for any value,
which the GPU tests check for the pool); jump_max_crossings changes only which crossings are jumped, i.e. float rounding of the
positions, with the same draws; chain_lanes decides which stream each packet draws from, so results change at
the statistical level and stay reproducible for the stored setting.  The best setting goes to a
JSON file keyed by device name and workload signature (which includes the decade of the packet budget);
engine runs pick it up through `tuned(...)`.
"""
from __future__ import annotations

import hashlib
import itertools
import json
import os
import time

import numpy as np

CACHE = os.environ.get("YMC3D_TUNE_CACHE", os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                                          "autotune.json"))
SEARCH = {
    "pool_frac": (1.0, 0.5),
    "jump_max_crossings": (16, 64),
    "relaunch_lanes": (8, 12, 16),
    "chain_factor": (1, 4, 16),
}


def budget_decade(params: dict) -> int:
    """log10 of the packet budget, rounded: the best pool and chains depend on it (a 1e6-packet
    run is bound by its longest packets, a 1e8-packet run by throughput)."""
    total = float(np.sum(params["budgets"])) if "budgets" in params else float(params.get("num_photons", 0))
    return int(np.floor(np.log10(total) + 0.5)) if total > 0 else 0


def signature(scene, params: dict) -> str:
    """Workload key: geometry shape and voxel, table count, illumination, tally layout and the decade
    of the packet budget."""
    h = hashlib.sha1()
    h.update(np.ascontiguousarray(scene.file_values).tobytes())
    h.update(np.ascontiguousarray(scene.records, dtype=np.float64).tobytes())
    for k in ("illumination", "tally_nx", "tally_ny", "radial_bins"):
        h.update(f"{k}={params.get(k, 0)};".encode())
    h.update(f"budget_decade={budget_decade(params)};".encode())
    return h.hexdigest()[:16]


def device_name() -> str:
    """The GPU's marketing name when rocminfo reports one (e.g. "AMD Instinct MI355X"), else its
    architecture (e.g. "gfx950": some hosts leave the marketing name empty); "gpu" without rocminfo."""
    try:
        import subprocess

        out = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=30).stdout
        arch = ""
        for line in out.splitlines():
            key, _, value = line.strip().partition(":")
            value = value.strip()
            if key == "Marketing Name" and "Instinct" in value:
                return value
            if key == "Name" and value.startswith("gfx") and not arch:
                arch = value
        if arch:
            return arch
    except Exception:
        pass
    return "gpu"


def _load() -> dict:
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            return json.load(f)
    return {}


def tuned(scene, params: dict) -> dict:
    """Launch parameters stored for this workload (empty if never tuned)."""
    return dict(_load().get(device_name(), {}).get(signature(scene, params), {}).get("params", {}))


def tune(engine, scene, sample_photons: int | None = None, repeats: int = 2, verbose: bool = True) -> dict:
    """Times every combination of SEARCH on a sample budget; stores and returns the best."""
    params = dict(scene.params)
    if sample_photons:
        if "budgets" in params:
            b = np.asarray(params["budgets"], dtype=np.float64)
            params["budgets"] = np.maximum(1, b * sample_photons / b.sum()).astype(np.uint64)
        else:
            params["num_photons"] = sample_photons
    scene.load(engine)
    engine.prepare(**params)
    engine.run()
    full = engine.pool_lanes()
    results = []
    for pf, jc, rl, cf in itertools.product(*SEARCH.values()):
        cand = {"pool_lanes": max(256, int(full * pf) // 256 * 256), "jump_max_crossings": jc,
                "relaunch_lanes": rl, "chain_lanes": int(full * cf)}
        engine.prepare(**params, **cand)
        ms = min(engine.run() for _ in range(repeats))
        results.append((ms, cand))
        if verbose:
            print(json.dumps({"kernel_ms": round(ms, 3), **cand}), flush=True)
    results.sort(key=lambda r: r[0])
    best_ms, best = results[0]
    db = _load()
    entry = db.setdefault(device_name(), {})
    entry[signature(scene, scene.params)] = {"params": best, "kernel_ms": best_ms, "scene": scene.name,
                                             "sample_photons": sample_photons, "tuned_at": time.time(),
                                             "worst_ms": results[-1][0]}
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    with open(CACHE + ".tmp", "w") as f:
        json.dump(db, f, indent=1)
    os.replace(CACHE + ".tmp", CACHE)
    return best


if __name__ == "__main__":
    import argparse
    import sys

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import ymc3d_amd as y
    from ymc3d_amd import configs

    ap = argparse.ArgumentParser(description="tune launch parameters for a workload")
    ap.add_argument("workload", choices=["skin_vessels", "skin_slab_sphere", "batched_spectra"])
    ap.add_argument("--photons", type=int, default=0, help="budget of the tuned scene (0: the scene default)")
    ap.add_argument("--sample", type=int, default=0, help="packets per timed run (0: the full budget; tail "
                                                             "effects depend on it)")
    a = ap.parse_args()
    eng = y.Engine(0)
    sc = getattr(configs, a.workload)(photons=a.photons) if a.photons else getattr(configs, a.workload)()
    print("best:", json.dumps(tune(eng, sc, a.sample)))
