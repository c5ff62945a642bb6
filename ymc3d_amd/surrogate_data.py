"""Monte Carlo data for the reflectance surrogate: batched runs of the two-layer-plus-vessel skin
model (configs.skin2_*), one parametric geometry and optical-property table per sample, thousands
of samples per persistent launch (each with its own packet budget) and several launches in flight
on separate engines, pencil beam at the surface centre, 32 radial reflectance bins plus total
reflectance per sample."""
from __future__ import annotations

import time

import numpy as np

from . import _abi, configs

# Samples per launch, and engines (each with its own HIP stream) driven concurrently: the last,
# long-path packets of one launch run serially while the other launches' packets fill the device.
# Concurrent launches need a hardware queue each: GPU_MAX_HW_QUEUES (read when HIP starts, default
# 4) at least ENGINES.  On 4096 flagged queries of the 256x256x64 serving cube (1e6 packets each) 8 engines on 16
# queues took 47 s, 6 engines on 4 queues 53-56 s.
TABLES_PER_LAUNCH = 2048
ENGINES = 8


# chain_lanes of the launches.  Packet costs differ by orders of magnitude between samples (weakly
# absorbing samples walk far longer), so chains stay short (the first ones hold 1/32 of a lane's
# share of the budget): a long chain of slow packets would keep its lane busy long after the rest
# of the pool has drained.
CHAIN_LANES = 16 * 458752


def _launch(engine, P, photons, seed, params, counters: bool = False):
    props, geo = configs.skin2_tables(P)
    engine.set_layered(geo, configs.SKIN2_SHAPE, (configs.SKIN2_VOXEL,) * 3)
    engine.set_optprops(*props)
    engine.prepare(**params, budgets=photons, seed=seed, flags=_abi.FLAG_COUNTERS if counters else 0,
                   chain_lanes=CHAIN_LANES)
    ms = engine.run()
    t = engine.read_tally()
    return ms * 1e-3, t


def run_mc(engine, P: np.ndarray, photons, seed: int = 1, tables_per_launch: int = TABLES_PER_LAUNCH,
           engines: int = ENGINES, progress_every: float = 20.0, label: str = "mc", counters: bool = False):
    """Runs every parameter row of P; returns dict of radial [n,32], radial_se, total [n], total_se,
    packets, steps (reference steps, counted only with `counters`, which runs the slower counting
    kernel; None otherwise), summed kernel seconds and wall seconds.  `engine` runs the first share of
    the launches; engines-1 more engines are created for the rest."""
    import concurrent.futures as cf

    from .engine import Engine

    P = np.atleast_2d(P)
    n = P.shape[0]
    photons = np.broadcast_to(np.asarray(photons, dtype=np.uint64), (n,))
    nr = configs.SKIN2_RADIAL[0]
    out = {"radial": np.zeros((n, nr)), "radial_se": np.zeros((n, nr)), "total": np.zeros(n),
           "total_se": np.zeros(n), "packets": int(photons.sum()), "kernel_s": 0.0, "wall_s": 0.0, "steps": 0,
           "launches": 0}
    params = configs.skin2_params(0)
    params.pop("num_photons")
    slices = [slice(b, min(n, b + tables_per_launch)) for b in range(0, n, tables_per_launch)]
    workers = max(1, min(engines, len(slices)))
    pool_engines = [engine] + [Engine(0, engine.multipliers) for _ in range(workers - 1)]
    t0 = time.perf_counter()
    last = t0

    import queue

    free = queue.Queue()
    for e in pool_engines:
        free.put(e)

    def work(sl):
        eng = free.get()
        try:
            return (sl, *_launch(eng, P[sl], photons[sl], seed + sl.start, params, counters))
        finally:
            free.put(eng)

    done = 0
    with cf.ThreadPoolExecutor(workers) as ex:
        futs = [ex.submit(work, sl) for sl in slices]
        for f in cf.as_completed(futs):
            sl, ks, t = f.result()
            out["kernel_s"] += ks
            out["launches"] += 1
            if counters:
                out["steps"] += int(t.totals["steps"].sum())
            out["radial"][sl] = t.radial()
            out["radial_se"][sl] = t.radial_se()
            out["total"][sl] = t.total_R()
            out["total_se"][sl] = t.total_R_se()
            done += sl.stop - sl.start
            now = time.perf_counter()
            if now - last > progress_every:
                print(f"[{label}] {done}/{n} samples, {now - t0:.0f} s", flush=True)
                last = now
    for e in pool_engines[1:]:
        e.close()
    out["wall_s"] = time.perf_counter() - t0
    if not counters:
        out["steps"] = None
    return out


def generate(engine, n_samples: int, photons: int = 100_000, seed: int = 0, outlier_frac: float = 0.0, **kw):
    P = configs.sample_skin2_params(n_samples, seed, outlier_frac)
    return P, run_mc(engine, P, photons, seed=seed * 7919 + 1, **kw)


def hyperspectral_cube(height: int, width: int, wavelengths: int = 64, outlier_frac: float = 0.05, seed: int = 5):
    """Per-pixel, per-wavelength parameter vectors [height*width*wavelengths, 12].

    Each pixel draws base parameters from the training distribution (an outlier pixel, with
    probability outlier_frac, from a range extended by half its width on both sides); absorption
    and reduced scattering then vary smoothly with wavelength (a seeded spectral shape per pixel for
    absorption, a power law for scattering), clipped to the base range for in-range pixels."""
    rng = np.random.default_rng(seed)
    npix = height * width
    base = configs.sample_skin2_params(npix, seed)
    outl = rng.uniform(size=npix) < outlier_frac
    if outl.any():
        base[outl] = configs.sample_skin2_params(int(outl.sum()), seed + 1, outlier_frac=1.0)
    lam = np.linspace(450.0, 850.0, wavelengths)
    centre = rng.uniform(450, 650, npix)
    width_nm = rng.uniform(40, 120, npix)
    shape = 0.4 + 1.6 * np.exp(-0.5 * ((lam[None, :] - centre[:, None]) / width_nm[:, None]) ** 2)  # [npix, L]
    b = rng.uniform(0.5, 1.5, npix)
    power = (lam[None, :] / 650.0) ** (-b[:, None])
    Q = np.repeat(base[:, None, :], wavelengths, axis=1)  # [npix, L, 12]
    for c in range(3):
        Q[:, :, c] = base[:, None, c] * shape
        Q[:, :, 3 + c] = base[:, None, 3 + c] * power
    Q = Q.reshape(-1, 12)
    inr = ~np.repeat(outl, wavelengths)
    Q[inr] = np.clip(Q[inr], configs.PARAM_LO, configs.PARAM_HI)
    return Q, np.repeat(outl, wavelengths)
