"""End-to-end reflectance surrogate: Monte Carlo training data, GPU training of an MLP ensemble,
evaluation against Monte Carlo, and serving a hyperspectral cube with Monte Carlo fallback.

    python -m ymc3d_amd.surrogate_pipeline --samples 200000 --photons 100000 --out report.json

Steps:
 1. training data: `samples` parameter vectors (configs.sample_skin2_params), batched Monte Carlo at
    `photons` packets each on the 128x128x64 two-layer-plus-vessel model, radial bins + total R;
 2. training: an ensemble of `members` MLPs 12-256-256-256-33 (GELU), BF16 matrix-core kernels with
    FP32 accumulation and Adam, each member on a bootstrap resample;
 3. evaluation: `test` held-out samples re-simulated at `test_photons` packets; relative error of the
    ensemble mean (linear reflectance) per bin and for the total, the Monte Carlo noise of that
    reference, and the photon count plain Monte Carlo needs for the same median error;
 4. serving: a height x width x 64-wavelength cube (5% out-of-range pixels); queries out of range or
    with ensemble spread above the threshold re-run by Monte Carlo, plus a 1% audit subset re-run to
    measure the served error;
 5. with vs without the surrogate: spectra/s of surrogate serving (fallback included) against plain
    Monte Carlo at matched accuracy.

`--cube 0 0 --save-ensemble e.npz` stops after step 3 and stores the ensemble with its threshold and
the Monte Carlo figures step 5 needs; `--load-ensemble e.npz` then runs steps 4-5 alone.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def rel_err(pred_log, radial, total):
    """|pred - mc| / mc in linear reflectance, for the 32 bins and the total."""
    mc = np.concatenate([radial, total[:, None]], axis=1)
    pred = 10.0 ** pred_log
    return np.abs(pred - mc) / np.maximum(mc, 1e-12), mc


def main(argv=None) -> int:
    # Monte Carlo launches stream through surrogate_data.ENGINES concurrent engines, each on its own
    # hardware queue (HIP's default is 4; read when HIP starts, before the first engine).
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
        os.environ["GPU_MAX_HW_QUEUES"] = "16"
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=200_000)
    ap.add_argument("--photons", type=int, default=100_000)
    ap.add_argument("--members", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--test", type=int, default=1024)
    ap.add_argument("--test-photons", type=int, default=1_000_000)
    ap.add_argument("--cube", type=int, nargs=2, default=(256, 256))
    ap.add_argument("--fallback-photons", type=int, default=1_000_000)
    ap.add_argument("--spread-quantile", type=float, default=0.97)
    ap.add_argument("--out", default="surrogate_report.json")
    ap.add_argument("--save-ensemble", default="", help="write the trained ensemble (.npz) after evaluation")
    ap.add_argument("--load-ensemble", default="", help="serve a saved ensemble: skips data, training and "
                                                        "evaluation")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)

    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import ymc3d_amd as y
    from ymc3d_amd import configs, surrogate as S, surrogate_data as D

    rep = {"config": vars(a)}
    eng = y.Engine(0)
    if a.load_ensemble:
        ens, extra = S.Ensemble.load(a.load_ensemble)
        rep["loaded"] = {k: v.tolist() for k, v in extra.items()}
        threshold = float(extra["threshold"])
        n_star = float(extra["n_star"])
        rep["data"] = {"packets_per_s": float(extra["mc_packets_per_s"])}
        return serve_and_compare(a, rep, eng, ens, threshold, n_star)

    # 1. training data
    P, mc = D.generate(eng, a.samples, a.photons, seed=a.seed, label="train-mc")
    Yt = S.encode_targets(mc["radial"], mc["total"])
    rep["data"] = {"samples": a.samples, "packets": mc["packets"], "wall_s": mc["wall_s"], "kernel_s": mc["kernel_s"],
                   "packets_per_s": mc["packets"] / mc["wall_s"],
                   "steps_per_packet": mc["steps"] / mc["packets"] if mc["steps"] is not None else None,
                   "launches": mc["launches"]}
    print("data", json.dumps(rep["data"]), flush=True)

    # 2. training
    t0 = time.perf_counter()
    ens = S.Ensemble.train(P, Yt, members=a.members, epochs=a.epochs, batch=a.batch, seed=a.seed + 1, verbose=True)
    rep["training"] = {"members": a.members, "epochs": a.epochs, "batch": a.batch, "wall_s": time.perf_counter() - t0,
                       "params_per_member": ens.members[0].n_params}
    print("training", json.dumps(rep["training"]), flush=True)

    # 3. evaluation against higher-count Monte Carlo on held-out samples
    Pt, mct = D.generate(eng, a.test, a.test_photons, seed=a.seed + 100, label="test-mc",
                         tables_per_launch=max(16, a.test // 8), progress_every=0.0)
    mean_t, spread_t = ens.predict(Pt)
    err, ref = rel_err(mean_t, mct["radial"], mct["total"])
    ref_se = np.concatenate([mct["radial_se"], mct["total_se"][:, None]], axis=1) / np.maximum(ref, 1e-12)
    med = float(np.median(err))
    # plain MC needs N* packets for the same median relative error: rel SE scales as 1/sqrt(N)
    s1 = np.median(ref_se) * np.sqrt(a.test_photons)
    n_star = float((s1 / med) ** 2)
    rep["evaluation"] = {"test_samples": a.test, "test_photons": a.test_photons,
                         "median_rel_err_all": med, "median_rel_err_total": float(np.median(err[:, -1])),
                         "median_rel_err_bins_0_15": float(np.median(err[:, :16])),
                         "p90_rel_err_all": float(np.quantile(err, 0.9)),
                         "mc_reference_median_rel_se": float(np.median(ref_se)),
                         "mc_packets_for_same_median_error": n_star}
    threshold = float(np.quantile(spread_t.max(1), a.spread_quantile))
    rep["evaluation"]["spread_threshold_log10"] = threshold
    print("evaluation", json.dumps(rep["evaluation"]), flush=True)
    if a.save_ensemble:
        ens.save(a.save_ensemble, threshold=threshold, n_star=n_star, mc_packets_per_s=rep["data"]["packets_per_s"],
                 median_rel_err=med, median_rel_err_total=rep["evaluation"]["median_rel_err_total"],
                 test_photons=a.test_photons, test_samples=a.test,
                 mc_reference_median_rel_se=rep["evaluation"]["mc_reference_median_rel_se"])
    return serve_and_compare(a, rep, eng, ens, threshold, n_star)


def measure_cube(ens, eng, threshold: float, n_star: float, cube=(256, 256), wavelengths: int = 64,
                 outlier_frac: float = 0.05, fallback_photons: int = 1_000_000, audit_frac: float = 0.01,
                 fallback_sample: int = 256, audit_sample: int = 256, seed: int = 0,
                 tables_per_launch: int = 1024, engines: int = 8) -> dict:
    """Serving a hyperspectral cube with Monte Carlo fallback, measured on the whole cube.

    Every query of the height x width x `wavelengths` cube (`outlier_frac` of the pixels out of range)
    goes through the ensemble (fused passes of 4096 rows) and is triaged (S.triage).  Flagged queries
    and the audit share are re-run by Monte Carlo with `fallback_photons` packets each.  Seeded samples
    of `fallback_sample` flagged and `audit_sample` audit queries run one after the other, each
    streamed through `engines` concurrent engines in launches of `tables_per_launch` queries; each
    sample's time per query, times the number of flagged or audit queries, gives the Monte Carlo time
    of the whole cube.  The audit
    sample gives the surrogate's median relative error against Monte Carlo (with that Monte Carlo's
    own noise, and the error left once that noise is removed in quadrature), and its packet rate gives
    plain Monte Carlo's spectra/s at the surrogate's accuracy (`n_star` packets per query: the count
    at which plain Monte Carlo's median relative error equals the surrogate's).
    The flagged queries' time is an upper bound: some out-of-range queries (g near 0.98, mu_s near
    1800/cm, mu_a near 0.01/cm) have packets that walk 4e6 reference steps, and a launch waits for
    them.  While one engine's launch finishes such walks the other engines' launches fill the device,
    which takes as many hardware queues as engines (GPU_MAX_HW_QUEUES, read when HIP starts; HIP's
    default is 4); the sample still ends on its own last walks, which a whole-cube run of hundreds of
    launches overlaps with other launches."""
    from ymc3d_amd import surrogate as S, surrogate_data as D

    H, W = cube
    Q, outl = D.hyperspectral_cube(H, W, wavelengths, outlier_frac, seed=seed + 200)
    tr = S.triage(ens, Q, threshold, audit_frac, seed=seed + 300)
    flagged, audit = tr["flagged"], tr["audit"]
    rng = np.random.default_rng(seed + 400)

    def sample(mask, k):
        idx = np.flatnonzero(mask)
        return np.sort(rng.choice(idx, size=min(k, idx.size), replace=False)) if idx.size else idx

    fs, aus = sample(flagged, fallback_sample), sample(audit, audit_sample)
    mc_kw = dict(tables_per_launch=tables_per_launch, engines=engines, progress_every=1e9)
    # untimed warm-up launch (first use of the parametric kernel and of the fallback engines)
    D.run_mc(eng, Q[fs[:8]] if fs.size else Q[:8], 10_000, seed=seed + 700, label="warm-up", **mc_kw)
    fb = D.run_mc(eng, Q[fs], fallback_photons, seed=seed + 500, label="fallback-sample", **mc_kw) if fs.size else None
    au = D.run_mc(eng, Q[aus], fallback_photons, seed=seed + 600, label="audit-sample", **mc_kw) if aus.size else None
    out = {"cube": [H, W, wavelengths], "queries": int(len(Q)), "outlier_pixels_frac": float(outl.mean()),
           "surrogate_s": tr["surrogate_s"], "surrogate_queries_per_s": len(Q) / tr["surrogate_s"],
           "flagged": int(flagged.sum()), "flagged_frac": float(flagged.mean()),
           "out_of_range": int((~S.in_range(Q)).sum()), "audit": int(audit.sum()),
           "fallback_photons": fallback_photons, "fallback_sample": int(fs.size), "audit_sample": int(aus.size),
           "mean": tr["mean"], "spread": tr["spread"], "flagged_mask": flagged, "audit_mask": audit,
           "outlier_mask": outl, "Q": Q, "fallback_rows": fs, "audit_rows": aus, "fallback_mc": fb, "audit_mc": au}
    t_fb = fb["wall_s"] / fs.size * flagged.sum() if fb else 0.0
    t_au = au["wall_s"] / aus.size * audit.sum() if au else 0.0
    out["fallback_s_per_query"] = fb["wall_s"] / fs.size if fb else 0.0
    out["audit_s_per_query"] = au["wall_s"] / aus.size if au else 0.0
    out["mc_s_whole_cube"] = t_fb + t_au
    out["total_s"] = tr["surrogate_s"] + t_fb + t_au
    out["spectra_per_s_with_fallback"] = H * W / out["total_s"]
    out["surrogate_only_spectra_per_s"] = H * W / tr["surrogate_s"]
    if au:
        out["mc_packets_per_s"] = au["packets"] / au["wall_s"]
        out["mc_spectra_per_s_at_matched_error"] = out["mc_packets_per_s"] / (wavelengths * n_star)
        out["speedup_with_fallback"] = out["spectra_per_s_with_fallback"] / out["mc_spectra_per_s_at_matched_error"]
    if au:
        err, ref = rel_err(tr["mean"][aus], au["radial"], au["total"])
        ase = np.concatenate([au["radial_se"], au["total_se"][:, None]], axis=1) / np.maximum(ref, 1e-12)
        out["audit_median_rel_err"] = float(np.median(err))
        out["audit_mc_median_rel_se"] = float(np.median(ase))
        out["audit_rel_err_less_mc_noise"] = float(np.sqrt(max(np.median(err) ** 2 - np.median(ase) ** 2, 0.0)))
    return out


def serve_and_compare(a, rep, eng, ens, threshold, n_star) -> int:
    from ymc3d_amd import surrogate as S, surrogate_data as D

    if a.cube[0] <= 0 or a.cube[1] <= 0:
        with open(a.out, "w") as f:
            json.dump(rep, f, indent=1)
        ens.close()
        eng.close()
        return 0

    # 4. serving a hyperspectral cube with Monte Carlo fallback
    H, W = a.cube
    Q, outl = D.hyperspectral_cube(H, W, 64, 0.05, seed=a.seed + 200)
    sv = S.serve(ens, Q, eng, threshold, a.fallback_photons, audit_frac=0.01, seed=a.seed + 300,
                 tables_per_launch=1024, progress_every=0.0, label="fallback-mc")
    qmean, flag, audit, todo, mcq, t_sur = sv["mean"], sv["flagged"], sv["audit"], sv["rows"], sv["mc"], sv["surrogate_s"]
    inr = S.in_range(Q)
    aud_idx = np.flatnonzero(audit[todo])
    aerr, aref = rel_err(qmean[todo][aud_idx], mcq["radial"][aud_idx], mcq["total"][aud_idx])
    # the audit Monte Carlo has noise of its own: its relative SE, and the surrogate error left after
    # removing that noise in quadrature (median over bins)
    ase = np.concatenate([mcq["radial_se"][aud_idx], mcq["total_se"][aud_idx, None]], axis=1) / np.maximum(aref, 1e-12)
    a_net = np.sqrt(np.maximum(np.median(aerr) ** 2 - np.median(ase) ** 2, 0.0))
    t_total = t_sur + mcq["wall_s"]
    spectra = H * W
    rep["serving"] = {"pixels": spectra, "wavelengths": 64, "queries": len(Q), "outlier_pixels_frac": float(outl.mean() if outl.size else 0),
                      "flagged_frac": float(flag.mean()), "flagged_out_of_range": int((~inr).sum()),
                      "audit_queries": int(audit.sum()), "audit_median_rel_err": float(np.median(aerr)),
                      "audit_mc_median_rel_se": float(np.median(ase)), "audit_rel_err_less_mc_noise": float(a_net),
                      "surrogate_s": t_sur, "surrogate_queries_per_s": len(Q) / t_sur,
                      "fallback_queries": int(flag.sum()), "fallback_packets": mcq["packets"],
                      "fallback_s": mcq["wall_s"], "total_s": t_total, "spectra_per_s": spectra / t_total,
                      "surrogate_only_spectra_per_s": spectra / t_sur}
    print("serving", json.dumps(rep["serving"]), flush=True)

    # 5. with vs without the surrogate at matched accuracy
    # in-range Monte Carlo throughput (the training-data run: same model, same parameter distribution)
    mc_rate = rep["data"]["packets_per_s"]
    mc_spectra_per_s = mc_rate / (64 * n_star)
    rep["comparison"] = {"mc_packets_per_s": mc_rate, "mc_packets_per_spectrum_at_matched_error": 64 * n_star,
                         "mc_spectra_per_s_at_matched_error": mc_spectra_per_s,
                         "surrogate_spectra_per_s_with_fallback": rep["serving"]["spectra_per_s"],
                         "speedup_with_fallback": rep["serving"]["spectra_per_s"] / mc_spectra_per_s,
                         "speedup_surrogate_only": rep["serving"]["surrogate_only_spectra_per_s"] / mc_spectra_per_s}
    print("comparison", json.dumps(rep["comparison"]), flush=True)
    with open(a.out, "w") as f:
        json.dump(rep, f, indent=1)
    ens.close()
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
