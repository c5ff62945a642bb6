"""Reflectance surrogate: an ensemble of MLPs 12 -> 256 -> 256 -> 256 -> 33 (GELU) mapping the
twelve skin parameters of configs.sample_skin2_params to log10 of 32 radial reflectance bins and of
the total diffuse reflectance.  Training and inference run on the device (libymc3d.so, BF16
matrix-core kernels with FP32 accumulation, Adam on FP32 master weights).  The spread of the
ensemble members is the error estimate; queries above threshold, or outside the training ranges,
fall back to Monte Carlo.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import configs
from ._abi import ptr
from .engine import check, load_library

WIDTHS = (12, 256, 256, 256, 33)
LOG_FLOOR = 1e-7  # reflectance per packet below this is treated as this when taking log10

# Trained ensemble shipped with the package (Ensemble.save format, with `threshold`, `n_star` and
# `mc_packets_per_s`): 4 members trained on 2e5 samples x 1e5 packets of the two-layer-plus-vessel
# model by `python -m ymc3d_amd.surrogate_pipeline --samples 200000 --photons 100000 --members 4
# --epochs 40 --test 512 --test-photons 4000000 --cube 0 0 --save-ensemble <path>`.
DEFAULT_ENSEMBLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "surrogate_skin2_ensemble.npz")


class MLP:
    """One network on the device (ymc_mlp)."""

    def __init__(self, widths=WIDTHS, seed: int = 0, device: int = 0):
        self._lib = load_library()
        self.widths = tuple(int(w) for w in widths)
        w = np.array(self.widths, np.int32)
        self._h = C.c_void_p()
        check(self._lib.ymc_mlp_create(device, ptr(w, C.c_int32), w.size, seed, C.byref(self._h)), "ymc_mlp_create")
        self.n_params = int(self._lib.ymc_mlp_num_params(self._h))

    def close(self):
        if self._h:
            self._lib.ymc_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        check(self._lib.ymc_mlp_get_params(self._h, ptr(p, C.c_float), p.size), "ymc_mlp_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check(self._lib.ymc_mlp_set_params(self._h, ptr(p, C.c_float), p.size), "ymc_mlp_set_params")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((x.shape[0], self.widths[-1]), np.float32)
        check(self._lib.ymc_mlp_forward(self._h, ptr(x, C.c_float), x.shape[0], ptr(y, C.c_float)), "ymc_mlp_forward")
        return y

    def gradients(self, x: np.ndarray, y: np.ndarray) -> tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        check(self._lib.ymc_mlp_gradients(self._h, ptr(x, C.c_float), ptr(y, C.c_float), x.shape[0],
                                          ptr(g, C.c_float), C.byref(loss)), "ymc_mlp_gradients")
        return loss.value, g

    def fit(self, x: np.ndarray, y: np.ndarray, epochs: int, batch: int = 512, lr: float = 2e-3,
            lr_final: float = 2e-5, seed: int = 0) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        check(self._lib.ymc_mlp_set_data(self._h, ptr(x, C.c_float), ptr(y, C.c_float), x.shape[0]),
              "ymc_mlp_set_data")
        losses = np.zeros(epochs, np.float32)
        check(self._lib.ymc_mlp_train(self._h, epochs, batch, lr, lr_final, seed, ptr(losses, C.c_float)),
              "ymc_mlp_train")
        return losses


def _handles(nets):
    h = (C.c_void_p * len(nets))(*[n._h.value for n in nets])
    return C.cast(h, C.POINTER(C.c_void_p)), h


def forward_ensemble(nets, x: np.ndarray) -> np.ndarray:
    """Outputs of every network for the same inputs, [len(nets), rows, out], in one fused launch per
    4096 rows (ymc_mlp_forward_ensemble)."""
    lib = load_library()
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty((len(nets), x.shape[0], nets[0].widths[-1]), np.float32)
    hp, keep = _handles(nets)
    check(lib.ymc_mlp_forward_ensemble(hp, len(nets), ptr(x, C.c_float), x.shape[0], ptr(y, C.c_float)),
          "ymc_mlp_forward_ensemble")
    return y


def time_forward(nets, rows: int, reps: int) -> float:
    """Device milliseconds per fused ensemble pass of `rows` device-resident rows."""
    lib = load_library()
    ms = C.c_double()
    hp, keep = _handles(nets)
    check(lib.ymc_mlp_time_forward(hp, len(nets), rows, reps, C.byref(ms)), "ymc_mlp_time_forward")
    return ms.value


def encode_params(P: np.ndarray) -> np.ndarray:
    """Parameters to network inputs in [-1, 1] over the training ranges (log scale where sampled so)."""
    P = np.atleast_2d(np.asarray(P, np.float64))
    lo, hi = configs.PARAM_LO, configs.PARAM_HI
    v = np.where(configs.PARAM_LOG, np.log(np.maximum(P, 1e-30)), P)
    a = np.where(configs.PARAM_LOG, np.log(lo), lo)
    b = np.where(configs.PARAM_LOG, np.log(hi), hi)
    return (2.0 * (v - a) / (b - a) - 1.0).astype(np.float32)


def in_range(P: np.ndarray) -> np.ndarray:
    P = np.atleast_2d(P)
    return np.all((P >= configs.PARAM_LO * (1 - 1e-6)) & (P <= configs.PARAM_HI * (1 + 1e-6)), axis=1)


def encode_targets(radial: np.ndarray, total: np.ndarray) -> np.ndarray:
    """log10 of the 32 radial bins and the total reflectance (per packet)."""
    return np.log10(np.maximum(np.concatenate([radial, total[:, None]], axis=1), LOG_FLOOR)).astype(np.float32)


@dataclass
class Ensemble:
    members: list = field(default_factory=list)
    y_mean: np.ndarray | None = None
    y_std: np.ndarray | None = None

    @classmethod
    def train(cls, P: np.ndarray, targets: np.ndarray, members: int = 4, epochs: int = 60, batch: int = 512,
              lr: float = 2e-3, seed: int = 0, verbose: bool = False):
        x = encode_params(P)
        ens = cls()
        ens.y_mean = targets.mean(0)
        ens.y_std = targets.std(0) + 1e-6
        yz = ((targets - ens.y_mean) / ens.y_std).astype(np.float32)
        def fit(m):
            net = MLP(seed=seed * 1000 + m)
            # each member sees its own bootstrap resample of the data
            idx = np.random.default_rng(seed * 1000 + m).integers(0, len(x), len(x))
            return net, net.fit(x[idx], yz[idx], epochs, batch, lr, lr * 0.01, seed * 1000 + m)

        # Members train concurrently, each from its own host thread on its own stream: a training step
        # is a chain of small kernels, and one thread's launches leave the device mostly idle.  Each
        # member's result does not depend on the others (identical parameters either way).
        import concurrent.futures as cf

        with cf.ThreadPoolExecutor(max(1, members)) as ex:
            fitted = list(ex.map(fit, range(members)))
        for m, (net, losses) in enumerate(fitted):
            if verbose:
                print(f"member {m}: loss {losses[0]:.4f} -> {losses[-1]:.5f}", flush=True)
            ens.members.append(net)
        return ens

    def predict(self, P: np.ndarray, chunk: int = 1 << 18) -> tuple[np.ndarray, np.ndarray]:
        """Mean prediction (log10 targets) and the member standard deviation (log10 units)."""
        P = np.atleast_2d(P)
        mean = np.empty((P.shape[0], WIDTHS[-1]), np.float32)
        spread = np.empty_like(mean)
        for r0 in range(0, P.shape[0], chunk):
            x = encode_params(P[r0:r0 + chunk])
            outs = forward_ensemble(self.members, x) * self.y_std + self.y_mean
            mean[r0:r0 + chunk] = outs.mean(0)
            spread[r0:r0 + chunk] = outs.std(0)
        return mean, spread

    def close(self):
        for m in self.members:
            m.close()

    def save(self, path: str, **extra) -> None:
        """Parameters of every member, the target scaling and `extra` values, as an .npz file."""
        np.savez(path, params=np.stack([m.get_params() for m in self.members]), y_mean=self.y_mean,
                 y_std=self.y_std, **{k: np.asarray(v) for k, v in extra.items()})

    @classmethod
    def load(cls, path: str):
        """An ensemble saved by `save`, and the extra values stored with it."""
        with np.load(path, allow_pickle=False) as f:
            data = {k: f[k] for k in f.files}
        ens = cls(y_mean=data.pop("y_mean"), y_std=data.pop("y_std"))
        for q, p in enumerate(data.pop("params")):
            net = MLP(seed=q)
            net.set_params(p)
            ens.members.append(net)
        return ens, data


def triage(ens: Ensemble, P: np.ndarray, threshold: float, audit_frac: float = 0.0, seed: int = 0) -> dict:
    """The surrogate half of `serve`: ensemble mean and spread of every row of P (fused passes of 4096
    rows), the rows to re-run by Monte Carlo (`flagged`: outside the training ranges, or largest member
    standard deviation over the 33 outputs above `threshold`, log10 units) and a seeded `audit_frac`
    share of the others (`audit`).  `surrogate_s` is the wall time of the inference."""
    import time

    P = np.atleast_2d(P)
    t0 = time.perf_counter()
    mean, spread = ens.predict(P)
    t_sur = time.perf_counter() - t0
    flagged = (~in_range(P)) | (spread.max(1) > threshold)
    audit = np.zeros(len(P), bool)
    if audit_frac > 0:
        audit = (~flagged) & (np.random.default_rng(seed).uniform(size=len(P)) < audit_frac)
    return {"mean": mean, "spread": spread, "flagged": flagged, "audit": audit, "surrogate_s": t_sur}


def serve(ens: Ensemble, P: np.ndarray, engine, threshold: float, fallback_photons: int, audit_frac: float = 0.0,
          seed: int = 0, **mc_kw) -> dict:
    """Answers parameter rows P with the surrogate, falling back to Monte Carlo.

    A row is re-run by batched Monte Carlo (`fallback_photons` packets, surrogate_data.run_mc) when it
    lies outside the training ranges or when the largest ensemble standard deviation over its 33
    outputs exceeds `threshold` (log10 units).  A seeded `audit_frac` share of the other rows is
    re-run too, to measure the served error; those keep the surrogate answer.  The Monte Carlo
    launches stream through surrogate_data.ENGINES concurrent engines, which need as many hardware
    queues (GPU_MAX_HW_QUEUES=16 in the environment before HIP starts; HIP's default is 4).  Returns `pred`
    (log10 targets, Monte Carlo values in flagged rows), `flagged`, `audit`, `rows` (indices
    re-run), `mc` (the run_mc result for `rows`) and `surrogate_s` (inference seconds)."""
    from . import surrogate_data

    P = np.atleast_2d(P)
    tr = triage(ens, P, threshold, audit_frac, seed)
    mean, spread, flagged, audit, t_sur = tr["mean"], tr["spread"], tr["flagged"], tr["audit"], tr["surrogate_s"]
    rows = np.flatnonzero(flagged | audit)
    mc = None
    pred = mean.copy()
    if rows.size:
        mc = surrogate_data.run_mc(engine, P[rows], fallback_photons, seed=seed + 1, **mc_kw)
        fb = flagged[rows]
        pred[rows[fb]] = encode_targets(mc["radial"][fb], mc["total"][fb])
    return {"pred": pred, "mean": mean, "spread": spread, "flagged": flagged, "audit": audit, "rows": rows,
            "mc": mc, "surrogate_s": t_sur}
