"""ctypes loader for the CPU oracle (libymc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, never by the ymc3d_amd package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from ymc3d_amd import _abi
from ymc3d_amd._abi import ptr

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libymc_oracle.so")
_lib = None

P = C.POINTER
_SIGS = {
    "ymc_oracle_last_error": (C.c_char_p, []),
    "ymc_oracle_mwc_step": (C.c_uint32, [P(C.c_uint64), C.c_uint32]),
    "ymc_oracle_u32_to_float_rz": (C.c_float, [C.c_uint32]),
    "ymc_oracle_mwc_fill": (None, [P(C.c_uint64), P(C.c_uint32), C.c_uint32, C.c_uint32, C.c_int32, P(C.c_float),
                                   P(C.c_uint64)]),
    "ymc_oracle_is_prime": (C.c_int32, [C.c_uint64]),
    "ymc_oracle_is_multiplier": (C.c_int32, [C.c_uint32]),
    "ymc_oracle_init_rng": (C.c_int32, [P(C.c_uint32), C.c_uint32, C.c_uint64, C.c_uint64, P(C.c_uint64),
                                        P(C.c_uint32)]),
    "ymc_oracle_grid_coord": (C.c_int32, [C.c_float, C.c_float, C.c_float]),
    "ymc_oracle_intersection": (C.c_float, [C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float,
                                            P(C.c_float), P(C.c_float)]),
    "ymc_oracle_fresnel": (C.c_float, [C.c_float, C.c_float, C.c_float, C.c_int32, P(C.c_float)]),
    "ymc_oracle_hg_costheta": (C.c_float, [C.c_float, C.c_float]),
    "ymc_oracle_specular": (C.c_float, [C.c_float, C.c_float]),
    "ymc_oracle_build_chains": (C.c_uint64, [P(C.c_uint64), C.c_int32, C.c_int32, C.c_void_p, C.c_uint64]),
    "ymc_oracle_run": (C.c_int32, [P(_abi.Geometry), P(_abi.OptProps), P(_abi.Params), C.c_void_p, P(C.c_uint64),
                                   P(C.c_uint32), C.c_uint64, P(_abi.Tally), C.c_void_p, C.c_uint64, C.c_int32]),
    "ymc_oracle_run_abs": (C.c_int32, [P(_abi.Geometry), P(_abi.OptProps), P(_abi.Params), C.c_void_p,
                                       P(C.c_uint64), P(C.c_uint32), C.c_uint64, P(_abi.Tally), C.c_void_p,
                                       C.c_uint64, C.c_int32, C.c_void_p]),
    "ymc_oracle_max_threads": (C.c_int32, []),
    "ymc_oracle_step": (C.c_int32, [P(_abi.Geometry), P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_float),
                                    C.c_int32, C.c_int32, C.c_uint64, P(C.c_float), P(C.c_int32), P(C.c_uint64),
                                    C.c_uint32, P(C.c_int32)]),
}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = C.CDLL(LIB)
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def mwc_fill(x, a, draws: int, kind: int = 0):
    xx = np.ascontiguousarray(x, dtype=np.uint64)
    aa = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.empty((draws, xx.size), dtype=np.float32)
    xf = np.empty_like(xx)
    lib().ymc_oracle_mwc_fill(ptr(xx, C.c_uint64), ptr(aa, C.c_uint32), xx.size, draws, kind, ptr(out, C.c_float),
                              ptr(xf, C.c_uint64))
    return out, xf


def init_rng(table, seed: int, n: int):
    t = np.ascontiguousarray(table, dtype=np.uint32)
    x = np.empty(n, dtype=np.uint64)
    a = np.empty(n, dtype=np.uint32)
    st = lib().ymc_oracle_init_rng(ptr(t, C.c_uint32), t.size, seed, n, ptr(x, C.c_uint64), ptr(a, C.c_uint32))
    return st, x, a


def step(rows, voxel, table, state, idx, x: int, a: int, fresnel_mode: int = 0, kill_zero_step: int = 0,
         index: int = 0):
    """One oracle step (src/mc3d_gpu.cu:181-353) of one packet.  rows [nz, ny, nx] uint8; table = (mua, mus,
    g, n) arrays over the rows; state = (x, y, z, ux, uy, uz, W, s); idx = (i, j, k, t).  Returns (frozen,
    state', idx', x', fresnel) with x' the MWC state after the step's draws."""
    r = np.ascontiguousarray(rows, dtype=np.uint8)
    nz, ny, nx = r.shape
    geo = _abi.Geometry(nx, ny, nz, voxel[0], voxel[1], voxel[2], ptr(r, C.c_uint8))
    cols = [np.ascontiguousarray(v, dtype=np.float32) for v in table]
    st = np.array(state, dtype=np.float32)
    ix = np.array(idx, dtype=np.int32)
    xs = np.array([x], dtype=np.uint64)
    fr = C.c_int32(0)
    frozen = lib().ymc_oracle_step(C.byref(geo), *(ptr(c, C.c_float) for c in cols), fresnel_mode, kill_zero_step,
                                   index, ptr(st, C.c_float), ptr(ix, C.c_int32), ptr(xs, C.c_uint64), a,
                                   C.byref(fr))
    return bool(frozen), st, ix, int(xs[0]), fr.value


def build_chains(budgets, chain_lanes: int):
    b = np.ascontiguousarray(np.atleast_1d(budgets), dtype=np.uint64)
    n = lib().ymc_oracle_build_chains(ptr(b, C.c_uint64), b.size, chain_lanes, None, 0)
    out = np.empty(n, dtype=_abi.CHAIN_DTYPE)
    lib().ymc_oracle_build_chains(ptr(b, C.c_uint64), b.size, chain_lanes, out.ctypes.data_as(C.c_void_p), n)
    return out


class OracleResult:
    pass


def run(scene_rows, voxel, optprops, chains, x, a, records: int = 0, threads: int = 0, absorption: bool = False,
        **params):
    """Runs the oracle over explicit chains.  Returns an engine-shaped TallyResult (+ records, and with
    absorption=True the fixed-point absorption grid [tables, nz, ny, nx] as `.absorption`)."""
    from ymc3d_amd.engine import TallyResult

    rows = np.ascontiguousarray(scene_rows, dtype=np.uint8)
    nz, ny, nx = rows.shape
    geo = _abi.Geometry(nx, ny, nz, voxel[0], voxel[1], voxel[2], ptr(rows, C.c_uint8))
    arrs = []
    for v in optprops:
        v = np.asarray(v, dtype=np.float32)
        arrs.append(np.ascontiguousarray(v[None] if v.ndim == 1 else v))
    L, T = arrs[0].shape
    op = _abi.OptProps(T, L, *(ptr(v, C.c_float) for v in arrs))
    params = dict(params)
    cancel = params.pop("cancel", None)
    budgets = params.pop("budgets", None)
    prm = _abi.default_params(**params)
    if cancel is not None:
        prm.cancel = C.cast(cancel, C.POINTER(C.c_int32))
    tnx = prm.tally_nx if prm.tally_nx > 0 else nx
    tny = prm.tally_ny if prm.tally_ny > 0 else ny
    nr = max(prm.radial_bins, 0)
    mk = lambda dt, *s: np.zeros(s, dtype=dt)
    res = TallyResult(mk(np.uint64, L, tny, tnx), mk(np.uint64, L, tny, tnx), mk(np.uint32, L, tny, tnx),
                      mk(np.uint64, L, tny, tnx), mk(np.uint64, L, tny, tnx), mk(np.uint32, L, tny, tnx),
                      mk(np.uint64, L, nr), mk(np.uint64, L, nr), mk(np.uint32, L, nr),
                      np.zeros(L, dtype=_abi.TOTALS_DTYPE))
    t = _abi.Tally(ptr(res.R_w, C.c_uint64), ptr(res.R_w2, C.c_uint64), ptr(res.R_n, C.c_uint32),
                   ptr(res.T_w, C.c_uint64), ptr(res.T_w2, C.c_uint64), ptr(res.T_n, C.c_uint32),
                   ptr(res.Rr_w, C.c_uint64), ptr(res.Rr_w2, C.c_uint64), ptr(res.Rr_n, C.c_uint32),
                   res.totals.ctypes.data_as(C.c_void_p))
    ch = np.ascontiguousarray(chains, dtype=_abi.CHAIN_DTYPE)
    xx = np.ascontiguousarray(x, dtype=np.uint64)
    aa = np.ascontiguousarray(a, dtype=np.uint32)
    rec = np.zeros(records, dtype=_abi.RECORD_DTYPE)
    grid = np.zeros((L, nz, ny, nx), dtype=np.uint64) if absorption else None
    st = lib().ymc_oracle_run_abs(C.byref(geo), C.byref(op), C.byref(prm), ch.ctypes.data_as(C.c_void_p),
                                  ptr(xx, C.c_uint64), ptr(aa, C.c_uint32), ch.size, C.byref(t),
                                  rec.ctypes.data_as(C.c_void_p) if records else None, records, threads,
                                  grid.ctypes.data_as(C.c_void_p) if absorption else None)
    if st not in (0, _abi.ERR_CANCELLED):
        raise RuntimeError(f"oracle run failed ({st}): {lib().ymc_oracle_last_error().decode()}")
    res.status = st
    res.records = rec
    res.absorption = grid
    return res


def run_scene(scene, multipliers, chain_lanes: int = 524288, records: int = 0, threads: int = 0, **overrides):
    """Runs a configs.Scene on the oracle with the engine's default schedule and seeding."""
    params = dict(scene.params)
    params.update(overrides)
    budgets = params.pop("budgets", None)
    n = params.pop("num_photons", None)
    L = scene.records.shape[0]
    if budgets is None:
        budgets = np.full(L, n, dtype=np.uint64)
    chains = build_chains(budgets, chain_lanes)
    st, x, a = init_rng(multipliers, params.get("seed", 1), chains.size)
    if st:
        raise RuntimeError(f"oracle seeding failed ({st})")
    return run(scene.rows, scene.voxel, scene.optprops(), chains, x, a, records=records, threads=threads, **params)
