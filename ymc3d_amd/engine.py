"""Host-side engine API over the C ABI of libymc3d.so.

The counterpart of the reference's run loop ``run_3dmc_gpu`` (src/mc3d_gpu.cu:388-601): the
caller sets a voxel geometry and optical-property tables, prepares a photon budget (stream
seeding and chain schedule on the host), runs the persistent HIP transport kernel and reads the
fixed-point tallies back.  There is no CPU fallback: without the built library every entry point
raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import FX_ONE, ptr

_LIB: C.CDLL | None = None


class YmcError(RuntimeError):
    """A C-ABI call returned a non-OK status."""

    def __init__(self, status: int, where: str, message: str):
        self.status = status
        super().__init__(f"{where}: {_abi.STATUS_NAMES.get(status, status)}: {message}")


def load_library() -> C.CDLL:
    """Loads the built engine (libymc3d.so next to this file); raises if it is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(_abi.LIB_PATH):
            raise ImportError(
                f"{_abi.LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C ymc3d_amd/csrc`")
        _LIB = _abi.bind(C.CDLL(_abi.LIB_PATH))
        if _LIB.ymc_abi_version() != _abi.ABI_VERSION:
            raise ImportError("libymc3d.so ABI version mismatch; rebuild")
    return _LIB


def check(status: int, where: str) -> None:
    if status != _abi.OK:
        msg = load_library().ymc_last_error().decode(errors="replace")
        raise YmcError(status, where, msg)


MULTIPLIER_CACHE = os.path.join(_abi.PACKAGE_DIR, "data", "multipliers.bin")
DEFAULT_MULTIPLIERS = (1 << 20) + 1


def generate_multipliers(count: int) -> np.ndarray:
    """The `count` largest a < 2^32 with a*2^32-1 and a*2^31-1 prime, descending."""
    out = np.empty(count, dtype=np.uint32)
    check(load_library().ymc_multipliers_generate(count, ptr(out, C.c_uint32)), "ymc_multipliers_generate")
    return out


def read_multipliers_text(path: str, cap: int = DEFAULT_MULTIPLIERS) -> np.ndarray:
    """First column of a safeprimes_base32.txt-style file (the reference's RNG table format)."""
    out = np.empty(cap, dtype=np.uint32)
    n = C.c_uint32(0)
    check(load_library().ymc_multipliers_read_text(path.encode(), ptr(out, C.c_uint32), cap, C.byref(n)),
          "ymc_multipliers_read_text")
    return out[: n.value].copy()


def default_multipliers() -> np.ndarray:
    """Multiplier table used when none is given: the build-time cache, else generated now."""
    if os.path.exists(MULTIPLIER_CACHE):
        tab = np.fromfile(MULTIPLIER_CACHE, dtype="<u4")
        if tab.size >= 2:
            return tab
    return generate_multipliers(DEFAULT_MULTIPLIERS)


def seed_streams(table: np.ndarray, seed: int, n: int) -> tuple[np.ndarray, np.ndarray]:
    """init_RNG (gpumcml_rng/gpumcml_rng.cu:59-118) for n streams; returns (x, a)."""
    table = np.ascontiguousarray(table, dtype=np.uint32)
    x = np.empty(n, dtype=np.uint64)
    a = np.empty(n, dtype=np.uint32)
    check(load_library().ymc_seed_streams(ptr(table, C.c_uint32), table.size, seed, n, ptr(x, C.c_uint64),
                                          ptr(a, C.c_uint32)), "ymc_seed_streams")
    return x, a


def build_chains(budgets, chain_lanes: int = 524288) -> np.ndarray:
    """Guided chain schedule (structured array of (count, table))."""
    b = np.ascontiguousarray(np.atleast_1d(budgets), dtype=np.uint64)
    lib = load_library()
    n = C.c_uint64(0)
    check(lib.ymc_build_chains(ptr(b, C.c_uint64), b.size, chain_lanes, None, 0, C.byref(n)), "ymc_build_chains")
    out = np.empty(n.value, dtype=_abi.CHAIN_DTYPE)
    check(lib.ymc_build_chains(ptr(b, C.c_uint64), b.size, chain_lanes, out.ctypes.data_as(C.c_void_p), n.value,
                               C.byref(n)), "ymc_build_chains")
    return out


@dataclass
class TallyResult:
    """Fixed-point tallies of one run (scale 2^32).  Maps are [tables, ny, nx], x fastest."""

    R_w: np.ndarray
    R_w2: np.ndarray
    R_n: np.ndarray
    T_w: np.ndarray
    T_w2: np.ndarray
    T_n: np.ndarray
    Rr_w: np.ndarray
    Rr_w2: np.ndarray
    Rr_n: np.ndarray
    totals: np.ndarray  # structured, one record per table (ymc_totals)
    kernel_ms: float = float("nan")

    @property
    def photons(self) -> np.ndarray:
        return self.totals["photons"].astype(np.float64)

    def _norm(self, w: np.ndarray) -> np.ndarray:
        n = self.photons.reshape((-1,) + (1,) * (w.ndim - 1))
        return w.astype(np.float64) / FX_ONE / n

    def drs(self) -> np.ndarray:
        """Diffuse reflectance per bin, sum(W)/N (the _drs.bin quantity, src/mc3d_io.cu:218-227)."""
        return self._norm(self.R_w)

    def drs_se(self) -> np.ndarray:
        return _se(self.R_w, self.R_w2, self.photons)

    def radial(self) -> np.ndarray:
        return self._norm(self.Rr_w)

    def radial_se(self) -> np.ndarray:
        return _se(self.Rr_w, self.Rr_w2, self.photons)

    def total_R(self) -> np.ndarray:
        return self.totals["R_w"].astype(np.float64) / FX_ONE / self.photons

    def total_R_se(self) -> np.ndarray:
        return _se(self.totals["R_w"], self.totals["R_w2"], self.photons)

    def total_T(self) -> np.ndarray:
        return self.totals["T_w"].astype(np.float64) / FX_ONE / self.photons


def _se(w, w2, n):
    """Standard error of the per-packet mean: sqrt((E[X^2] - E[X]^2) / (N - 1))."""
    n = np.asarray(n, dtype=np.float64).reshape((-1,) + (1,) * (np.ndim(w) - 1))
    m1 = np.asarray(w, dtype=np.float64) / FX_ONE / n
    m2 = np.asarray(w2, dtype=np.float64) / FX_ONE / n
    return np.sqrt(np.maximum(m2 - m1 * m1, 0.0) / np.maximum(n - 1, 1))


def _as_tables(v, name: str) -> np.ndarray:
    a = np.asarray(v, dtype=np.float32)
    if a.ndim == 1:
        a = a[None, :]
    if a.ndim != 2:
        raise ValueError(f"{name} must be [types] or [tables, types]")
    return np.ascontiguousarray(a)


class Engine:
    """One device context (ymc_engine)."""

    def __init__(self, device: int = 0, multipliers: np.ndarray | None = None):
        self._lib = load_library()
        self._h = C.c_void_p()
        tab = np.ascontiguousarray(default_multipliers() if multipliers is None else multipliers, dtype=np.uint32)
        check(self._lib.ymc_engine_create(device, ptr(tab, C.c_uint32), tab.size, C.byref(self._h)),
              "ymc_engine_create")
        self.multipliers = tab
        self._shape = None
        self._tables = 0
        self._prm = None
        self._keep = []

    def close(self) -> None:
        if self._h:
            self._lib.ymc_engine_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- inputs ------------------------------------------------------------------------------
    def set_geometry(self, tissue: np.ndarray, dx: float, dy: float, dz: float) -> None:
        """tissue: uint8 rows [nz, ny, nx] (0 = ambient), as produced by io.read_geometry."""
        t = np.ascontiguousarray(tissue, dtype=np.uint8)
        if t.ndim != 3:
            raise ValueError("tissue must be [nz, ny, nx]")
        nz, ny, nx = t.shape
        g = _abi.Geometry(nx, ny, nz, dx, dy, dz, ptr(t, C.c_uint8))
        check(self._lib.ymc_engine_set_geometry(self._h, C.byref(g)), "ymc_engine_set_geometry")
        self._shape = (nx, ny, nz, float(dx), float(dy), float(dz))

    def set_layered(self, per_table: np.ndarray, shape: tuple[int, int, int], voxel: tuple[float, float, float]) -> None:
        """Parametric geometry per table (structured LAYERED_DTYPE array); shape is (nx, ny, nz)."""
        g = np.ascontiguousarray(per_table, dtype=_abi.LAYERED_DTYPE)
        nx, ny, nz = shape
        check(self._lib.ymc_engine_set_layered(self._h, nx, ny, nz, *voxel, g.ctypes.data_as(C.c_void_p), g.size),
              "ymc_engine_set_layered")
        self._shape = (nx, ny, nz, *map(float, voxel))

    def set_optprops(self, mua, mus, g, n) -> None:
        """Arrays [types] or [tables, types]; row 0 is the ambient medium."""
        arrs = [_as_tables(v, k) for v, k in ((mua, "mua"), (mus, "mus"), (g, "g"), (n, "n"))]
        L, T = arrs[0].shape
        if any(a.shape != (L, T) for a in arrs):
            raise ValueError("optical-property arrays differ in shape")
        op = _abi.OptProps(T, L, *(ptr(a, C.c_float) for a in arrs))
        check(self._lib.ymc_engine_set_optprops(self._h, C.byref(op)), "ymc_engine_set_optprops")
        self._tables = L

    # -- runs --------------------------------------------------------------------------------
    def _params(self, kw) -> _abi.Params:
        kw = dict(kw)
        budgets = kw.pop("budgets", None)
        cancel = kw.pop("cancel", None)
        p = _abi.default_params(**kw)
        keep = []
        if budgets is not None:
            b = np.ascontiguousarray(budgets, dtype=np.uint64)
            if b.size != self._tables:
                raise ValueError("budgets must have one entry per table")
            p.budgets = ptr(b, C.c_uint64)
            keep.append(b)
        if cancel is not None:
            p.cancel = C.cast(cancel, C.POINTER(C.c_int32))
            keep.append(cancel)
        self._keep = keep
        return p

    def prepare(self, **params) -> None:
        """Seeds streams and uploads the chain schedule (fields of ymc_params as keywords)."""
        p = self._params(params)
        check(self._lib.ymc_engine_prepare(self._h, C.byref(p)), "ymc_engine_prepare")
        self._prm = p

    def prepare_chains(self, chains: np.ndarray, x: np.ndarray, a: np.ndarray, **params) -> None:
        p = self._params(params)
        ch = np.ascontiguousarray(chains, dtype=_abi.CHAIN_DTYPE)
        xx = np.ascontiguousarray(x, dtype=np.uint64)
        aa = np.ascontiguousarray(a, dtype=np.uint32)
        if not (ch.size == xx.size == aa.size):
            raise ValueError("chains, x and a must have equal length")
        check(self._lib.ymc_engine_prepare_chains(self._h, C.byref(p), ch.ctypes.data_as(C.c_void_p),
                                                  ptr(xx, C.c_uint64), ptr(aa, C.c_uint32), ch.size),
              "ymc_engine_prepare_chains")
        self._prm = p

    def run(self) -> float:
        """Runs the prepared budget; returns the transport kernel time in ms (HIP events)."""
        check(self._lib.ymc_engine_run(self._h), "ymc_engine_run")
        return self.last_kernel_ms()

    def last_kernel_ms(self) -> float:
        ms = C.c_double(0)
        check(self._lib.ymc_engine_last_kernel_ms(self._h, C.byref(ms)), "ymc_engine_last_kernel_ms")
        return ms.value

    def pool_lanes(self) -> int:
        v = C.c_int32(0)
        check(self._lib.ymc_engine_pool_lanes(self._h, C.byref(v)), "ymc_engine_pool_lanes")
        return v.value

    def read_tally(self) -> TallyResult:
        if self._prm is None or self._shape is None:
            raise RuntimeError("nothing prepared")
        p = self._prm
        nx, ny = self._shape[0], self._shape[1]
        tnx = p.tally_nx if p.tally_nx > 0 else nx
        tny = p.tally_ny if p.tally_ny > 0 else ny
        L, nr = self._tables, max(p.radial_bins, 0)
        mk = lambda dt, *s: np.zeros(s, dtype=dt)
        res = TallyResult(mk(np.uint64, L, tny, tnx), mk(np.uint64, L, tny, tnx), mk(np.uint32, L, tny, tnx),
                          mk(np.uint64, L, tny, tnx), mk(np.uint64, L, tny, tnx), mk(np.uint32, L, tny, tnx),
                          mk(np.uint64, L, nr), mk(np.uint64, L, nr), mk(np.uint32, L, nr),
                          np.zeros(L, dtype=_abi.TOTALS_DTYPE))
        t = _abi.Tally(ptr(res.R_w, C.c_uint64), ptr(res.R_w2, C.c_uint64), ptr(res.R_n, C.c_uint32),
                       ptr(res.T_w, C.c_uint64), ptr(res.T_w2, C.c_uint64), ptr(res.T_n, C.c_uint32),
                       ptr(res.Rr_w, C.c_uint64), ptr(res.Rr_w2, C.c_uint64), ptr(res.Rr_n, C.c_uint32),
                       res.totals.ctypes.data_as(C.c_void_p))
        check(self._lib.ymc_engine_read_tally(self._h, C.byref(t)), "ymc_engine_read_tally")
        res.kernel_ms = self.last_kernel_ms()
        return res

    def read_records(self, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=_abi.RECORD_DTYPE)
        check(self._lib.ymc_engine_read_records(self._h, out.ctypes.data_as(C.c_void_p), count),
              "ymc_engine_read_records")
        return out

    def read_absorption(self) -> np.ndarray:
        """Absorption grid of the last run made with FLAG_ABSORPTION: fixed-point absorbed weight
        (raw / 2^32) per voxel, [tables, nz, ny, nx]."""
        nx, ny, nz = self._shape[:3]
        out = np.zeros((self._tables, nz, ny, nx), dtype=np.uint64)
        check(self._lib.ymc_engine_read_absorption(self._h, out.ctypes.data_as(C.c_void_p), out.size),
              "ymc_engine_read_absorption")
        return out

    def simulate(self, **params) -> TallyResult:
        self.prepare(**params)
        self.run()
        return self.read_tally()

    def rng_fill(self, x: np.ndarray, a: np.ndarray, draws: int, kind: int = 0) -> tuple[np.ndarray, np.ndarray]:
        """Device MWC draws: out[draw, stream] as rand_MWC_co (kind 0) or rand_MWC_oc (kind 1)."""
        xx = np.ascontiguousarray(x, dtype=np.uint64)
        aa = np.ascontiguousarray(a, dtype=np.uint32)
        out = np.empty((draws, xx.size), dtype=np.float32)
        xf = np.empty_like(xx)
        check(self._lib.ymc_engine_rng_fill(self._h, ptr(xx, C.c_uint64), ptr(aa, C.c_uint32), xx.size, draws, kind,
                                            ptr(out, C.c_float), ptr(xf, C.c_uint64)), "ymc_engine_rng_fill")
        return out, xf

    def layered_words(self, geom, shape: tuple[int, int, int]) -> np.ndarray:
        """Device voxel words [nz, ny, nx] of one parametric geometry: row | zup << 8 | zdn << 16 | rxy << 24."""
        g = np.ascontiguousarray(np.atleast_1d(geom)[:1], dtype=_abi.LAYERED_DTYPE)
        nx, ny, nz = shape
        out = np.zeros((nz, ny, nx), dtype=np.uint32)
        check(self._lib.ymc_engine_layered_words(self._h, g.ctypes.data_as(C.c_void_p), nx, ny, nz,
                                                 ptr(out, C.c_uint32)), "ymc_engine_layered_words")
        return out

    def rng_skip(self, x: np.ndarray, a: np.ndarray, crossings: int) -> np.ndarray:
        """Device stream skip past the draws of `crossings` plain crossings (4 each, <= 63), with the
        modular table a box flight uses; returns the final states."""
        xx = np.ascontiguousarray(x, dtype=np.uint64)
        aa = np.ascontiguousarray(a, dtype=np.uint32)
        xf = np.empty_like(xx)
        check(self._lib.ymc_engine_rng_skip(self._h, ptr(xx, C.c_uint64), ptr(aa, C.c_uint32), xx.size, crossings,
                                            ptr(xf, C.c_uint64)), "ymc_engine_rng_skip")
        return xf


def layered_voxelize(geom, shape: tuple[int, int, int]) -> np.ndarray:
    """Rows [nz, ny, nx] of the voxel grid one parametric geometry describes."""
    g = np.ascontiguousarray(np.atleast_1d(geom)[:1], dtype=_abi.LAYERED_DTYPE)
    nx, ny, nz = shape
    rows = np.zeros((nz, ny, nx), dtype=np.uint8)
    check(load_library().ymc_layered_voxelize(g.ctypes.data_as(C.c_void_p), nx, ny, nz, ptr(rows, C.c_uint8)),
          "ymc_layered_voxelize")
    return rows


def run_3dmc_gpu(tissue_rows: np.ndarray, dx: float, dy: float, dz: float, mua, mus, g, n, num_photons: int,
                 seed: int = 1, multipliers: np.ndarray | None = None):
    """Drop-in for run_3dmc_gpu (src/mc3d_gpu.cu:388): returns (R, R_tot, T, T_tot, num_finished) with R and T
    holding raw weight sums per geometry bin, like the reference's output arrays."""
    lib = load_library()
    t = np.ascontiguousarray(tissue_rows, dtype=np.int32)
    nz, ny, nx = t.shape
    arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in (mua, mus, g, n)]
    tab = np.ascontiguousarray(default_multipliers() if multipliers is None else multipliers, dtype=np.uint32)
    R = np.zeros((ny, nx))
    T = np.zeros((ny, nx))
    A = np.zeros((nz, ny, nx), dtype=np.float32)
    rt, tt, nf = C.c_double(0), C.c_double(0), C.c_int32(0)
    check(lib.ymc_run_3dmc_gpu(nx, ny, nz, dx, dy, dz, ptr(t, C.c_int32), arrs[0].size,
                               *(ptr(a, C.c_float) for a in arrs), num_photons, seed, ptr(tab, C.c_uint32), tab.size,
                               ptr(R, C.c_double), C.byref(rt), ptr(T, C.c_double), C.byref(tt), ptr(A, C.c_float),
                               C.byref(nf)), "ymc_run_3dmc_gpu")
    return R, rt.value, T, tt.value, nf.value


def run_once(tissue_rows: np.ndarray, voxel, optprops, multipliers: np.ndarray | None = None, device: int = 0,
             **params) -> TallyResult:
    """One-shot run (ymc_run): uint8 remapped rows [nz, ny, nx], voxel sizes, optical properties
    (mua, mus, g, n as [tables, T] or [T]) and ymc_params fields as keywords (budgets for several
    tables).  With flags including FLAG_ABSORPTION the result carries `.absorption`
    [tables, nz, ny, nx]."""
    lib = load_library()
    rows = np.ascontiguousarray(tissue_rows, dtype=np.uint8)
    nz, ny, nx = rows.shape
    arrs = []
    for v in optprops:
        v = np.asarray(v, dtype=np.float32)
        arrs.append(np.ascontiguousarray(v[None] if v.ndim == 1 else v))
    L, T = arrs[0].shape
    geo = _abi.Geometry(nx, ny, nz, voxel[0], voxel[1], voxel[2], ptr(rows, C.c_uint8))
    op = _abi.OptProps(T, L, *(ptr(v, C.c_float) for v in arrs))
    kw = dict(params)
    budgets = kw.pop("budgets", None)
    cancel = kw.pop("cancel", None)
    p = _abi.default_params(**kw)
    keep = []
    if budgets is not None:
        b = np.ascontiguousarray(budgets, dtype=np.uint64)
        if b.size != L:
            raise ValueError("budgets must have one entry per table")
        p.budgets = ptr(b, C.c_uint64)
        keep.append(b)
    if cancel is not None:
        p.cancel = C.cast(cancel, C.POINTER(C.c_int32))
    tnx = p.tally_nx if p.tally_nx > 0 else nx
    tny = p.tally_ny if p.tally_ny > 0 else ny
    nr = max(p.radial_bins, 0)
    mk = lambda dt, *s: np.zeros(s, dtype=dt)
    res = TallyResult(mk(np.uint64, L, tny, tnx), mk(np.uint64, L, tny, tnx), mk(np.uint32, L, tny, tnx),
                      mk(np.uint64, L, tny, tnx), mk(np.uint64, L, tny, tnx), mk(np.uint32, L, tny, tnx),
                      mk(np.uint64, L, nr), mk(np.uint64, L, nr), mk(np.uint32, L, nr),
                      np.zeros(L, dtype=_abi.TOTALS_DTYPE))
    t = _abi.Tally(ptr(res.R_w, C.c_uint64), ptr(res.R_w2, C.c_uint64), ptr(res.R_n, C.c_uint32),
                   ptr(res.T_w, C.c_uint64), ptr(res.T_w2, C.c_uint64), ptr(res.T_n, C.c_uint32),
                   ptr(res.Rr_w, C.c_uint64), ptr(res.Rr_w2, C.c_uint64), ptr(res.Rr_n, C.c_uint32),
                   res.totals.ctypes.data_as(C.c_void_p))
    grid = np.zeros((L, nz, ny, nx), dtype=np.uint64) if p.flags & _abi.FLAG_ABSORPTION else None
    tab = np.ascontiguousarray(default_multipliers() if multipliers is None else multipliers, dtype=np.uint32)
    st = lib.ymc_run(device, ptr(tab, C.c_uint32), tab.size, C.byref(geo), C.byref(op), C.byref(p), C.byref(t),
                     grid.ctypes.data_as(C.c_void_p) if grid is not None else None)
    if st not in (_abi.OK, _abi.ERR_CANCELLED):
        check(st, "ymc_run")
    res.status = st
    res.absorption = grid
    return res
