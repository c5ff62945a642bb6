"""Seeded scene generators for the benchmark workloads, in the reference's input layout.

Lengths are in cm and coefficients in 1/cm.  Each scene holds the geometry in file numbering
(0-based optical-property records, 255 = ambient) plus the remapped rows the engine uploads, the
optical-property records (one table per wavelength / property set) and the run parameters.

- skin_slab_sphere (64x64x32, voxel 0.01): slab mua=1 mus=100 g=0.9 n=1.4 with a spherical
  inclusion r=0.05 centred 0.1 below the surface (mua=20); pencil beam at the surface centre,
  1e6 packets, 64x64 reflectance grid.
- skin_vessels (256x256x128, voxel 0.002): epidermis 100 um (n=1.44), dermis 1.5 mm (n=1.39),
  subcutis (n=1.44), eight seeded horizontal vessels (diameter 30-200 um, axis depth 0.2-1 mm,
  n=1.39); per-tissue mua ~ logU(0.1, 50), mus ~ logU(50, 500), g ~ U(0.7, 0.95); uniform
  illumination, 1e8 packets, 256x256 grid.
- batched_spectra: the slab-sphere geometry with 64 wavelengths, seeded smooth spectra
  (mua a sum of Gaussians, mus' = a*(lambda/500nm)^-b, mus = mus'/(1-g) as prepare_optprops.m:35),
  budgets ~ logU(1e5, 1e7), 256x256 grid per wavelength.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Scene:
    name: str
    file_values: np.ndarray  # int32 [nz, ny, nx]
    voxel: tuple[float, float, float]
    records: np.ndarray      # float64 [tables, T, 4] (mua, mus, g, n) per file record
    params: dict = field(default_factory=dict)

    @property
    def rows(self) -> np.ndarray:
        v = self.file_values
        return np.where(v == 255, 0, v + 1).astype(np.uint8)

    @property
    def shape(self) -> tuple[int, int, int]:
        nz, ny, nx = self.file_values.shape
        return nx, ny, nz

    def optprops(self) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
        """(mua, mus, g, n) as float32 [tables, T+1] with row 0 the ambient medium."""
        L, T, _ = self.records.shape
        out = []
        for q, amb in enumerate((0.0, 0.0, 0.0, 1.0)):
            a = np.empty((L, T + 1), dtype=np.float32)
            a[:, 0] = amb
            a[:, 1:] = self.records[:, :, q].astype(np.float32)
            out.append(a)
        return tuple(out)

    def load(self, engine) -> None:
        engine.set_geometry(self.rows, *self.voxel)
        engine.set_optprops(*self.optprops())


def _centres(n: int, d: float) -> np.ndarray:
    return (np.arange(n, dtype=np.float64) + 0.5) * d


def skin_slab_sphere(photons: int = 1_000_000, seed: int = 1) -> Scene:
    nx, ny, nz, d = 64, 64, 32, 0.01
    vals = np.zeros((nz, ny, nx), dtype=np.int32)
    z, y, x = np.meshgrid(_centres(nz, d), _centres(ny, d), _centres(nx, d), indexing="ij")
    inside = (x - 0.32) ** 2 + (y - 0.32) ** 2 + (z - 0.10) ** 2 <= 0.05 ** 2
    vals[inside] = 1
    rec = np.array([[[1.0, 100.0, 0.9, 1.4], [20.0, 100.0, 0.9, 1.4]]])
    params = dict(num_photons=photons, seed=seed, illumination=1, pencil_x=0.32, pencil_y=0.32)
    return Scene("skin_slab_sphere_64x64x32", vals, (d, d, d), rec, params)


def _vessel_mask(rng, nx, ny, nz, d, count, dmin, dmax, zmin, zmax):
    z, y, x = np.meshgrid(_centres(nz, d), _centres(ny, d), _centres(nx, d), indexing="ij")
    mask = np.zeros((nz, ny, nx), dtype=bool)
    Lx, Ly = nx * d, ny * d
    for _ in range(count):
        diam = rng.uniform(dmin, dmax)
        depth = rng.uniform(zmin, zmax)
        theta = rng.uniform(0.0, np.pi)
        x0, y0 = rng.uniform(0, Lx), rng.uniform(0, Ly)
        # distance from the voxel centre to a horizontal axis through (x0, y0, depth)
        ux, uy = np.cos(theta), np.sin(theta)
        px, py = x - x0, y - y0
        along = px * ux + py * uy
        perp2 = (px - along * ux) ** 2 + (py - along * uy) ** 2 + (z - depth) ** 2
        mask |= perp2 <= (diam / 2) ** 2
    return mask


def skin_vessels(photons: int = 100_000_000, seed: int = 2) -> Scene:
    nx, ny, nz, d = 256, 256, 128, 0.002
    rng = np.random.default_rng(seed)
    vals = np.empty((nz, ny, nx), dtype=np.int32)
    epi, der = 5, 5 + 75  # 100 um and 1.5 mm in 20 um voxels
    vals[:epi] = 0
    vals[epi:der] = 1
    vals[der:] = 2
    mask = _vessel_mask(rng, nx, ny, nz, d, 8, 0.003, 0.020, 0.02, 0.10)
    mask[:epi] = False
    mask[der:] = False
    vals[mask] = 3
    T = 4
    mua = np.exp(rng.uniform(np.log(0.1), np.log(50.0), T))
    mus = np.exp(rng.uniform(np.log(50.0), np.log(500.0), T))
    g = rng.uniform(0.7, 0.95, T)
    n = np.array([1.44, 1.39, 1.44, 1.39])
    rec = np.stack([mua, mus, g, n], axis=1)[None]
    params = dict(num_photons=photons, seed=seed, illumination=0)
    return Scene("skin_vessels_256x256x128", vals, (d, d, d), rec, params)


def batched_spectra(tables: int = 64, seed: int = 3, budget_range=(1e5, 1e7), tally: int = 256) -> Scene:
    base = skin_slab_sphere()
    rng = np.random.default_rng(seed)
    lam = np.linspace(450.0, 850.0, tables)
    T = base.records.shape[1]
    rec = np.empty((tables, T, 4))
    for t in range(T):
        amp = rng.uniform(0.5, 20.0, 3) * (4.0 if t == 1 else 1.0)
        cen = rng.uniform(420.0, 880.0, 3)
        wid = rng.uniform(20.0, 80.0, 3)
        mua = 0.2 + (amp[None] * np.exp(-0.5 * ((lam[:, None] - cen[None]) / wid[None]) ** 2)).sum(1)
        a, b = rng.uniform(8.0, 20.0), rng.uniform(0.5, 2.0)
        g = base.records[0, t, 2]
        mus = a * (lam / 500.0) ** (-b) / (1.0 - g)
        rec[:, t, 0] = mua
        rec[:, t, 1] = mus
        rec[:, t, 2] = g
        rec[:, t, 3] = base.records[0, t, 3]
    budgets = np.exp(rng.uniform(np.log(budget_range[0]), np.log(budget_range[1]), tables)).astype(np.uint64)
    nx, ny, _ = base.shape
    params = dict(base.params)
    params.pop("num_photons")
    params.update(budgets=budgets, tally_nx=tally, tally_ny=tally, tally_dx=nx * 0.01 / tally,
                  tally_dy=ny * 0.01 / tally)
    return Scene(f"batched_spectra_{tables}x", base.file_values, base.voxel, rec, params)


def van_de_hulst_slab(photons: int = 1_000_000, seed: int = 4) -> Scene:
    """Index-matched slab (n=1) mua=10 mus=90 g=0.75, thickness 0.02 over ambient voxels."""
    nx, ny, nz, d = 8, 8, 14, 0.002
    vals = np.full((nz, ny, nx), 255, dtype=np.int32)
    vals[:10] = 0
    rec = np.array([[[10.0, 90.0, 0.75, 1.0]]])
    params = dict(num_photons=photons, seed=seed, illumination=0)
    return Scene("van_de_hulst_slab", vals, (0.01, 0.01, d), rec, params)


# ------------------------------------------------------------------------------------------------
# Two-layer skin with one vessel, 128x128x64 voxels of 0.002 cm, one parametric geometry per
# table: the training workload of the reflectance surrogate.  Twelve parameters per sample:
#   0-2  mua of epidermis, dermis, blood  [1/cm]   logU(0.1, 50), logU(0.1, 50), logU(1, 300)
#   3-5  mus' of epidermis, dermis, blood [1/cm]   logU(5, 50)
#   6-8  g of epidermis, dermis, blood             U(0.7, 0.95)
#   9    epidermis thickness [cm]                  2..8 voxels (40-160 um), uniform
#   10   vessel axis depth [cm]                    U(0.02, 0.1)
#   11   vessel diameter [cm]                      U(0.003, 0.02)
# mus = mus'/(1-g) (prepare_optprops.m:35); n = 1.44 (epidermis), 1.39 (dermis, blood).  Pencil beam
# at the surface centre; targets are 32 radial reflectance bins of 0.004 cm plus total reflectance.

SKIN2_SHAPE = (128, 128, 64)
SKIN2_VOXEL = 0.002
SKIN2_RADIAL = (32, 0.004)
PARAM_NAMES = ("mua_epi", "mua_der", "mua_blood", "musp_epi", "musp_der", "musp_blood", "g_epi", "g_der",
               "g_blood", "epi_thickness", "vessel_depth", "vessel_diameter")
PARAM_LOG = np.array([1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0], dtype=bool)
PARAM_LO = np.array([0.1, 0.1, 1.0, 5.0, 5.0, 5.0, 0.7, 0.7, 0.7, 2 * SKIN2_VOXEL, 0.02, 0.003])
PARAM_HI = np.array([50.0, 50.0, 300.0, 50.0, 50.0, 50.0, 0.95, 0.95, 0.95, 8 * SKIN2_VOXEL, 0.10, 0.020])


def sample_skin2_params(n: int, seed: int, outlier_frac: float = 0.0) -> np.ndarray:
    """Seeded parameter vectors [n, 12]; outlier_frac of them drawn from ranges widened by a quarter of
    their (log) width on both sides, so some parameters fall outside the training ranges."""
    rng = np.random.default_rng(seed)
    lo, hi = PARAM_LO.copy(), PARAM_HI.copy()
    u = rng.uniform(size=(n, 12))
    if outlier_frac > 0:
        out = rng.uniform(size=n) < outlier_frac
        u[out] = rng.uniform(-0.25, 1.25, size=(out.sum(), 12))
    P = np.where(PARAM_LOG, np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo))), lo + u * (hi - lo))
    P[:, 9] = np.clip(np.rint(P[:, 9] / SKIN2_VOXEL), 1, 20) * SKIN2_VOXEL  # whole voxels
    P[:, 6:9] = np.clip(P[:, 6:9], 0.5, 0.98)
    P[:, 10:12] = np.maximum(P[:, 10:12], SKIN2_VOXEL)
    return P


def skin2_tables(P: np.ndarray):
    """Optical-property tables [tables, 4 rows] (row 0 ambient) and layered geometries for parameter rows."""
    import ymc3d_amd._abi as abi

    P = np.atleast_2d(P)
    L = P.shape[0]
    mua = np.zeros((L, 4), np.float32)
    mus = np.zeros((L, 4), np.float32)
    g = np.zeros((L, 4), np.float32)
    n = np.ones((L, 4), np.float32)
    for r in range(3):
        mua[:, r + 1] = P[:, r]
        g[:, r + 1] = P[:, 6 + r]
        mus[:, r + 1] = P[:, 3 + r] / (1.0 - P[:, 6 + r])
    n[:, 1], n[:, 2], n[:, 3] = 1.44, 1.39, 1.39
    geo = np.zeros(L, dtype=abi.LAYERED_DTYPE)
    nx = SKIN2_SHAPE[0]
    geo["top_voxels"] = np.rint(P[:, 9] / SKIN2_VOXEL).astype(np.int32)
    geo["row_top"], geo["row_bottom"], geo["row_vessel"] = 1, 2, 3
    geo["vessel_x"] = nx / 2.0
    geo["vessel_z"] = P[:, 10] / SKIN2_VOXEL
    geo["vessel_r"] = P[:, 11] / 2.0 / SKIN2_VOXEL
    return (mua, mus, g, n), geo


def skin2_params(photons: int) -> dict:
    nx, ny, _ = SKIN2_SHAPE
    return dict(num_photons=photons, illumination=1, pencil_x=nx * SKIN2_VOXEL / 2, pencil_y=ny * SKIN2_VOXEL / 2,
                radial_bins=SKIN2_RADIAL[0], radial_dr=SKIN2_RADIAL[1], tally_nx=1, tally_ny=1,
                tally_dx=nx * SKIN2_VOXEL, tally_dy=ny * SKIN2_VOXEL)


def skin2_scene(p: np.ndarray, photons: int = 100_000, seed: int = 1) -> Scene:
    """One parameter vector as a materialised voxel-grid Scene (for the oracle and for tests)."""
    import ymc3d_amd as y

    props, geo = skin2_tables(p)
    rows = y.layered_voxelize(geo[0], SKIN2_SHAPE)
    vals = np.where(rows == 0, 255, rows.astype(np.int32) - 1)
    rec = np.stack([props[0][0, 1:], props[1][0, 1:], props[2][0, 1:], props[3][0, 1:]], axis=1)[None].astype(np.float64)
    params = dict(skin2_params(photons), seed=seed)
    return Scene("skin2_vessel_128x128x64", vals.astype(np.int32), (SKIN2_VOXEL,) * 3, rec, params)
