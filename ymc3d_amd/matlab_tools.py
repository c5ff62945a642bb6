"""Host counterparts of the reference's MATLAB model tools (matlab/geometry/*.m, matlab/read/*.m).

The reference ships MATLAB scripts that build a model and read results; this module does the same
in numpy, writing and reading the same binary files, so existing model scripts keep working:

- `prepare_geometry`        matlab/geometry/prepare_geometry.m:1-56 (layered skin with one vessel, mm)
- `write_geometry_file`     matlab/geometry/write_geometry_file.m:1-22
- `melanin`, `musr`, `g`    matlab/geometry/prepare_optprops.m:1-14 (spectral models)
- `optprops_table`,
  `prepare_optprops`        matlab/geometry/prepare_optprops.m:16-57
- `write_optical_props_file` matlab/geometry/write_optical_props_file.m:1-10
- `read_optical_props_file` reads either layout (see below)
- `prepare_model`           matlab/geometry/prepare_model.m:1-4, with the tissue-count check the
                            script's own comment (:1) says it lacks
- `read_general_mcfile`,
  `read_ymc3d_outputfile`   matlab/read/read_general_mcfile.m, read_ymc3d_outputfile.m

Optical-property file layout pitfall.  write_optical_props_file.m:8 writes the T x 4 MATLAB matrix
with fwrite, i.e. column-major: all T values of mua, then all mus, all g, all n.  The C reader
(opticalprops_read_from_file, src/mc3d_io.cu:16-66) reads T records of (mua, mus, g, n) instead, so a
file written by the MATLAB script is read transposed by gpu-mc (for T > 1).  Both layouts are offered
here, chosen explicitly: layout="records" (what the C reader and the engine's CLI expect) or
layout="matlab" (what the MATLAB writer produces).  The MATLAB writer also names the file
[fname, 'bid'] with no dot (:6); `write_optical_props_file` keeps that naming.

Other quirks of the scripts, and what is done here:
- prepare_geometry.m:28-31 passes length/d quotients straight to zeros(); 1.7/0.02 is
  84.99999999999999 in binary64, which MATLAB rejects as a size.  The counts are rounded here.
- prepare_optprops.m:27 builds mua_blood from function handles multiplied by numbers (an error in
  MATLAB) and :54 names the file from an undefined `lam`; the intended expressions are used:
  mua_blood(oxy, x) = mua_oxy(x) oxy + mua_deoxy(x) (1 - oxy), file name + floor(lambda).
- the blood absorption spectra come from muab.mat, which the reference does not ship: the caller
  passes them (callables of wavelength).  Parity unpinned for the blood terms.
"""
from __future__ import annotations

import math
from typing import Callable

import numpy as np

from . import io

# Tissue codes of prepare_geometry.m:7-12 (file values; the C reader maps value v to row v + 1).
EPI, PAPDER, RETDER, SUBC, MOLE, VESSEL = 0, 1, 2, 3, 4, 5


def prepare_geometry(filename: str | None = None, dx: float = 20e-3, dy: float = 20e-3, dz: float = 20e-3,
                     length=(1.0, 1.0, 1.7), epi_length: float = 0.1, der_length: float = 1.5,
                     vessel_x: float | None = None, vessel_radius: float = 0.120 / 2.0, vessel_depth: float = 0.250):
    """The layered skin model of prepare_geometry.m, in mm.

    Layers (:20-37): epidermis floor(0.1/dz) voxels, papillary dermis floor(0.1*1.5/dz) more, reticular
    dermis floor((1.5 - 0.15)/dz) more, subcutis to the bottom; each depth computed in binary64 as the
    script does.  One horizontal vessel along y (:40-54): voxel (i, j, k), 1-based as in the script, is
    vessel when (i dx - x)^2 + (k dz - d)^2 <= r^2, with x = length_x/2, r = 0.06 and d = 0.25 mm (the
    test uses i dx and k dz, the voxel's upper faces in 1-based indexing, as written).  Returns the
    file values int32 [nz, ny, nx] (x fastest, as write_geometry_file.m writes VOI(Nx,Ny,Nz)) and
    the voxel size; writes `filename` + '.bin' when a name is given (write_geometry_file.m:14)."""
    lx, ly, lz = length
    nx, ny, nz = (int(round(lx / dx)), int(round(ly / dy)), int(round(lz / dz)))
    pap_length = 0.1 * der_length
    ret_length = der_length - pap_length
    epi_vox = math.floor(epi_length / dz)
    pap_vox = epi_vox + math.floor(pap_length / dz)
    ret_vox = pap_vox + math.floor(ret_length / dz)
    voi = np.empty((nz, ny, nx), np.int32)
    voi[:epi_vox] = EPI
    voi[epi_vox:pap_vox] = PAPDER
    voi[pap_vox:ret_vox] = RETDER
    voi[ret_vox:] = SUBC
    x = lx / 2.0 if vessel_x is None else vessel_x
    i1 = np.arange(1, nx + 1, dtype=np.float64)
    k1 = np.arange(1, nz + 1, dtype=np.float64)
    inside = ((i1[None, :] * dx - x) ** 2 + (k1[:, None] * dz - vessel_depth) ** 2) <= vessel_radius ** 2  # [nz, nx]
    voi[np.broadcast_to(inside[:, None, :], voi.shape)] = VESSEL
    if filename is not None:
        write_geometry_file(filename, voi, dx, dy, dz)
    return voi, (dx, dy, dz)


def write_geometry_file(fname: str, values: np.ndarray, dx: float, dy: float, dz: float) -> str:
    """write_geometry_file.m: int32 3, int32 Nx Ny Nz, f64 dx dy dz, int32 values x fastest, to
    fname + '.bin'.  values: int32 [nz, ny, nx].  Returns the path written."""
    path = fname + ".bin"
    io.write_geometry(path, values, dx, dy, dz)
    return path


def melanin(lam):
    """Melanin absorption spectrum shape (lambda/694)^-3.46 (prepare_optprops.m:1-3), lambda in nm."""
    return (np.asarray(lam, np.float64) / 694.0) ** (-3.46)


def musr(lam):
    """Reduced scattering 100 (aMie (lambda/500)^-bMie + aRay (lambda/500)^-4), aMie 18.780, bMie 0.22,
    aRay 17.6 (prepare_optprops.m:5-10)."""
    l5 = np.asarray(lam, np.float64) / 500.0
    return 100.0 * (18.780 * l5 ** (-0.22) + 17.6 * l5 ** (-4))


def g(lam):
    """Anisotropy 0.62 + lambda 29e-5 (prepare_optprops.m:12-14)."""
    return 0.62 + np.asarray(lam, np.float64) * 29e-05


def optprops_table(lam: float, mua_oxy: Callable[[float], float], mua_deoxy: Callable[[float], float]) -> np.ndarray:
    """The 3 x 4 table of prepare_optprops.m:27-49 at wavelength lam (nm): rows epidermis, papillary
    dermis, reticular dermis; columns mua, mus, g, n.  mua_oxy / mua_deoxy: blood absorption spectra
    (the script's muab.mat), in the units of the other terms; the 0.01 factors of the script apply."""
    blood = lambda oxy: mua_oxy(lam) * oxy + mua_deoxy(lam) * (1 - oxy)
    gg = float(g(lam))
    mus = float(musr(lam)) / (1 - gg) * 0.01
    t = np.empty((3, 4), np.float64)
    t[0] = ((227 * float(melanin(lam)) + 0.002 * blood(0.5) + (1 - 0.01 - 0.002) * 25) * 0.01, mus, gg, 1.4)
    bvf = 0.01
    t[1] = ((bvf * blood(0.5) + (1 - bvf) * 25) * 0.01, mus, gg, 1.4)
    t[2] = ((bvf * blood(0.8) + (1 - bvf) * 25) * 0.01, mus, gg, 1.4)
    return t


def prepare_optprops(lambdas, name: str, mua_oxy, mua_deoxy, layout: str = "matlab") -> list[str]:
    """One optical-property file per wavelength, named name + floor(lambda) + 'bid' as the script does
    (prepare_optprops.m:53-55 with write_optical_props_file.m:6); layout as for
    write_optical_props_file.  Returns the paths written."""
    paths = []
    for lam in np.atleast_1d(lambdas):
        paths.append(write_optical_props_file(f"{name}{int(math.floor(lam))}",
                                              optprops_table(float(lam), mua_oxy, mua_deoxy), layout))
    return paths


def write_optical_props_file(fname: str, optical_data: np.ndarray, layout: str = "matlab") -> str:
    """write_optical_props_file.m: int32 T, then the T x 4 doubles (mua, mus, g, n per tissue), to
    fname + 'bid'.  layout="matlab" writes them column-major as the MATLAB fwrite does (all mua, then
    mus, g, n); layout="records" writes one (mua, mus, g, n) record per tissue, the layout the C reader
    expects.  Returns the path written."""
    d = np.asarray(optical_data, np.float64)
    if d.ndim != 2 or d.shape[1] != 4:
        raise ValueError("optical_data must be [T, 4] (mua, mus, g, n)")
    path = fname + "bid"
    if layout == "records":
        io.write_optprops(path, d)
    elif layout == "matlab":
        with open(path, "wb") as f:
            f.write(np.int32(d.shape[0]).tobytes())
            f.write(np.ascontiguousarray(d.T, dtype="<f8").tobytes())
    else:
        raise ValueError("layout must be 'records' or 'matlab'")
    return path


def read_optical_props_file(path: str, layout: str = "records"):
    """(mua, mus, g, n), each T+1 long with row 0 the ambient medium (0, 0, 0, 1), as
    opticalprops_read_from_file (src/mc3d_io.cu:16-66) fills them.  layout="records" is that reader
    (through the engine's C ABI); layout="matlab" reads the column-major file the MATLAB writer
    produces."""
    if layout == "records":
        return io.read_optprops(path)
    if layout != "matlab":
        raise ValueError("layout must be 'records' or 'matlab'")
    with open(path, "rb") as f:
        T = int(np.frombuffer(f.read(4), "<i4")[0])
        vals = np.frombuffer(f.read(), "<f8")
    if T < 1 or vals.size != 4 * T:
        raise ValueError(f"{path}: {T} tissues but {vals.size} values")
    cols = vals.reshape(4, T).astype(np.float32)
    amb = np.array([0.0, 0.0, 0.0, 1.0], np.float32)
    return tuple(np.concatenate([[amb[c]], cols[c]]) for c in range(4))


def check_model(values: np.ndarray, num_tissues: int) -> None:
    """Raise when a geometry's file values name tissues the optical-property file lacks.  The C reader
    maps file value v to table row v + 1 (255 to the ambient row 0), so every other value must be
    below the file's tissue count; the engine refuses such a pair only when a run is prepared."""
    v = np.asarray(values)
    used = np.unique(v[v != 255])
    bad = used[(used < 0) | (used >= num_tissues)]
    if bad.size:
        raise ValueError(f"geometry uses tissue values {bad.tolist()} but the optical properties have "
                         f"{num_tissues} tissues (values 0..{num_tissues - 1})")


def prepare_model(lambdas, optprops_name: str, geometry_name: str, mua_oxy, mua_deoxy, layout: str = "matlab"):
    """prepare_model.m: the optical-property files for `lambdas`, then the geometry file.  The
    script's example is inconsistent by its own comment (:1): its geometry uses tissue values 3 and 5
    while its optical-property files hold three tissues; this raises on such a pair (check_model)
    after writing both, so the files can still be inspected.  Returns (optprops paths, geometry path)."""
    paths = prepare_optprops(lambdas, optprops_name, mua_oxy, mua_deoxy, layout)
    voi, (dx, dy, dz) = prepare_geometry()
    gpath = write_geometry_file(geometry_name, voi, dx, dy, dz)
    check_model(voi, 3)
    return paths, gpath


def read_general_mcfile(filename: str, datatype: str = "double"):
    """read_general_mcfile.m: int32 dims, int32 num_x, num_y[, num_z], f64 dx, dy[, dz], then the data
    in MATLAB order, returned as data[i, j, k] of shape (num_x, num_y, num_z) (num_z = 1 and dz = 0 for
    2-D files) with dx, dy, dz.  datatype: 'double', 'float', 'int32' or 'uint8'."""
    dt = {"double": "<f8", "float": "<f4", "single": "<f4", "int32": "<i4", "integer*4": "<i4", "uint8": "u1"}[datatype]
    with open(filename, "rb") as f:
        dims = int(np.frombuffer(f.read(4), "<i4")[0])
        nx, ny = (int(v) for v in np.frombuffer(f.read(8), "<i4"))
        nz = int(np.frombuffer(f.read(4), "<i4")[0]) if dims > 2 else 1
        dx, dy = (float(v) for v in np.frombuffer(f.read(16), "<f8"))
        dz = float(np.frombuffer(f.read(8), "<f8")[0]) if dims > 2 else 0.0
        data = np.frombuffer(f.read(np.dtype(dt).itemsize * nx * ny * nz), dt)
    if data.size != nx * ny * nz:
        raise ValueError(f"{filename}: expected {nx * ny * nz} values, found {data.size}")
    return data.reshape(nz, ny, nx).transpose(2, 1, 0), dx, dy, dz


def read_ymc3d_outputfile(filename: str):
    """read_ymc3d_outputfile.m: read_general_mcfile with doubles (the _drs.bin / _abs.bin outputs)."""
    return read_general_mcfile(filename, "double")
