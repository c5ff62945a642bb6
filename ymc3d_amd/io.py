"""The reference's binary file formats (src/mc3d_io.cu), through the engine's C ABI.

- geometry.bin: int32 3, int32 nx, ny, nz, f64 dx, dy, dz, int32 values [nz][ny][nx] (x fastest);
  value v in 0..254 is optical-property record v, 255 is the ambient medium.
- optprops (*.bid): int32 T, then T records of f64 (mua, mus, g, n).
- save_property output (_drs.bin): int32 dims, int32 nx, ny[, nz], f64 dx, dy[, dz], f64 data.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._abi import ptr
from .engine import check, load_library


def read_geometry(path: str) -> tuple[np.ndarray, tuple[float, float, float]]:
    """Returns remapped uint8 rows [nz, ny, nx] (0 = ambient) and the voxel size
    (geometry_read_from_file, src/mc3d_io.cu:107-169)."""
    lib = load_library()
    nx, ny, nz = C.c_int32(), C.c_int32(), C.c_int32()
    dx, dy, dz = C.c_double(), C.c_double(), C.c_double()
    buf = C.POINTER(C.c_uint8)()
    check(lib.ymc_read_geometry(path.encode(), C.byref(nx), C.byref(ny), C.byref(nz), C.byref(dx), C.byref(dy),
                                C.byref(dz), C.byref(buf)), "ymc_read_geometry")
    try:
        n = nx.value * ny.value * nz.value
        rows = np.ctypeslib.as_array(buf, shape=(n,)).copy().reshape(nz.value, ny.value, nx.value)
    finally:
        lib.ymc_free(C.cast(buf, C.c_void_p))
    return rows, (dx.value, dy.value, dz.value)


def write_geometry(path: str, file_values: np.ndarray, dx: float, dy: float, dz: float) -> None:
    """file_values: int32 [nz, ny, nx] in file numbering (0-based records, 255 = ambient)."""
    v = np.ascontiguousarray(file_values, dtype=np.int32)
    nz, ny, nx = v.shape
    check(load_library().ymc_write_geometry(path.encode(), nx, ny, nz, dx, dy, dz, ptr(v, C.c_int32)),
          "ymc_write_geometry")


def rows_to_file_values(rows: np.ndarray) -> np.ndarray:
    """Inverse of the reader's remap: row 0 -> 255, row r -> r - 1."""
    r = np.asarray(rows, dtype=np.int32)
    return np.where(r == 0, 255, r - 1).astype(np.int32)


def file_values_to_rows(values: np.ndarray) -> np.ndarray:
    v = np.asarray(values, dtype=np.int32)
    if v.min(initial=0) < 0 or v.max(initial=0) > 255:
        raise ValueError("geometry values must lie in 0..255")
    return np.where(v == 255, 0, v + 1).astype(np.uint8)


def read_optprops(path: str) -> tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """(mua, mus, g, n), each T+1 long with row 0 the ambient medium (src/mc3d_io.cu:16-66)."""
    lib = load_library()
    T = C.c_int32()
    bufs = [C.POINTER(C.c_float)() for _ in range(4)]
    check(lib.ymc_read_optprops(path.encode(), C.byref(T), *(C.byref(b) for b in bufs)), "ymc_read_optprops")
    out = []
    for b in bufs:
        out.append(np.ctypeslib.as_array(b, shape=(T.value,)).copy())
        lib.ymc_free(C.cast(b, C.c_void_p))
    return tuple(out)


def write_optprops(path: str, records: np.ndarray) -> None:
    """records: [T, 4] (mua, mus, g, n) per tissue, written record by record (the C reader's layout)."""
    r = np.ascontiguousarray(records, dtype=np.float64)
    if r.ndim != 2 or r.shape[1] != 4:
        raise ValueError("records must be [T, 4]")
    check(load_library().ymc_write_optprops(path.encode(), r.shape[0], ptr(r, C.c_double)), "ymc_write_optprops")


def write_property(path: str, data: np.ndarray, d: tuple[float, ...]) -> None:
    """save_property (src/mc3d_io.cu:175-216): data [ny, nx] or [nz, ny, nx] (x fastest)."""
    a = np.ascontiguousarray(data, dtype=np.float64)
    if a.ndim == 2:
        ny, nx = a.shape
        check(load_library().ymc_write_property(path.encode(), 2, nx, ny, 1, d[0], d[1], 0.0, ptr(a, C.c_double)),
              "ymc_write_property")
    elif a.ndim == 3:
        nz, ny, nx = a.shape
        check(load_library().ymc_write_property(path.encode(), 3, nx, ny, nz, d[0], d[1], d[2], ptr(a, C.c_double)),
              "ymc_write_property")
    else:
        raise ValueError("data must be 2-D or 3-D")


def read_property(path: str) -> tuple[np.ndarray, tuple[float, ...]]:
    """Reader of the save_property layout (matlab/read/read_general_mcfile.m): returns data shaped
    [ny, nx] or [nz, ny, nx] and the bin sizes."""
    with open(path, "rb") as f:
        dims = int(np.frombuffer(f.read(4), "<i4")[0])
        if dims not in (2, 3):
            raise ValueError(f"{path}: {dims} dimensions")
        shape = np.frombuffer(f.read(4 * dims), "<i4")
        d = tuple(float(v) for v in np.frombuffer(f.read(8 * dims), "<f8"))
        data = np.frombuffer(f.read(), "<f8")
    nx, ny = int(shape[0]), int(shape[1])
    nz = int(shape[2]) if dims == 3 else 1
    if data.size != nx * ny * nz:
        raise ValueError(f"{path}: expected {nx * ny * nz} values, found {data.size}")
    return (data.reshape(ny, nx) if dims == 2 else data.reshape(nz, ny, nx)), d
