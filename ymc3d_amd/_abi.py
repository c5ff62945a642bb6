"""ctypes mirror of include/ymc3d.h (structs, constants and function signatures)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ABI_VERSION = 1
FX_ONE = 4294967296.0  # fixed-point scale of every weight sum

OK, ERR_ARG, ERR_IO, ERR_FORMAT, ERR_SEED, ERR_TABLE, ERR_DEVICE, ERR_NOMEM, ERR_CANCELLED, ERR_STATE = range(10)
STATUS_NAMES = {
    0: "OK", 1: "ERR_ARG", 2: "ERR_IO", 3: "ERR_FORMAT", 4: "ERR_SEED", 5: "ERR_TABLE",
    6: "ERR_DEVICE", 7: "ERR_NOMEM", 8: "ERR_CANCELLED", 9: "ERR_STATE",
}
ILLUM_UNIFORM, ILLUM_PENCIL = 0, 1
FRESNEL_YMC3D, FRESNEL_EXACT = 0, 1
FLAG_COUNTERS, FLAG_RECORD, FLAG_ABSORPTION = 0x1, 0x2, 0x4


class Geometry(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("tissue", C.POINTER(C.c_uint8))]


class Layered(C.Structure):
    _fields_ = [("top_voxels", C.c_int32), ("row_top", C.c_int32), ("row_bottom", C.c_int32),
                ("row_vessel", C.c_int32), ("vessel_x", C.c_float), ("vessel_z", C.c_float),
                ("vessel_r", C.c_float), ("reserved", C.c_float)]


LAYERED_DTYPE = np.dtype([("top_voxels", np.int32), ("row_top", np.int32), ("row_bottom", np.int32),
                          ("row_vessel", np.int32), ("vessel_x", np.float32), ("vessel_z", np.float32),
                          ("vessel_r", np.float32), ("reserved", np.float32)])


class OptProps(C.Structure):
    _fields_ = [("num_types", C.c_int32), ("num_tables", C.c_int32),
                ("mua", C.POINTER(C.c_float)), ("mus", C.POINTER(C.c_float)),
                ("g", C.POINTER(C.c_float)), ("n", C.POINTER(C.c_float))]


class Params(C.Structure):
    _fields_ = [("num_photons", C.c_uint64), ("budgets", C.POINTER(C.c_uint64)),
                ("seed", C.c_uint64), ("cancel", C.POINTER(C.c_int32)),
                ("illumination", C.c_int32), ("pencil_x", C.c_float), ("pencil_y", C.c_float),
                ("tally_nx", C.c_int32), ("tally_ny", C.c_int32),
                ("tally_dx", C.c_float), ("tally_dy", C.c_float),
                ("radial_bins", C.c_int32), ("radial_dr", C.c_float),
                ("radial_x0", C.c_float), ("radial_y0", C.c_float),
                ("fresnel_mode", C.c_int32), ("kill_zero_step", C.c_int32),
                ("chain_lanes", C.c_int32), ("pool_lanes", C.c_int32),
                ("compact_interval", C.c_int32), ("flags", C.c_uint32),
                ("jump_mode", C.c_int32), ("jump_max_crossings", C.c_int32),
                ("relaunch_lanes", C.c_int32), ("relaunch_wait", C.c_int32)]


class Totals(C.Structure):
    _fields_ = [("photons", C.c_uint64), ("R_w", C.c_uint64), ("R_w2", C.c_uint64),
                ("T_w", C.c_uint64), ("T_w2", C.c_uint64), ("R_n", C.c_uint64), ("T_n", C.c_uint64),
                ("A_w", C.c_uint64), ("roulette_w", C.c_int64), ("killed_w", C.c_uint64),
                ("killed", C.c_uint64), ("steps", C.c_uint64), ("fresnel", C.c_uint64)]


TOTALS_DTYPE = np.dtype([(name, np.int64 if name == "roulette_w" else np.uint64) for name, _ in Totals._fields_])


class Tally(C.Structure):
    _fields_ = [("R_w", C.POINTER(C.c_uint64)), ("R_w2", C.POINTER(C.c_uint64)), ("R_n", C.POINTER(C.c_uint32)),
                ("T_w", C.POINTER(C.c_uint64)), ("T_w2", C.POINTER(C.c_uint64)), ("T_n", C.POINTER(C.c_uint32)),
                ("Rr_w", C.POINTER(C.c_uint64)), ("Rr_w2", C.POINTER(C.c_uint64)), ("Rr_n", C.POINTER(C.c_uint32)),
                ("totals", C.c_void_p)]


class PhotonRecord(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float),
                ("steps", C.c_uint32), ("fate", C.c_int32)]


RECORD_DTYPE = np.dtype([("x", np.float32), ("y", np.float32), ("z", np.float32), ("w", np.float32),
                         ("steps", np.uint32), ("fate", np.int32)])
CHAIN_DTYPE = np.dtype([("count", np.uint32), ("table", np.uint32)])


def ptr(arr: np.ndarray | None, ctype):
    """Pointer to a contiguous numpy array (or NULL)."""
    if arr is None:
        return C.POINTER(ctype)()
    assert arr.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return arr.ctypes.data_as(C.POINTER(ctype))


def default_params(**kw) -> Params:
    p = Params()
    p.seed = 1
    p.illumination = ILLUM_UNIFORM
    p.pencil_x = p.pencil_y = -1.0
    p.radial_x0 = p.radial_y0 = -1.0
    p.kill_zero_step = 1
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"ymc_params has no field {k!r}")
        setattr(p, k, v)
    return p


P = C.POINTER
_SIGS = {
    "ymc_last_error": (C.c_char_p, []),
    "ymc_abi_version": (C.c_int32, []),
    "ymc_multipliers_generate": (C.c_int, [C.c_uint32, P(C.c_uint32)]),
    "ymc_multipliers_read_text": (C.c_int, [C.c_char_p, P(C.c_uint32), C.c_uint32, P(C.c_uint32)]),
    "ymc_seed_streams": (C.c_int, [P(C.c_uint32), C.c_uint32, C.c_uint64, C.c_uint64, P(C.c_uint64), P(C.c_uint32)]),
    "ymc_build_chains": (C.c_int, [P(C.c_uint64), C.c_int32, C.c_int32, C.c_void_p, C.c_uint64, P(C.c_uint64)]),
    "ymc_read_geometry": (C.c_int, [C.c_char_p, P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_double),
                                    P(C.c_double), P(C.c_double), P(P(C.c_uint8))]),
    "ymc_write_geometry": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                     C.c_double, P(C.c_int32)]),
    "ymc_read_optprops": (C.c_int, [C.c_char_p, P(C.c_int32), P(P(C.c_float)), P(P(C.c_float)),
                                    P(P(C.c_float)), P(P(C.c_float))]),
    "ymc_write_optprops": (C.c_int, [C.c_char_p, C.c_int32, P(C.c_double)]),
    "ymc_write_property": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                     C.c_double, C.c_double, P(C.c_double)]),
    "ymc_free": (None, [C.c_void_p]),
    "ymc_safe_radius": (C.c_int, [P(C.c_uint8), C.c_int32, C.c_int32, C.c_int32, P(C.c_uint8)]),
    "ymc_safe_box": (C.c_int, [P(C.c_uint8), C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(C.c_uint8),
                               P(C.c_uint8), P(C.c_uint8)]),
    "ymc_layered_voxelize": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, P(C.c_uint8)]),
    "ymc_engine_set_layered": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                         C.c_float, C.c_void_p, C.c_int32]),
    "ymc_engine_create": (C.c_int, [C.c_int32, P(C.c_uint32), C.c_uint32, P(C.c_void_p)]),
    "ymc_engine_destroy": (None, [C.c_void_p]),
    "ymc_engine_set_geometry": (C.c_int, [C.c_void_p, P(Geometry)]),
    "ymc_engine_set_optprops": (C.c_int, [C.c_void_p, P(OptProps)]),
    "ymc_engine_prepare": (C.c_int, [C.c_void_p, P(Params)]),
    "ymc_engine_prepare_chains": (C.c_int, [C.c_void_p, P(Params), C.c_void_p, P(C.c_uint64), P(C.c_uint32),
                                            C.c_uint64]),
    "ymc_engine_run": (C.c_int, [C.c_void_p]),
    "ymc_engine_read_tally": (C.c_int, [C.c_void_p, P(Tally)]),
    "ymc_engine_read_records": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "ymc_engine_read_absorption": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "ymc_engine_last_kernel_ms": (C.c_int, [C.c_void_p, P(C.c_double)]),
    "ymc_engine_pool_lanes": (C.c_int, [C.c_void_p, P(C.c_int32)]),
    "ymc_engine_rng_fill": (C.c_int, [C.c_void_p, P(C.c_uint64), P(C.c_uint32), C.c_uint32, C.c_uint32,
                                      C.c_int32, P(C.c_float), P(C.c_uint64)]),
    "ymc_engine_layered_words": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, P(C.c_uint32)]),
    "ymc_engine_rng_skip": (C.c_int, [C.c_void_p, P(C.c_uint64), P(C.c_uint32), C.c_uint32, C.c_uint32,
                                      P(C.c_uint64)]),
    "ymc_probe_fma_rate": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(C.c_double),
                                     P(C.c_double)]),
    "ymc_run_3dmc_gpu": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float,
                                   P(C.c_int32), C.c_int32, P(C.c_float), P(C.c_float), P(C.c_float),
                                   P(C.c_float), C.c_int32, C.c_uint64, P(C.c_uint32), C.c_uint32,
                                   P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_float),
                                   P(C.c_int32)]),
    "ymc_run": (C.c_int, [C.c_int32, P(C.c_uint32), C.c_uint32, P(Geometry), P(OptProps), P(Params), P(Tally),
                          C.c_void_p]),
    "ymc_mlp_create": (C.c_int, [C.c_int32, P(C.c_int32), C.c_int32, C.c_uint64, P(C.c_void_p)]),
    "ymc_mlp_destroy": (None, [C.c_void_p]),
    "ymc_mlp_num_params": (C.c_uint64, [C.c_void_p]),
    "ymc_mlp_set_params": (C.c_int, [C.c_void_p, P(C.c_float), C.c_uint64]),
    "ymc_mlp_get_params": (C.c_int, [C.c_void_p, P(C.c_float), C.c_uint64]),
    "ymc_mlp_forward": (C.c_int, [C.c_void_p, P(C.c_float), C.c_int64, P(C.c_float)]),
    "ymc_mlp_forward_ensemble": (C.c_int, [P(C.c_void_p), C.c_int32, P(C.c_float), C.c_int64, P(C.c_float)]),
    "ymc_mlp_forward_ensemble_noalloc": (C.c_int, [P(C.c_void_p), C.c_int32, P(C.c_float), C.c_int64,
                                                   P(C.c_float)]),
    "ymc_mlp_time_forward": (C.c_int, [P(C.c_void_p), C.c_int32, C.c_int32, C.c_int32, P(C.c_double)]),
    "ymc_mlp_gradients": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int32, P(C.c_float), P(C.c_float)]),
    "ymc_mlp_set_data": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float), C.c_int64]),
    "ymc_mlp_train": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_uint64, P(C.c_float)]),
}

# Entry points of the product library, in the order include/ymc3d.h declares them.
EXPORTED = tuple(_SIGS)


def bind(lib: C.CDLL) -> C.CDLL:
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PACKAGE_DIR, "libymc3d.so")
