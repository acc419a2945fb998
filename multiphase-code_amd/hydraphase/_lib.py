"""ctypes binding of libhydra_hip.so (include/hydra_hip.h).

Loading fails loudly when the library has not been built: there is no
CPU fallback for the hot path.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libhydra_hip.so")

PARAM_D = ["pmax", "pmin", "tgmax", "tgmin", "ugmax", "ugmin", "vgmax", "vgmin", "wgmax", "wgmin",
           "upmax", "upmin", "vpmax", "vpmin", "wpmax", "wpmin", "tpmax", "tpmin",
           "pinf", "gamap", "gamag", "cpp", "cpg", "amolg", "prg",
           "epslnmax", "epslnmin", "epslninf", "phipzero",
           "cfl", "diap", "surfacetension", "as_", "ts"]
PARAM_I = ["niter", "itersub", "istart", "myflux", "myorder", "precision"]
# largest neles hp_dt_min_noalloc takes (include/hydra_hip.h)
NOALLOC_MAX_CELLS = 67108864
# hp_health status bits (include/hydra_hip.h)
HEALTH_NONFINITE = 1
HEALTH_PRESSURE = 2
PARAM_X = ["co2", "co4", "c23", "cmug", "rrg", "gravity", "prp", "foursigmabyd"]


class HpParams(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_double) for n in PARAM_D] + [(n, ctypes.c_int) for n in PARAM_I]
                + [(n, ctypes.c_double) for n in PARAM_X])


def make_params(d: dict) -> HpParams:
    p = HpParams()
    for n in PARAM_D + PARAM_I:
        setattr(p, n, d.get(n, 0))
    return p  # derived constants left at zero: the library derives them


class HydraError(RuntimeError):
    pass


_lib = None

EXPORTS = ["hp_create", "hp_destroy", "hp_upload_state", "hp_download_state", "hp_primitive",
           "hp_dt_min", "hp_set_dt", "hp_smooth", "hp_viscous", "hp_flux", "hp_rk_stage", "hp_rk",
           "hp_end_step", "hp_step", "hp_synchronize", "hp_last_error", "hp_get_array",
           "hp_set_array", "hp_last_timing", "hp_probe_cell", "hp_health", "hp_batch_health", "hp_check_division",
           "hp_batch_create", "hp_batch_destroy", "hp_batch_add", "hp_batch_remove", "hp_batch_set_grid",
           "hp_batch_set_block",
           "hp_batch_step", "hp_batch_synchronize", "hp_batch_last_ms", "hp_batch_last_error",
           "hp_case_bytes", "hp_set_donors", "hp_apply_donors", "hp_dt_min_noalloc",
           "hp_stream_bandwidth", "hp_step_diffs", "hp_last_diffs", "hp_set_flux_timing"]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HydraError(f"{LIB_PATH} is missing: build it first (python __graft_entry__.py build "
                         "or python multiphase-code_amd/hydraphase/build.py)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    c_int, c_double, c_size = ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    L.hp_create.argtypes = ([ctypes.POINTER(vp), c_int, c_int, c_int, dp, dp, dp, ip, c_int]
                            + [ip] * 9 + [ctypes.POINTER(HpParams)])
    L.hp_destroy.argtypes = [vp]
    L.hp_upload_state.argtypes = [vp, dp, c_int, dp, c_double, c_int]
    L.hp_download_state.argtypes = [vp, dp, c_int, dp, dp, ip]
    for n in ["hp_primitive", "hp_smooth", "hp_viscous", "hp_flux", "hp_rk", "hp_synchronize"]:
        getattr(L, n).argtypes = [vp]
    L.hp_dt_min.argtypes = [vp, dp]
    L.hp_set_dt.argtypes = [vp, c_double, c_double, c_int]
    L.hp_rk_stage.argtypes = [vp, c_int]
    L.hp_end_step.argtypes = [vp, dp]
    L.hp_step.argtypes = [vp, c_int]
    L.hp_step_diffs.argtypes = [vp, c_int, dp]
    L.hp_last_diffs.argtypes = [vp, dp]
    L.hp_set_flux_timing.argtypes = [vp, c_int]
    L.hp_last_error.argtypes = [vp]
    L.hp_last_error.restype = ctypes.c_char_p
    L.hp_get_array.argtypes = [vp, c_int, dp, c_size]
    L.hp_set_array.argtypes = [vp, c_int, dp, c_size]
    L.hp_last_timing.argtypes = [vp, dp, dp, ip]
    L.hp_probe_cell.argtypes = [vp, c_int, dp]
    L.hp_health.argtypes = [vp, ip]
    L.hp_batch_health.argtypes = [vp, ip, c_int]
    L.hp_check_division.argtypes = [dp, dp, c_size, ctypes.POINTER(c_size)]
    L.hp_batch_create.argtypes = [ctypes.POINTER(vp), c_int, c_int, c_int]
    L.hp_batch_destroy.argtypes = [vp]
    L.hp_batch_add.argtypes = [vp, vp, ip]
    L.hp_batch_remove.argtypes = [vp, vp]
    L.hp_batch_set_grid.argtypes = [vp, c_int, c_int, c_int]
    L.hp_batch_set_block.argtypes = [vp, c_int]
    L.hp_batch_step.argtypes = [vp, c_int, c_int]
    L.hp_batch_synchronize.argtypes = [vp]
    L.hp_batch_last_ms.argtypes = [vp, dp]
    L.hp_batch_last_error.argtypes = [vp]
    L.hp_batch_last_error.restype = ctypes.c_char_p
    L.hp_set_donors.argtypes = [vp, c_int, ip, ip, c_int, dp]
    L.hp_apply_donors.argtypes = [vp]
    L.hp_dt_min_noalloc.argtypes = [vp, c_size, vp, c_int, c_double, c_double, c_double, vp, vp]
    L.hp_stream_bandwidth.argtypes = [c_size, c_int, dp, dp]
    L.hp_case_bytes.argtypes = [c_int, c_int, c_int, c_int]
    L.hp_case_bytes.restype = c_size
    for n in EXPORTS:
        if n not in ("hp_last_error", "hp_batch_last_error", "hp_case_bytes"):
            getattr(L, n).restype = c_int
    _lib = L
    return L
