"""ctypes binding of the HIP library (athad_cond_amd/libathad_hip.so).

The library is the product: every operator runs on the GPU through the C-ABI declared in
include/athad_hip.h. There is no CPU fallback; loading fails loudly when the library has
not been built.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libathad_hip.so")

FIELDS = [
    "t", "u", "v", "w", "c", "cloud", "ice", "gr", "co2", "tke", "dis",
    "tn", "un", "vn", "wn", "cn", "cloudn", "icen", "grn", "co2n", "tken", "disn",
    "p_dyn", "p_stat", "r_humid", "r_dry", "t_eq", "nue", "prod", "tke_source",
    "dis_source", "aux_u", "aux_v", "aux_w", "S_c_c",
]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
FIELDS2D = ["vel_star", "t_top_init"]
PROGNOSTIC = FIELDS[:11]

FP64, MIXED = 0, 1
OK, NOT_READY = 0, 6
LEX, JACOBI, SYMMETRIC = 0, 1, 2

KERNEL_ID = {"rk4": 0, "rk4_plane": 1, "level_mean": 2, "poisson": 3, "satadj": 4, "damp": 5,
             "densities": 6, "turb": 7, "bc": 8, "solid": 9, "polar": 10, "radial": 11,
             "store": 12, "init": 13, "ice": 14, "conv": 15}

# Exported entry points (include/athad_hip.h)
EXPORTS = [
    "athad_create", "athad_destroy", "athad_last_error", "athad_footprint_bytes", "athad_init_state",
    "athad_set_field", "athad_get_field", "athad_set_field2d", "athad_get_field2d", "athad_upload_rows",
    "athad_download_rows", "athad_set_iter", "athad_get_iter", "athad_poisson_sweep", "athad_satadj",
    "athad_damp_moist", "athad_densities", "athad_turbulence", "athad_boundary", "athad_rk4",
    "athad_filters", "athad_store", "athad_step", "athad_run", "athad_sync", "athad_profile",
    "athad_profile_result", "athad_version", "athad_dew_point", "athad_member_from_surface",
    "athad_query", "athad_device_memory", "athad_damp_tile", "athad_set_damp_tile",
    "athad_field_stats", "athad_set_member_params", "athad_restart_bytes", "athad_save_state_mem",
    "athad_load_state_mem", "athad_save_state", "athad_load_state", "athad_field_minmax", "athad_residuum",
    "athad_meridional_streamfunction", "athad_get_slice", "athad_radiation", "athad_get_radiation",
    "athad_set_radiation", "athad_moist_convection", "athad_set_conv_field", "athad_get_conv_field",
    "athad_ice_scheme", "athad_set_ice_scheme", "athad_set_mp_field", "athad_get_mp_field",
    "athad_field_minmax_noalloc", "athad_project_initial_velocity", "athad_set_moist_convection",
    "athad_get_conv_cape", "athad_set_turbulence_model", "athad_turbulence_init",
]

# Turbulence closures, TurbulenceAtm.h:24 order (ATHAD_TURB_* in include/athad_hip.h)
TURB_MODELS = ["k_epsilon", "k_omega", "k_omega_SST"]

# Arrays of the two-category ice scheme (athad_mp_field in include/athad_hip.h)
CONV_FIELDS = ["s", "M_u", "M_d", "s_u", "q_v_u", "q_c_u", "c_u", "e_l", "e_d", "e_p", "g_p", "s_d", "q_v_d",
               "P_conv", "E_u", "D_u", "u_u", "v_u", "w_u", "E_d", "D_d", "u_d", "v_d", "w_d", "MC_t", "MC_q",
               "MC_v", "MC_w", "Deep_beg", "Deep_end", "CloudBase", "LevelFreeSinking", "Evaporation"]
MP_FIELDS = ["P_rain", "P_snow", "P_rainn", "P_snown", "Precipitation", "S_v", "S_c", "S_i", "S_r", "S_s"]

# restart_arrays() of the reference checkpoint (atmosphere/FileIO_Atm.cpp:297-304), in file
# order; the four precipitation arrays are the ice scheme's (zeros where it has never run)
RESTART_ARRAYS = ["t", "u", "v", "w", "c", "cloud", "ice", "gr", "co2",
                  "tn", "un", "vn", "wn", "cn", "cloudn", "icen", "grn", "co2n",
                  "p_dyn", "p_stat", "tke", "dis", "tken", "disn", "nue",
                  "P_rain", "P_snow", "P_rainn", "P_snown"]
RESTART_MAGIC = 0x41544D31


class Grid(ctypes.Structure):
    _fields_ = [("im", ctypes.c_int), ("jm", ctypes.c_int), ("km", ctypes.c_int)]


class MemberParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "t_surf_equator", "t_surf_pole", "r_air", "x_H2O", "x_CO2", "x_N2",
        "c_0", "co2_0", "c_h2o_dry_top")]


def default_params() -> MemberParams:
    """cli/config_cond.xml values."""
    return MemberParams(513.15, 503.15, 40.8, 0.5578, 0.4109, 0.0313, 0.3464, 0.6233, 0.00423)


_lib = None


def lib():
    """Loads the library; raises when it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"athad_cond_amd: HIP library not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        dp = ctypes.POINTER(ctypes.c_double)
        L.athad_create.argtypes = [i, ctypes.POINTER(Grid), ctypes.POINTER(MemberParams), i, i,
                                   ctypes.POINTER(vp)]
        L.athad_destroy.argtypes = [vp]
        L.athad_last_error.restype = ctypes.c_char_p
        L.athad_last_error.argtypes = [vp]
        L.athad_footprint_bytes.restype = ctypes.c_int64
        L.athad_footprint_bytes.argtypes = [ctypes.POINTER(Grid), i]
        L.athad_init_state.argtypes = [vp, ctypes.c_uint64]
        for name in ("athad_set_field", "athad_get_field", "athad_set_field2d", "athad_get_field2d"):
            getattr(L, name).argtypes = [vp, i, i, dp]
        L.athad_upload_rows.argtypes = [vp, i, i, ctypes.POINTER(dp)]
        L.athad_download_rows.argtypes = [vp, i, i, ctypes.POINTER(dp)]
        L.athad_set_iter.argtypes = [vp, i, i, i]
        L.athad_get_iter.argtypes = [vp, i, ctypes.POINTER(i), ctypes.POINTER(i)]
        for name in ("athad_poisson_sweep", "athad_satadj", "athad_damp_moist", "athad_turbulence",
                     "athad_boundary", "athad_rk4", "athad_filters", "athad_store", "athad_step",
                     "athad_sync", "athad_turbulence_init"):
            getattr(L, name).argtypes = [vp]
        L.athad_densities.argtypes = [vp, i, i]
        L.athad_set_turbulence_model.argtypes = [vp, i]
        L.athad_project_initial_velocity.argtypes = [vp, i]
        L.athad_run.argtypes = [vp, i, i]
        L.athad_profile.argtypes = [vp, i, i]
        L.athad_profile_result.argtypes = [vp, i, ctypes.POINTER(d), ctypes.POINTER(i)]
        L.athad_version.restype = ctypes.c_char_p
        L.athad_query.argtypes = [vp]
        L.athad_field_stats.argtypes = [vp, i, i, dp, dp]
        L.athad_set_member_params.argtypes = [vp, i, ctypes.POINTER(MemberParams)]
        L.athad_damp_tile.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(i)]
        L.athad_set_damp_tile.argtypes = [vp, i, i]
        L.athad_device_memory.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
        L.athad_dew_point.restype = d
        L.athad_dew_point.argtypes = [d]
        L.athad_member_from_surface.argtypes = [d, d, d, d, ctypes.POINTER(MemberParams)]
        L.athad_restart_bytes.restype = ctypes.c_int64
        L.athad_restart_bytes.argtypes = [ctypes.POINTER(Grid)]
        L.athad_save_state_mem.argtypes = [vp, i, vp, ctypes.c_int64]
        L.athad_load_state_mem.argtypes = [vp, i, vp, ctypes.c_int64]
        L.athad_save_state.argtypes = [vp, i, ctypes.c_char_p]
        L.athad_load_state.argtypes = [vp, i, ctypes.c_char_p]
        ip = ctypes.POINTER(i)
        L.athad_field_minmax.argtypes = [vp, i, i, dp, ip, dp, ip]
        L.athad_meridional_streamfunction.argtypes = [vp, i, dp, dp, dp]
        L.athad_get_slice.argtypes = [vp, i, i, i, i, dp]
        L.athad_radiation.argtypes = [vp]
        L.athad_get_radiation.argtypes = [vp, i, dp, dp, dp, dp]
        L.athad_ice_scheme.argtypes = [vp]
        L.athad_set_radiation.argtypes = [vp, i]
        L.athad_set_ice_scheme.argtypes = [vp, i]
        L.athad_set_moist_convection.argtypes = [vp, i]
        L.athad_set_mp_field.argtypes = [vp, i, i, dp]
        L.athad_moist_convection.argtypes = [vp]
        L.athad_set_conv_field.argtypes = [vp, i, i, dp]
        L.athad_get_conv_field.argtypes = [vp, i, i, dp]
        L.athad_get_conv_cape.argtypes = [vp, i, dp]
        L.athad_get_mp_field.argtypes = [vp, i, i, dp]
        L.athad_residuum.argtypes = [vp, i, dp, ip, dp, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64), ip]
        L.athad_field_minmax_noalloc.argtypes = [vp, i, i, i, i, vp, dp, ctypes.POINTER(i), dp, ctypes.POINTER(i)]
        for name in EXPORTS:
            if name not in ("athad_last_error", "athad_footprint_bytes", "athad_version", "athad_dew_point",
                            "athad_restart_bytes"):
                getattr(L, name).restype = i
        _lib = L
    return _lib


class AthadError(RuntimeError):
    pass


def check(status, ctx=None):
    if status != 0:
        msg = lib().athad_last_error(ctx)
        raise AthadError(f"athad status {status}: {msg.decode() if msg else ''}")
