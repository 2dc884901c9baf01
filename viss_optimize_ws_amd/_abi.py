"""ctypes mirror of include/vissws.h (structures and function signatures).

The same stage signatures are exported by the gfx950 library (prefix ``vw_``)
and by the CPU oracle used in tests (prefix ``vo_``), so one wrapper class
drives either.
"""
import ctypes as C

VW_OK = 0
VW_ERR_ARG = 1
VW_ERR_DEVICE = 2
VW_ERR_STATE = 3
VW_ERR_RANGE = 4
VW_ERR_IO = 5
VW_ERR_HALFSYM = 6

VW_GEOM_LUNAR8 = 0
VW_GEOM_CIRCULAR = 1
VW_PREC_PLAIN = 0
VW_PREC_MIXED = 1
VW_RECON_MIRROR = 0
VW_RECON_FULL = 1
VW_VISS_AUTO = 0
VW_VISS_ROWS = 1
VW_VISS_ANT = 2
VW_VISS_ANT_CHECK = 3

CLS_MIXED, CLS_VISIBLE, CLS_HIDDEN = 0, 1, 2


class vw_config(C.Structure):
    _fields_ = [
        ("nside", C.c_int32),
        ("freq_hz", C.c_double),
        ("geometry", C.c_int32),
        ("n_ant", C.c_int32),
        ("n_epochs", C.c_int32),
        ("dt_s", C.c_double),
        ("orbit_seed", C.c_uint64),
        ("day", C.c_int32),
        ("segs", C.c_int32),
        ("seg_tlen", C.POINTER(C.c_int32)),
        ("n_seg_explicit", C.c_int32),
        ("blockage", C.c_int32),
        ("occ_mode", C.c_int32),
        ("recon_mode", C.c_int32),
        ("tile_pix", C.c_int32),
        ("precision", C.c_int32),
        ("bl_max_m", C.c_double),
        ("occ_radius_km", C.c_double),
        ("occ_altitude_km", C.c_double),
        ("img_n", C.c_int32),
        ("healpix_out", C.c_int32),
        ("r_min_km", C.c_double),
        ("r_max_km", C.c_double),
        ("gm_km3_s2", C.c_double),
        ("class_budget_bytes", C.c_int64),
        ("serial_stages", C.c_int32),
        ("viss_form", C.c_int32),
        ("nsky", C.c_int32),
    ]


class vw_plan(C.Structure):
    _fields_ = [
        ("npix", C.c_int64),
        ("ntile", C.c_int64),
        ("n_img_pts", C.c_int64),
        ("n_img_tiles", C.c_int64),
        ("n_pairs", C.c_int32),
        ("n_epochs", C.c_int32),
        ("segs", C.c_int32),
        ("max_tlen", C.c_int32),
        ("orbit_res", C.c_int64),
        ("seg_len", C.c_int64),
        ("lambda_m", C.c_double),
        ("bl_max_m", C.c_double),
        ("cosphi", C.c_double),
        ("pix_area", C.c_double),
        ("device_bytes", C.c_int64),
    ]


class vw_seg_info(C.Structure):
    _fields_ = [
        ("t0", C.c_int32),
        ("tlen", C.c_int32),
        ("n_half", C.c_int64),
        ("fa", C.c_double),
        ("fb", C.c_double),
        ("denom", C.c_double),
        ("res", C.c_int64),
        ("half", C.c_int64),
        ("du", C.c_double),
        ("nuniq", C.c_int64),
        ("nkeys", C.c_int64),
    ]


class vw_geom_host(C.Structure):
    _fields_ = [(k, C.POINTER(C.c_double)) for k in ("u", "v", "w", "xa", "ya", "za", "xb", "yb", "zb")]


class vw_timings(C.Structure):
    _fields_ = [
        ("gen_ms", C.c_double),
        ("viss_ms", C.c_double),
        ("conj_ms", C.c_double),
        ("group_ms", C.c_double),
        ("recon_ms", C.c_double),
        ("finalize_ms", C.c_double),
        ("viss_kernel_ms", C.c_double),
        ("recon_kernel_ms", C.c_double),
        ("n_viss_launch", C.c_int32),
        ("n_recon_launch", C.c_int32),
    ]


P = C.c_void_p
PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)
PI32 = C.POINTER(C.c_int32)
PU32 = C.POINTER(C.c_uint32)
PU8 = C.POINTER(C.c_uint8)
i32, i64 = C.c_int32, C.c_int64

# name -> (restype, argtypes); shared by the vw_ (device) and vo_ (oracle) prefixes
STAGE_FUNCS = {
    "create": (C.c_int, [C.POINTER(vw_config), C.POINTER(P)]),
    "destroy": (None, [P]),
    "last_error": (C.c_char_p, [P]),
    "get_plan": (C.c_int, [P, C.POINTER(vw_plan)]),
    "get_schedule": (C.c_int, [P, PI32, PI32]),
    "set_sky": (C.c_int, [P, PD, i64, i32]),
    "gen_segment": (C.c_int, [P, i32, i32, C.POINTER(vw_geom_host)]),
    "conj_fill": (C.c_int, [P, PD, PD]),
    "recon_segment": (C.c_int, [P, C.POINTER(vw_seg_info)]),
    "finalize": (C.c_int, [P, PD, PU32, PD, PU32]),
    "reset_accumulators": (C.c_int, [P]),
    "get_lmn": (C.c_int, [P, PD, PD, PD, PD]),
    "get_tiles": (C.c_int, [P, i32, PD]),
    "get_img_points": (C.c_int, [P, PI64, PD, PD, PD]),
    "get_positions": (C.c_int, [P, PD]),
    "get_rows": (C.c_int, [P, PD, PD, PD]),
    "get_tile_classes": (C.c_int, [P, i32, PU8]),
    "get_masks": (C.c_int, [P, i64, i64, PU8]),
    "get_groups": (C.c_int, [P, C.POINTER(vw_seg_info), PI64, PI32, PI32, PI32]),
}

DEVICE_ONLY = {
    "config_init": (C.c_int, [C.POINTER(vw_config), i32]),
    "viss_half": (C.c_int, [P, PD]),
    "viss_adjoint": (C.c_int, [P, PD, PD]),
    "run_all": (C.c_int, [P]),
    "set_day": (C.c_int, [P, i32]),
    "get_timings": (C.c_int, [P, C.POINTER(vw_timings)]),
    "get_accumulators": (C.c_int, [P, PD, PU32, PD, PU32, PI64]),
    "set_accumulators": (C.c_int, [P, PD, PU32, PD, PU32, i64]),
    "get_viss_counts": (C.c_int, [P, PI64]),
    "set_batch_targets": (C.c_int, [P, i64, i64]),
    "get_batch_targets": (C.c_int, [P, PI64, PI64]),
    "autotune": (C.c_int, [P, i32, PI64, PI64]),
    "run_all_graph": (C.c_int, [P]),
    "set_sky_schedules": (C.c_int, [P, C.POINTER(i32), C.POINTER(i32)]),
    "make_sky": (C.c_int, [C.c_uint64, i64, i32, PD]),
    "io_load_sky": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, i64, PD, C.c_char_p, i32]),
    "io_write_map": (C.c_int, [C.c_char_p, PD, i64, i32]),
    "io_write_viss": (C.c_int, [C.c_char_p, PD, i64, i32]),
    "io_write_cols3": (C.c_int, [C.c_char_p, PD, PD, PD, i64, i32]),
    "version": (C.c_char_p, []),
}

PF = C.POINTER(C.c_float)

# include/vw_unet.h (learned deconvolver)
UNET_FUNCS = {
    "unet_create": (C.c_int, [i32, i32, i32, C.c_uint64, C.POINTER(P)]),
    "unet_destroy": (None, [P]),
    "unet_last_error": (C.c_char_p, [P]),
    "unet_param_count": (C.c_int64, [P]),
    "unet_get_params": (C.c_int, [P, PF]),
    "unet_set_params": (C.c_int, [P, PF]),
    "unet_loss_grad": (C.c_int, [P, PF, PF, i32, PD, PF]),
    "unet_set_data": (C.c_int, [P, PF, PF, i64]),
    "unet_train": (C.c_int, [P, i64, i32, C.c_float, C.c_uint64, PF]),
    "unet_predict": (C.c_int, [P, PF, i32, PF]),
    "unet_predict_noalloc": (C.c_int, [P, PF, i32, PF]),
    "unet_train_ms": (C.c_int, [P, PD]),
}

ORACLE_ONLY = {
    "viss_half": (C.c_int, [P, i64, i64, PD, PI64]),
    "set_viss_raw": (C.c_int, [P, PD]),
    "group_segment": (C.c_int, [P, C.POINTER(vw_seg_info)]),
    "recon_points": (C.c_int, [P, i32, PI64, i64, PD, PU32]),
    "get_predicate": (C.c_int, [P, i64, i64, PU8]),
    "viss_adjoint": (C.c_int, [P, PD, PD]),
}


def bind(lib, prefix, table):
    for name, (res, args) in table.items():
        f = getattr(lib, prefix + name)
        f.restype = res
        f.argtypes = args
