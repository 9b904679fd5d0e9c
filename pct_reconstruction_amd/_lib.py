"""ctypes binding of lib/libpct.so (include/pct.h).

The library is loaded from this package's lib/ directory; there is no fallback: if it is missing
or was built for another architecture, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libpct.so"

PCT_OK = 0
ERRORS = {-1: "invalid argument", -2: "stage called out of order", -3: "HIP runtime error",
          -4: "file I/O or format error", -5: "allocation failure", -6: "size beyond a documented limit"}

HULL_IMPORT, HULL_SC, HULL_MSC, HULL_SM, HULL_FBP = 0, 1, 2, 3, 4
FILTER_RAM_LAK, FILTER_SHEPP_LOGAN, FILTER_NONE = 0, 1, 2
X0_IMPORT, X0_HULL, X0_FBP, X0_HYBRID, X0_ZEROS = 0, 1, 2, 3, 4
MLP_STANDARD, MLP_TABULATED = 0, 1
WRITE_FBP, WRITE_HULL, WRITE_X0, WRITE_X = 1, 2, 4, 8
ACCUM_FP32, ACCUM_FP64, ACCUM_KAHAN32 = 0, 1, 2
DROP_SPATIAL_ORDER, DROP_XCD_SLABS, DROP_BRICKS, DROP_COLS_LDS, DROP_COLS_SLICED, DROP_BUCKETS = 1, 2, 4, 8, 16, 32
DROP_DEFAULT_FLAGS = DROP_SPATIAL_ORDER | DROP_BRICKS | DROP_COLS_SLICED  # the context's default (internal.h)

IMG = {"x": 0, "fbp": 1, "x0": 2, "hull": 3, "sc_hull": 4, "msc_counts": 5, "sm_counts": 6, "msc_hull": 7,
       "sm_hull": 8, "fbp_hull": 9, "fbp_median": 10, "hull_filtered": 11}
BINS = {"counts_pre": 0, "mean_wepl": 1, "mean_ut": 2, "mean_uv": 3, "std_wepl": 4, "std_ut": 5, "std_uv": 6,
        "counts_post": 7, "sinogram": 8, "sinogram_filtered": 9}
HIST = {"x_entry": 0, "y_entry": 1, "z_entry": 2, "x_exit": 3, "y_exit": 4, "z_exit": 5, "xy_entry_angle": 6,
        "xz_entry_angle": 7, "xy_exit_angle": 8, "xz_exit_angle": 9, "wepl": 10, "bin_num": 11, "first_voxel": 12,
        "path_count": 13, "chord": 14, "gantry_angle": 15}
HIST_DTYPE = {"bin_num": np.int32, "first_voxel": np.uint32, "path_count": np.uint32, "chord": np.float64}
COUNTS = {"loaded": 0, "recon_volume": 1, "post_cuts": 2, "reconstruction": 3, "list_entries": 4, "current": 5}

_D, _I = C.c_double, C.c_int32


class Params(C.Structure):
    """pct_params (include/pct.h); field names follow the settings.cfg keys in lower case."""

    _fields_ = [
        ("gantry_angle_interval", _D), ("angular_bin_size", _D), ("ssd_t_size", _D), ("ssd_v_size", _D),
        ("t_shift", _D), ("u_shift", _D), ("v_shift", _D), ("t_bin_size", _D), ("v_bin_size", _D),
        ("sigmas_2_keep", _I), ("recon_cyl_radius", _D), ("columns", _I), ("rows", _I), ("slice_thickness", _D),
        ("msc_diff_thresh", _I), ("sc_lower_threshold", _D), ("sc_upper_threshold", _D),
        ("msc_lower_threshold", _D), ("msc_upper_threshold", _D), ("sm_lower_threshold", _D),
        ("sm_upper_threshold", _D), ("sm_scale_threshold", _D), ("sc_on", _I), ("msc_on", _I), ("sm_on", _I),
        ("fbp_on", _I), ("hull_type", _I), ("fbp_filter_type", _I), ("x_0_type", _I), ("fbp_threshold", _D),
        ("lambda_", _D), ("iterations", _I), ("drop_block_size", _I), ("histories_per_block", _I),
        ("min_mlp_length", _I), ("trig_table_min_pis", _D), ("trig_table_max_pis", _D),
        ("trig_table_step_degrees", _D), ("coeff_table_range", _D), ("coeff_table_step", _D),
        ("poly_table_range", _D), ("poly_table_step", _D), ("e_0", _D), ("x_0_rad_length", _D), ("rsp_air", _D),
        ("source_radius", _D), ("data_in_mm", _I), ("quirk_a1_bin0", _I), ("quirk_a3_drop_floor", _I),
        ("quirk_voxel_decode", _I), ("drop_accum", _I), ("median_filter_fbp", _I), ("fbp_med_filter_radius", _I),
        ("avg_filter_fbp", _I), ("avg_filter_hull", _I), ("hull_avg_filter_radius", _I), ("mlp_algorithm", _I),
        ("hull_avg_filter_threshold", _D),
        ("tvs_on", _I), ("tvs_first", _I), ("tvs_conditioned", _I), ("tvs_old", _I), ("tvs_repetitions", _I),
        ("tvs_l0", _I), ("tvs_seed", _I), ("list_segment_histories", _I), ("tvs_a", _D), ("tvs_min_eta", _D),
        ("tv_threshold", _D),
        # derived
        ("t_bins", _I), ("v_bins", _I), ("angular_bins", _I), ("num_bins", _I), ("slices", _I), ("num_voxels", _I),
        ("recon_cyl_height", _D), ("image_width", _D), ("image_height", _D), ("image_thickness", _D),
        ("voxel_width", _D), ("voxel_height", _D), ("voxel_thickness", _D), ("x_zero_coordinate", _D),
        ("y_zero_coordinate", _D), ("z_zero_coordinate", _D), ("mlp_u_step", _D), ("ram_lak_tau", _D), ("pi", _D),
        ("trig_table_min", _D), ("trig_table_max", _D), ("trig_table_step", _D), ("trig_table_elements", _I),
        ("coeff_table_elements", _I), ("poly_table_elements", _I),
    ]

    def copy(self) -> "Params":
        q = Params()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(Params))
        return q

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Object(C.Structure):
    _fields_ = [("kind", _I), ("cx", _D), ("cy", _D), ("cz", _D), ("a", _D), ("b", _D), ("c", _D), ("phi_deg", _D),
                ("rsp", _D)]


class Phantom(C.Structure):
    _fields_ = [("n_objects", _I), ("obj", Object * 16)]


class Beam(C.Structure):
    _fields_ = [("energy_mev", _D), ("u_planes", _D * 4), ("t_half_range", _D), ("v_half_range", _D),
                ("continuous_angles", _I), ("n_angles", _I), ("angle_interval", _D), ("wepl_noise_sigma", _D),
                ("outlier_fraction", _D), ("scatter", _I)]


_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run pct_reconstruction_amd/build.py (hipcc, gfx950)")
    L = C.CDLL(str(LIB_PATH), mode=C.RTLD_GLOBAL)
    vp, i32, i64, dbl, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_size_t
    P = C.POINTER
    fptr4 = C.c_void_p * 4
    sigs = {
        "pct_params_size": ([], sz),
        "pct_params_default": ([P(Params)], None),
        "pct_params_load_cfg": ([P(Params), C.c_char_p, C.c_char_p, sz], C.c_int),
        "pct_params_set": ([P(Params), C.c_char_p, C.c_char_p], C.c_int),
        "pct_params_derive": ([P(Params), C.c_char_p, sz], C.c_int),
        "pct_ctx_create": ([P(Params), C.c_int, P(vp)], C.c_int),
        "pct_ctx_destroy": ([vp], C.c_int),
        "pct_last_error": ([vp], C.c_char_p),
        "pct_ctx_synchronize": ([vp], C.c_int),
        "pct_get_count": ([vp, C.c_int, P(i64)], C.c_int),
        "pct_load_histories": ([vp, sz, fptr4, fptr4, fptr4, vp, vp], C.c_int),
        "pct_intersect_bin_carve": ([vp], C.c_int),
        "pct_hull_finish": ([vp], C.c_int),
        "pct_stats_and_cuts": ([vp], C.c_int),
        "pct_fbp": ([vp], C.c_int),
        "pct_select_hull": ([vp], C.c_int),
        "pct_initial_iterate": ([vp, vp], C.c_int),
        "pct_mlp_endpoints": ([vp], C.c_int),
        "pct_mlp_build_lists": ([vp], C.c_int),
        "pct_drop_iterate": ([vp, C.c_int], C.c_int),
        "pct_run_pipeline": ([vp], C.c_int),
        "pct_reset_iterate": ([vp], C.c_int),
        "pct_set_launch": ([vp, C.c_int, C.c_int, C.c_int], C.c_int),
        "pct_drop_profile": ([vp, C.c_int, P(dbl), P(dbl), P(i64), P(i64), P(i64), P(i64), P(i64)], C.c_int),
        "pct_get_image": ([vp, C.c_int, vp], C.c_int),
        "pct_set_image": ([vp, vp], C.c_int),
        "pct_set_hull": ([vp, vp], C.c_int),
        "pct_get_bins": ([vp, C.c_int, vp], C.c_int),
        "pct_get_histories": ([vp, C.c_int, vp, sz], C.c_int),
        "pct_get_lists": ([vp, vp, sz, vp, sz], C.c_int),
        "pct_residual": ([vp, P(dbl)], C.c_int),
        "pct_get_tvs_log": ([vp, C.c_size_t, vp, vp, vp, P(i64)], C.c_int),
        "pct_total_variation": ([vp, P(dbl)], C.c_int),
        "pct_export_preprocessed": ([vp, C.c_char_p, C.c_int], C.c_int),
        "pct_net_apply_iterate": ([vp, vp, P(C.c_int), P(dbl), P(dbl)], C.c_int),
        "pct_import_preprocessed": ([vp, C.c_char_p], C.c_int),
        "pct_metrics_json": ([vp], C.c_char_p),
        "pct_drop_iteration_device": ([i64, vp, vp, vp, vp, vp, vp, vp, i64, C.c_int, C.c_int, C.c_int, C.c_int,
                                       dbl, vp], C.c_int),
        "pct_write_reference_layout": ([vp, C.c_char_p, C.c_int], C.c_int),
        "pct_drop_iteration_device_noalloc": ([i64, vp, vp, vp, vp, vp, i64, C.c_int, C.c_int, C.c_int, C.c_int,
                                               dbl, vp], C.c_int),
        "pct_read_projection": ([C.c_char_p, P(i64), P(C.c_float), fptr4, fptr4, fptr4, vp, vp], C.c_int),
        "pct_write_projection": ([C.c_char_p, i64, C.c_float, C.c_float, fptr4, fptr4, fptr4, vp, vp], C.c_int),
        "pct_write_image_text": ([C.c_char_p, vp, C.c_int, C.c_int, C.c_int], C.c_int),
        "pct_write_image_dat": ([C.c_char_p, vp, C.c_int, C.c_int, C.c_int], C.c_int),
        "pct_write_image_png": ([C.c_char_p, vp, C.c_int, C.c_int, C.c_int], C.c_int),
        "pct_phantom_cylinder_c1": ([P(Phantom)], None),
        "pct_phantom_head": ([P(Phantom)], None),
        "pct_phantom_random": ([P(Phantom), C.c_uint64, C.c_int, dbl, dbl], None),
        "pct_generate_histories": ([P(Phantom), P(Beam), C.c_uint64, i64, fptr4, fptr4, fptr4, vp, vp], C.c_int),
        "pct_phantom_rasterize": ([P(Phantom), P(Params), vp], C.c_int),
        "pct_net_create": ([C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_int, P(vp)], C.c_int),
        "pct_net_destroy": ([vp], C.c_int),
        "pct_net_last_error": ([vp], C.c_char_p),
        "pct_net_num_params": ([vp], i64),
        "pct_net_get_params": ([vp, vp, i64], C.c_int),
        "pct_net_set_params": ([vp, vp, i64], C.c_int),
        "pct_net_loss_grad": ([vp, vp, vp, vp, P(dbl), vp], C.c_int),
        "pct_net_predict": ([vp, vp, vp, vp], C.c_int),
        "pct_net_add_sample": ([vp, vp, vp, vp], C.c_int),
        "pct_net_num_samples": ([vp], C.c_int),
        "pct_net_train": ([vp, C.c_int, dbl, C.c_uint64, P(dbl)], C.c_int),
        "pct_net_save": ([vp, C.c_char_p], C.c_int),
        "pct_net_load": ([vp, C.c_char_p], C.c_int),
        "pct_net_apply": ([vp, vp, P(C.c_int), P(dbl), P(dbl)], C.c_int),
    }
    for name, (argtypes, restype) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = argtypes
        fn.restype = restype
    if L.pct_params_size() != C.sizeof(Params):
        raise RuntimeError(f"pct_params layout mismatch: C {L.pct_params_size()} vs Python {C.sizeof(Params)}")
    _lib = L
    return L


def exported_symbols() -> list[str]:
    """Names of every function declared in include/pct.h (for the ABI test)."""
    import re
    hdr = (Path(__file__).resolve().parent.parent / "include" / "pct.h").read_text()
    body = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(pct_[a-z0-9_]+)\s*\(", body)))


def check(rc: int, ctx=None, what: str = "") -> None:
    if rc == PCT_OK:
        return
    msg = ERRORS.get(rc, f"error {rc}")
    if ctx is not None:
        detail = lib().pct_last_error(ctx)
        if detail:
            msg += ": " + detail.decode(errors="replace")
    raise RuntimeError(f"{what}: {msg}" if what else msg)


def ptr(a: np.ndarray) -> C.c_void_p:
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


def ptr4(arrs) -> "C.Array":
    return (C.c_void_p * 4)(*[a.ctypes.data for a in arrs])


def hip_device_count() -> int:
    """Number of HIP devices visible (0 without a GPU); does not initialise torch."""
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        return 0
    n = C.c_int(0)
    if hip.hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


os.environ.setdefault("HSA_ENABLE_IPC_MODE_LEGACY", "0")
