"""Host-side mirror of the reference's reconstruction stages over the C-ABI.

The reference drives its path from two functions, preprocessing() and image_reconstruction()
(src/pCT_Reconstruction_Data_Segments_Blake.cu:7394, :10650), each a fixed sequence of stage
functions over global arrays.  `Reconstruction` keeps those stage names and that order; each method
calls one C-ABI entry point of lib/libpct.so, whose kernels do the work on the GPU.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import Params, check, lib, ptr, ptr4


def default_params(**overrides) -> Params:
    """Defaults of the reference's `configurations` constructor, then `overrides` (field names)."""
    p = Params()
    lib().pct_params_default(C.byref(p))
    for k, v in overrides.items():
        setattr(p, "lambda_" if k == "lambda" else k, v)
    return p


def load_settings(path: str | Path, base: Params | None = None) -> Params:
    """Parse a settings.cfg file (typing of read_key_value_pair, src/pCT_Reconstruction.cu:9275)."""
    p = base.copy() if base is not None else default_params()
    err = C.create_string_buffer(512)
    rc = lib().pct_params_load_cfg(C.byref(p), str(path).encode(), err, len(err))
    if rc != _lib.PCT_OK:
        raise ValueError(err.value.decode() or _lib.ERRORS.get(rc, str(rc)))
    return p


def set_key(p: Params, key: str, value) -> None:
    """Set one settings.cfg key (e.g. "HULL_TYPE", 1)."""
    if isinstance(value, bool):
        value = "true" if value else "false"
    rc = lib().pct_params_set(C.byref(p), key.encode(), str(value).encode())
    if rc != _lib.PCT_OK:
        raise KeyError(key)


def derive(p: Params) -> Params:
    """Derived quantities (include/Configurations.h:1745-1789) on a copy of p."""
    q = p.copy()
    err = C.create_string_buffer(512)
    rc = lib().pct_params_derive(C.byref(q), err, len(err))
    if rc != _lib.PCT_OK:
        raise ValueError(err.value.decode())
    return q


class Reconstruction:
    """One reconstruction on one GPU: holds the histories, bins, images and MLP lists in HBM."""

    def __init__(self, params: Params, device: int = 0):
        self.params = derive(params)
        self._ctx = C.c_void_p()
        rc = lib().pct_ctx_create(C.byref(self.params), device, C.byref(self._ctx))
        if rc != _lib.PCT_OK:
            raise RuntimeError(f"pct_ctx_create failed: {_lib.ERRORS.get(rc, rc)} (is a HIP device visible?)")

    # -- lifetime -------------------------------------------------------------------------------
    def close(self) -> None:
        if self._ctx:
            lib().pct_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name: str, *args) -> None:
        check(getattr(lib(), name)(self._ctx, *args), self._ctx, name)

    # -- input ----------------------------------------------------------------------------------
    def load_histories(self, t: np.ndarray, v: np.ndarray, u: np.ndarray, wepl: np.ndarray,
                       angle_deg: np.ndarray) -> None:
        """t, v, u: (4, n) float32 tracker coordinates (PCTD order); wepl, angle_deg: (n,) float32."""
        t = [np.ascontiguousarray(a, dtype=np.float32) for a in t]
        v = [np.ascontiguousarray(a, dtype=np.float32) for a in v]
        u = [np.ascontiguousarray(a, dtype=np.float32) for a in u]
        w = np.ascontiguousarray(wepl, dtype=np.float32)
        ang = np.ascontiguousarray(angle_deg, dtype=np.float32)
        n = w.shape[0]
        for a in (*t, *v, *u, ang):
            if a.shape != (n,):
                raise ValueError("all history arrays must have the same length")
        self._call("pct_load_histories", n, ptr4(t), ptr4(v), ptr4(u), ptr(w), ptr(ang))

    def load_projection_files(self, paths) -> None:
        """Read PCTD/PCTN projection files (read_data_chunk_v0) and load their histories."""
        from .io import read_projection
        for path in paths:
            d = read_projection(path)
            self.load_histories(d["t"], d["v"], d["u"], d["wepl"], d["angle_deg"])

    # -- stages, in the reference's order --------------------------------------------------------
    def import_and_process_data(self) -> None:
        """recon_volume_intersections + binning + hull-detection counts (DS.cu:5554)."""
        self._call("pct_intersect_bin_carve")

    def hull_detection_finish(self) -> None:
        """MSC / SM edge detection (DS.cu:7276)."""
        self._call("pct_hull_finish")

    def statistical_calculations_and_cuts(self) -> None:
        """Per-bin mean / sigma and the SIGMAS_2_KEEP cuts (DS.cu:5972)."""
        self._call("pct_stats_and_cuts")

    def generate_preprocessing_images(self, x0: np.ndarray | None = None) -> None:
        """construct_sinogram + FBP + hull_selection + define_initial_iterate (DS.cu:8647)."""
        self._call("pct_fbp")
        self._call("pct_select_hull")
        buf = None
        if x0 is not None:
            buf = np.ascontiguousarray(x0, dtype=np.float32).reshape(-1)
        self._call("pct_initial_iterate", ptr(buf) if buf is not None else None)

    def preprocessing(self, x0: np.ndarray | None = None) -> None:
        """The whole preprocessing() of the reference (DS.cu:7394-7403)."""
        self.import_and_process_data()
        self.hull_detection_finish()
        self.statistical_calculations_and_cuts()
        self.generate_preprocessing_images(x0)

    def reconstruction_cuts(self) -> None:
        """MLP endpoints on the hull and removal of histories that miss it (DS.cu:8296)."""
        self._call("pct_mlp_endpoints")

    def build_mlp_lists(self) -> None:
        """MLP voxel lists of every reconstruction history (find_MLP_path_GPU_tabulated, DS.cu:9104)."""
        self._call("pct_mlp_build_lists")

    def DROP(self, iterations: int | None = None, write_dir=None) -> None:
        """DROP iterations over the lists (DROP_GPU, DS.cu:9848).  With write_dir, every iterate is written
        there as x_<k>.txt/.dat/.png after its iteration (WRITE_X_KI, DS.cu:9915-9922)."""
        its = self.params.iterations if iterations is None else iterations
        if write_dir is None:
            self._call("pct_drop_iterate", int(its))
            return
        for _ in range(int(its)):
            self._call("pct_drop_iterate", 1)
            self.write_reference_layout(write_dir, _lib.WRITE_X)

    def write_reference_layout(self, directory, flags: int = 15) -> None:
        """FBP.txt, hull.*, x_0.* and x_<k>.* in the reference's output names (pct_write_reference_layout)."""
        from pathlib import Path
        Path(directory).mkdir(parents=True, exist_ok=True)
        self._call("pct_write_reference_layout", str(directory).encode(), int(flags))

    def image_reconstruction(self, iterations: int | None = None) -> None:
        """reconstruction_cuts + MLP + DROP (DS.cu:10650)."""
        self.reconstruction_cuts()
        self.build_mlp_lists()
        self.DROP(iterations)

    def run(self, x0: np.ndarray | None = None) -> np.ndarray:
        self.preprocessing(x0)
        self.image_reconstruction()
        return self.image("x")

    def synchronize(self) -> None:
        self._call("pct_ctx_synchronize")

    def set_launch(self, drop_row_lanes: int = 16, drop_col_voxels: int = 1, spatial_order: bool = True,
                   xcd_slabs: bool = False, bricks: bool = True, cols_lds: bool = False,
                   cols_sliced: bool = True, buckets: bool = False) -> None:
        """Launch shape and scheduling of the DROP kernels (pct_set_launch); results do not change.
        The defaults are the library's."""
        flags = ((_lib.DROP_SPATIAL_ORDER if spatial_order else 0) | (_lib.DROP_XCD_SLABS if xcd_slabs else 0)
                 | (_lib.DROP_BRICKS if bricks else 0) | (_lib.DROP_COLS_LDS if cols_lds else 0)
                 | (_lib.DROP_COLS_SLICED if cols_sliced else 0) | (_lib.DROP_BUCKETS if buckets else 0))
        self._call("pct_set_launch", int(drop_row_lanes), int(drop_col_voxels), flags)

    def reset_iterate(self) -> None:
        """x <- x0 on the device."""
        self._call("pct_reset_iterate")

    def profile_drop(self, iterations: int = 1) -> dict:
        """DROP iterations with per-launch HIP-event timing (see pct_drop_profile)."""
        ro, co = C.c_double(), C.c_double()
        la, hi, ga, sa, tu = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        self._call("pct_drop_profile", int(iterations), C.byref(ro), C.byref(co), C.byref(la), C.byref(hi),
                   C.byref(ga), C.byref(sa), C.byref(tu))
        return {"rows_ms": ro.value, "cols_ms": co.value, "blocks": la.value, "histories": hi.value,
                "gathered": ga.value, "summed": sa.value, "touched": tu.value}

    # -- outputs ----------------------------------------------------------------------------------
    @property
    def shape(self) -> tuple[int, int, int]:
        p = self.params
        return (p.slices, p.rows, p.columns)

    def image(self, kind: str = "x") -> np.ndarray:
        out = np.empty(self.params.num_voxels, dtype=np.float32)
        self._call("pct_get_image", _lib.IMG[kind], ptr(out))
        return out.reshape(self.shape)

    def set_image(self, x: np.ndarray) -> None:
        buf = np.ascontiguousarray(x, dtype=np.float32).reshape(-1)
        if buf.size != self.params.num_voxels:
            raise ValueError("image size mismatch")
        self._call("pct_set_image", ptr(buf))

    def set_hull(self, hull: np.ndarray) -> None:
        buf = np.ascontiguousarray(hull, dtype=np.uint8).reshape(-1)
        if buf.size != self.params.num_voxels:
            raise ValueError("hull size mismatch")
        self._call("pct_set_hull", ptr(buf))

    def bins(self, kind: str) -> np.ndarray:
        out = np.empty(self.params.num_bins, dtype=np.float32)
        self._call("pct_get_bins", _lib.BINS[kind], ptr(out))
        if kind.startswith("counts"):
            return out.astype(np.int32)
        return out

    def count(self, kind: str) -> int:
        n = C.c_int64()
        self._call("pct_get_count", _lib.COUNTS[kind], C.byref(n))
        return n.value

    def histories(self, kind: str) -> np.ndarray:
        """A per-history array of the histories held at the current stage."""
        n = self.count("current")
        out = np.empty(n, dtype=_lib.HIST_DTYPE.get(kind, np.float32))
        self._call("pct_get_histories", _lib.HIST[kind], ptr(out), n)
        return out

    def lists(self) -> tuple[np.ndarray, np.ndarray]:
        """(offsets[n+1], voxels) of the MLP lists in history order."""
        n = self.count("reconstruction")
        off = np.zeros(n + 1, dtype=np.int64)
        self._call("pct_get_lists", ptr(off), n + 1, None, 0)
        vox = np.empty(int(off[-1]), dtype=np.uint32)
        self._call("pct_get_lists", ptr(off), n + 1, ptr(vox) if vox.size else None, vox.size)
        return off, vox

    def residual(self) -> float:
        r = C.c_double()
        self._call("pct_residual", C.byref(r))
        return r.value

    def export_preprocessed(self, directory, exact_chords: bool = True) -> None:
        """accumulate_and_export_preprocessed_data (V.cu:6613): path lists, per-history data, hull, x0."""
        Path(directory).mkdir(parents=True, exist_ok=True)
        self._call("pct_export_preprocessed", str(directory).encode(), 1 if exact_chords else 0)

    def import_preprocessed(self, directory) -> None:
        """init_import_preprocessed_data (V.cu:6775): ready for DROP from the exported files."""
        self._call("pct_import_preprocessed", str(directory).encode())

    def total_variation(self) -> float:
        """TV of the current iterate (calculate_total_variation, DS.cu:10165), fixed-order FP64 sum."""
        r = C.c_double()
        self._call("pct_total_variation", C.byref(r))
        return r.value

    def tvs_log(self) -> list:
        """Per NTVS iteration since the initial iterate: (TV before, TV after, L after)."""
        n = C.c_int64()
        self._call("pct_get_tvs_log", 0, None, None, None, C.byref(n))
        k = n.value
        tb, ta = np.zeros(k, np.float64), np.zeros(k, np.float64)
        L = np.zeros(k, np.int32)
        self._call("pct_get_tvs_log", k, tb.ctypes.data_as(C.c_void_p), ta.ctypes.data_as(C.c_void_p),
                   L.ctypes.data_as(C.c_void_p), C.byref(n))
        return [(float(tb[i]), float(ta[i]), int(L[i])) for i in range(k)]

    def metrics(self) -> dict:
        return json.loads(lib().pct_metrics_json(self._ctx).decode())
