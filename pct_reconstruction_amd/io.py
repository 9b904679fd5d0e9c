"""Projection files (PCTD/PCTN) and image writers in the reference's formats."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from ._lib import check, lib, ptr, ptr4


def read_projection(path: str | Path) -> dict:
    """Read a projection file (read_data_chunk_v0, DS.cu:4361).  Values are as stored (mm for the
    reference's data; set DATA_IN_MM to convert)."""
    n = C.c_int64()
    ang = C.c_float()
    null4 = (C.c_void_p * 4)()
    check(lib().pct_read_projection(str(path).encode(), C.byref(n), C.byref(ang), null4, null4, null4, None, None),
          what=f"read {path}")
    n = n.value
    t = np.empty((4, n), np.float32)
    v = np.empty((4, n), np.float32)
    u = np.empty((4, n), np.float32)
    w = np.empty(n, np.float32)
    a = np.empty(n, np.float32)
    check(lib().pct_read_projection(str(path).encode(), C.byref(C.c_int64()), C.byref(ang), ptr4(t), ptr4(v),
                                    ptr4(u), ptr(w), ptr(a)), what=f"read {path}")
    return {"t": t, "v": v, "u": u, "wepl": w, "angle_deg": a, "header_angle_deg": ang.value}


def write_projection(path: str | Path, t, v, u, wepl, angle_deg: float, energy_mev: float = 200.0,
                     angles: np.ndarray | None = None) -> None:
    """Write a PCTD file (version 0) or, with per-history angles, a PCTN version-1 file."""
    t = [np.ascontiguousarray(x, np.float32) for x in t]
    v = [np.ascontiguousarray(x, np.float32) for x in v]
    u = [np.ascontiguousarray(x, np.float32) for x in u]
    w = np.ascontiguousarray(wepl, np.float32)
    a = None if angles is None else np.ascontiguousarray(angles, np.float32)
    check(lib().pct_write_projection(str(path).encode(), w.shape[0], float(angle_deg), float(energy_mev), ptr4(t),
                                     ptr4(v), ptr4(u), ptr(w), ptr(a) if a is not None else None),
          what=f"write {path}")


def write_image_text(path: str | Path, image: np.ndarray) -> None:
    """array_2_disk layout (DS.cu:10743): COLUMNS values per line, ROWS*SLICES lines."""
    img = np.ascontiguousarray(image, np.float32)
    s, r, c = img.shape
    check(lib().pct_write_image_text(str(path).encode(), ptr(img), c, r, s), what=f"write {path}")


def write_image_dat(path: str | Path, image: np.ndarray) -> None:
    """write_PNG's .dat (DS.cu:9939): float32 of value/2."""
    img = np.ascontiguousarray(image, np.float32)
    s, r, c = img.shape
    check(lib().pct_write_image_dat(str(path).encode(), ptr(img), c, r, s), what=f"write {path}")


def write_image_png(path: str | Path, image: np.ndarray) -> None:
    """write_PNG's PNG (DS.cu:9969): 16-bit gray, COLUMNS x ROWS*SLICES, round(clamp(value/2, 0, 1) * 65535)."""
    img = np.ascontiguousarray(image, np.float32)
    s, r, c = img.shape
    check(lib().pct_write_image_png(str(path).encode(), ptr(img), c, r, s), what=f"write {path}")
