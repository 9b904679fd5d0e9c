"""Seeded synthetic phantoms and proton histories (host generator in lib/libpct.so)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Beam, Object, Params, Phantom, check, lib, ptr, ptr4


def phantom_cylinder() -> Phantom:
    """Water cylinder (R 80 mm, H 80 mm) with RSP 0.3 / 1.1 / 1.7 inserts (R 10 mm at r 40 mm)."""
    ph = Phantom()
    lib().pct_phantom_cylinder_c1(C.byref(ph))
    return ph


def phantom_head() -> Phantom:
    """10-ellipsoid head phantom, RSP 0.0-1.8."""
    ph = Phantom()
    lib().pct_phantom_head(C.byref(ph))
    return ph


def phantom_random(seed: int, n_ellipsoids: int, radius: float, half_height: float) -> Phantom:
    """Random ellipsoids inside a body ellipsoid; RSP uniform in [0, 1.8] (painter's order)."""
    ph = Phantom()
    lib().pct_phantom_random(C.byref(ph), int(seed), int(n_ellipsoids), float(radius), float(half_height))
    return ph


def phantom_from_objects(objects) -> Phantom:
    """objects: iterable of (kind, cx, cy, cz, a, b, c, phi_deg, rsp); kind 0 cylinder, 1 ellipsoid."""
    ph = Phantom()
    objs = list(objects)
    if len(objs) > 16:
        raise ValueError("at most 16 objects")
    ph.n_objects = len(objs)
    for k, o in enumerate(objs):
        ph.obj[k] = Object(*o)
    return ph


def beam(energy_mev: float = 200.0, u_planes=(-20.0, -15.0, 15.0, 20.0), t_half_range: float = 11.0,
         v_half_range: float = 4.5, continuous_angles: bool = False, n_angles: int = 90, angle_interval: float = 4.0,
         wepl_noise_sigma: float = 0.3, outlier_fraction: float = 0.01, scatter: bool = True) -> Beam:
    b = Beam()
    b.energy_mev = energy_mev
    for k in range(4):
        b.u_planes[k] = u_planes[k]
    b.t_half_range, b.v_half_range = t_half_range, v_half_range
    b.continuous_angles = int(continuous_angles)
    b.n_angles, b.angle_interval = n_angles, angle_interval
    b.wepl_noise_sigma, b.outlier_fraction, b.scatter = wepl_noise_sigma, outlier_fraction, int(scatter)
    return b


def generate(phantom: Phantom, bm: Beam, seed: int, n: int) -> dict:
    """n histories: dict with t, v, u of shape (4, n), wepl (n,), angle_deg (n,), all float32 (cm, deg)."""
    t = np.empty((4, n), np.float32)
    v = np.empty((4, n), np.float32)
    u = np.empty((4, n), np.float32)
    w = np.empty(n, np.float32)
    a = np.empty(n, np.float32)
    check(lib().pct_generate_histories(C.byref(phantom), C.byref(bm), int(seed), int(n), ptr4(t), ptr4(v), ptr4(u),
                                       ptr(w), ptr(a)), what="pct_generate_histories")
    return {"t": t, "v": v, "u": u, "wepl": w, "angle_deg": a}


def rasterize(phantom: Phantom, params: Params) -> np.ndarray:
    """RSP at the voxel centres of the reconstruction grid, shape (slices, rows, columns)."""
    out = np.empty(params.num_voxels, np.float32)
    check(lib().pct_phantom_rasterize(C.byref(phantom), C.byref(params), ptr(out)), what="pct_phantom_rasterize")
    return out.reshape(params.slices, params.rows, params.columns)
