"""Workload presets: parameters, phantom and beam of the benchmark configurations.

  cylinder-2M   water cylinder + 3 inserts, 2M histories, 90 x 4 deg, 200x200x32 @ 1x1x2.5 mm,
                SC hull, Shepp-Logan FBP x0, 12 DROP iterations, block 3200, lambda 1e-4.
  head-50M      10-ellipsoid head, 50M histories, continuous angles, 256x256x96 @ ~1 mm, MSC + SM,
                3 sigma cuts, FBP x0, 20 iterations, block 60000.
  learned-128   random ellipsoid phantoms on 128x128x64 @ 2 mm (training pairs of the warm start).
  learned-head50m random ellipsoid phantoms with the head-50M geometry, beam and histories.
  large-200M    head, 200M histories at 230 MeV, 512x512x128 @ 0.5x0.5x1 mm, SM hull, block 120000.
"""
from __future__ import annotations

from . import synth
from .recon import default_params, derive


def cylinder_2m(n: int = 2_000_000):
    p = default_params(recon_cyl_radius=10.0, columns=200, rows=200, slice_thickness=0.25, ssd_v_size=9.0,
                       ssd_t_size=35.0, t_bin_size=0.1, v_bin_size=0.25, angular_bin_size=4.0, sigmas_2_keep=3,
                       sc_on=1, msc_on=0, sm_on=0, hull_type=1, fbp_filter_type=1, x_0_type=2, iterations=12,
                       drop_block_size=3200, lambda_=1e-4)
    ph = synth.phantom_cylinder()
    bm = synth.beam(energy_mev=200.0, t_half_range=11.0, v_half_range=4.5, continuous_angles=False, n_angles=90,
                    angle_interval=4.0, wepl_noise_sigma=0.3, outlier_fraction=0.01)
    return derive(p), ph, bm, n


def head_50m(n: int = 50_000_000):
    # SSD_V_SIZE 10.61 -> RECON_CYL_HEIGHT 9.61 cm -> int(9.61/0.1) = 96 slices of 0.1001 cm
    p = default_params(recon_cyl_radius=12.8, columns=256, rows=256, slice_thickness=0.1, ssd_v_size=10.61,
                       ssd_t_size=35.0, t_bin_size=0.1, v_bin_size=0.25, angular_bin_size=4.0, sigmas_2_keep=3,
                       sc_on=0, msc_on=1, sm_on=1, hull_type=2, fbp_filter_type=1, x_0_type=2, iterations=20,
                       drop_block_size=60000, lambda_=1e-4)
    ph = synth.phantom_head()
    bm = synth.beam(energy_mev=200.0, t_half_range=11.0, v_half_range=4.5, continuous_angles=True,
                    wepl_noise_sigma=0.25, outlier_fraction=0.01)
    return derive(p), ph, bm, n


def learned_128(seed: int, n: int = 10_000_000):
    p = default_params(recon_cyl_radius=12.8, columns=128, rows=128, slice_thickness=0.2, ssd_v_size=13.81,
                       ssd_t_size=35.0, t_bin_size=0.1, v_bin_size=0.25, angular_bin_size=4.0, sigmas_2_keep=3,
                       sc_on=0, msc_on=1, sm_on=0, hull_type=2, fbp_filter_type=1, x_0_type=2, iterations=40,
                       drop_block_size=60000, lambda_=1e-4)
    q = derive(p)
    ph = synth.phantom_random(seed, 6 + seed % 9, 10.0, q.recon_cyl_height / 2)
    bm = synth.beam(energy_mev=200.0, t_half_range=13.0, v_half_range=6.0, continuous_angles=True,
                    wepl_noise_sigma=0.25, outlier_fraction=0.01)
    return q, ph, bm, n


def learned_head50m(seed: int, n: int = 50_000_000):
    """Training phantoms for a warm start at the head-50M scale: the head-50M parameters and beam with a
    random ellipsoid phantom (6-14 ellipsoids, RSP 0-1.8)."""
    q, _, bm, n = head_50m(n)
    ph = synth.phantom_random(seed, 6 + seed % 9, 10.0, q.recon_cyl_height / 2)
    return q, ph, bm, n


def large_200m(n: int = 200_000_000):
    p = default_params(recon_cyl_radius=12.8, columns=512, rows=512, slice_thickness=0.1, ssd_v_size=13.81,
                       ssd_t_size=35.0, t_bin_size=0.1, v_bin_size=0.25, angular_bin_size=4.0, sigmas_2_keep=3,
                       sc_on=0, msc_on=0, sm_on=1, hull_type=3, fbp_filter_type=1, x_0_type=2, iterations=12,
                       drop_block_size=120000, lambda_=1e-4)
    ph = synth.phantom_head()
    bm = synth.beam(energy_mev=230.0, t_half_range=11.0, v_half_range=6.0, continuous_angles=True,
                    wepl_noise_sigma=0.25, outlier_fraction=0.01)
    return derive(p), ph, bm, n


PRESETS = {"cylinder-2M": cylinder_2m, "head-50M": head_50m, "large-200M": large_200m}
