"""Parity of the selectable stage options with the oracle, stage by stage (same bar as test_gpu_parity):

- HULL_TYPE: IMPORT (a caller's hull, import_hull V.cu:6659), SM (hull_selection DS.cu:7351-7393 on the
  SM counts) and FBP (FBP_image_2_hull_GPU DS.cu:6592-6603);
- X_0_TYPE: IMPORT, HULL and ZEROS (define_initial_iterate DS.cu:8612-8646);
- DATA_IN_MM with non-zero t/u/v shifts, the histories entering through projection files
  (read_data_chunk_v0 DS.cu:4361-4467, convert_mm_2_cm :4070, apply_tuv_shifts :4090).
"""
from __future__ import annotations

import numpy as np
import pytest

from cases import small_cylinder, small_head
from test_gpu_parity import _run_and_compare

pytestmark = pytest.mark.gpu


def _cylinder_mask(P, radius_cm):
    C, R, S = P.columns, P.rows, P.slices
    x = -P.recon_cyl_radius + (np.arange(C) + 0.5) * P.voxel_width
    y = P.recon_cyl_radius - (np.arange(R) + 0.5) * P.voxel_height
    r2 = x[None, :] ** 2 + y[:, None] ** 2
    return np.broadcast_to(r2 <= radius_cm ** 2, (S, R, C)).astype(np.uint8).reshape(-1)


def test_parity_import_hull_and_import_x0(oracle):
    P, data = small_cylinder(n=40000, seed=31, hull_type=0, x_0_type=0)
    hull = _cylinder_mask(P, 8.6)
    x0 = np.random.default_rng(7).uniform(0.0, 1.2, P.num_voxels).astype(np.float32)
    _run_and_compare(P, data, oracle, "import-hull-import-x0", hull_import=hull, x0_import=x0)


def test_parity_sm_hull_with_hull_x0(oracle):
    P, data = small_head(n=50000, seed=13, hull_type=3, x_0_type=1)
    _run_and_compare(P, data, oracle, "head-SM-hull-x0")


def test_parity_fbp_hull_with_zero_x0(oracle):
    P, data = small_cylinder(n=50000, seed=17, hull_type=4, x_0_type=4)
    _run_and_compare(P, data, oracle, "cylinder-FBP-hull-zeros")


def test_parity_mm_input_with_shifts_through_projection_files(oracle, tmp_path):
    from pct_reconstruction_amd import io
    P, data = small_cylinder(n=36000, seed=23, data_in_mm=1, t_shift=0.05, u_shift=-0.03, v_shift=0.02)
    order = np.argsort(data["angle_deg"], kind="stable")
    mm = {k: (np.asarray(data[k], np.float32)[..., order] * np.float32(10)).astype(np.float32)
          for k in ("t", "v", "u", "wepl")}
    mm["angle_deg"] = np.asarray(data["angle_deg"], np.float32)[order]
    files = []
    for a in np.unique(mm["angle_deg"]):
        sel = mm["angle_deg"] == a
        f = tmp_path / f"projection_{int(a):03d}.bin"
        io.write_projection(f, mm["t"][:, sel], mm["v"][:, sel], mm["u"][:, sel], mm["wepl"][sel], angle_deg=float(a))
        files.append(f)
    assert np.array_equal(mm["angle_deg"], np.trunc(mm["angle_deg"]))  # integer angles: PCTD v0 files
    _run_and_compare(P, mm, oracle, "mm-shifts-files", files=files)


def test_parity_untabulated_mlp(oracle):
    """MLP_ALGORITHM = STANDARD (find_MLP_path_GPU, DS.cu:8895-9103): polynomial Sigma terms, log() scattering
    coefficients and cos()/sin() rotation per history; lists, chords and the DROP image against the oracle."""
    P, data = small_head(n=40000, seed=29, mlp_algorithm=0)
    _run_and_compare(P, data, oracle, "head-untabulated-MLP")


def test_reference_layout_files(tmp_path):
    """FBP.txt, hull.*, x_0.* and x_<k>.* after every DROP iteration, in the reference's names and layouts
    (array_2_disk DS.cu:10743, write_PNG :9939; call sites :6112, :7365, :8637, :9915)."""
    from pct_reconstruction_amd import Reconstruction, io
    P, data = small_cylinder(n=20000, seed=5, iterations=2)
    with Reconstruction(P) as r:
        r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
        r.preprocessing()
        r.write_reference_layout(tmp_path)
        r.reconstruction_cuts()
        r.build_mlp_lists()
        r.DROP(2, write_dir=tmp_path)
        x, x0, hull = r.image("x"), r.image("x0"), r.image("hull")
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == sorted(["FBP.txt", "hull.txt", "hull.dat", "hull.png", "x_0.txt", "x_0.dat", "x_0.png",
                            "x_1.txt", "x_1.dat", "x_1.png", "x_2.txt", "x_2.dat", "x_2.png"])
    io.write_image_text(tmp_path / "check.txt", x)
    assert (tmp_path / "x_2.txt").read_bytes() == (tmp_path / "check.txt").read_bytes()
    assert np.array_equal(np.fromfile(tmp_path / "x_0.dat", np.float32), (x0 / 2).reshape(-1))
    lines = (tmp_path / "hull.txt").read_text().splitlines()
    assert len(lines) == P.rows * P.slices
    assert np.array_equal(np.array([ln.split() for ln in lines], float).reshape(-1), hull.reshape(-1))


@pytest.mark.parametrize("table,block,spill", [("0", "256", "512"), ("4096", "512", "1"), ("8192", "256", "512"),
                                               ("8192", "512", "512"), ("16384", "256", "3")])
def test_carve_variants_count_identically(oracle, monkeypatch, table, block, spill):
    """The hull counts of the fused, bin-ordered carve (LDS hash table of 4096/8192/16384 slots for 256 or 512
    lanes, visits that find no slot or reach the spill count going to global atomics) and of one walk per
    method (table 0) are the oracle's, bit for bit (hull_detection, DS.cu:6607; SC_GPU :6681, MSC_GPU :6799,
    SM_GPU :6997)."""
    from pct_reconstruction_amd import Reconstruction
    monkeypatch.setenv("PCT_CARVE_TABLE", table)
    monkeypatch.setenv("PCT_CARVE_BLOCK", block)
    monkeypatch.setenv("PCT_CARVE_SPILL", spill)
    P, data = small_head(n=60000, seed=37, sc_on=1, msc_on=1, sm_on=1)
    ref = oracle.run_pipeline(P, data, stop_after="cuts")
    with Reconstruction(P) as r:
        r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
        r.import_and_process_data()
        print(f"[carve table={table} block={block} spill={spill}] {r.count('recon_volume')} histories", flush=True)
        assert np.array_equal(r.image("sc_hull").reshape(-1), ref["sc_hull"])
        assert np.array_equal(r.image("msc_counts").reshape(-1), ref["msc_counts"])
        assert np.array_equal(r.image("sm_counts").reshape(-1), ref["sm_counts"])
        assert r.image("sm_counts").max() > 40  # many visits per voxel: the table aggregates
        r.hull_detection_finish()
        r.statistical_calculations_and_cuts()  # reuses the bin order of the carve
        assert np.array_equal(r.bins("counts_post"), ref["counts_post"])
        assert np.array_equal(r.bins("sinogram"), ref["sinogram_sum"])
        assert r.count("post_cuts") == ref["h_cut"].shape[1]


@pytest.mark.parametrize("group_entries,launch", [("1073741824", {}), ("150000", {}),
                                                  ("150000", {"bricks": False, "spatial_order": False})])
def test_pipelined_list_build_matches_the_serial_one(monkeypatch, group_entries, launch):
    """The MLP lists built group by group on a second stream while earlier groups are transposed give the
    same lists, the same DROP image and the same residual as the lists built before the plan.  Blocks of
    3200 histories at 700 per launch leave unprocessed tails in every block and a last block of 527
    histories that joins no group (quirk A3); at 150000 entries per group there are many groups.
    2 iterations (DROP_GPU, DS.cu:9848)."""
    from pct_reconstruction_amd import Reconstruction
    P, data = small_head(n=50000, seed=53, drop_block_size=3200, histories_per_block=700, iterations=2)
    monkeypatch.setenv("PCT_DROP_GROUP_ENTRIES", group_entries)
    out = {}
    for pipe in ("0", "1"):
        monkeypatch.setenv("PCT_DROP_PIPELINE", pipe)
        with Reconstruction(P) as r:
            if launch:
                r.set_launch(**launch)
            r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
            r.preprocessing()
            r.reconstruction_cuts()
            r.build_mlp_lists()
            off, vox = r.lists()
            r.DROP(2)
            out[pipe] = dict(off=off, vox=vox, chord=r.histories("chord"), x=r.image("x").reshape(-1),
                             res=r.residual(), entries=r.count("list_entries"), n=r.count("reconstruction"))
    a, b = out["0"], out["1"]
    print(f"[pipelined lists, group entries {group_entries}, launch {launch}] {a['n']} histories, {a['entries']} entries, "
          f"last block {a['n'] % 3200} histories", flush=True)
    assert a["n"] % 3200 < 700  # the last block's histories are not processed (quirk A3) and join no group
    for k in ("off", "vox", "chord", "x"):
        assert np.array_equal(a[k], b[k]), k
    assert a["entries"] == b["entries"] and a["res"] == b["res"]
