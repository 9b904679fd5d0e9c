"""Parity at the full size of one large-200M segment (25M histories, 512x512x128 at 0.5x0.5x1 mm,
230 MeV, SM hull, 12 DROP iterations, block 120000).  This grid needs 25-bit voxel indices and the
blocks carry about 4x the entries of a head-50M block, so the transposition keys, groups and column
slices run at sizes the smaller cases do not reach:

  - intersections and bin numbers of all 25M histories, oracle vs library, bit for bit;
  - MLP voxel lists and chords of 3000 sampled reconstruction histories, recomputed by the oracle
    from the library's own endpoints, bit for bit;
  - list invariants: counts sum to the entries, every index is a voxel, every list starts at its
    entry voxel;
  - DROP: the data residual falls with every iteration, a rerun gives the same bits, and the
    unsliced column layout and the row-major schedule give the same bits as the default.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H_NAMES = ["x_entry", "y_entry", "z_entry", "x_exit", "y_exit", "z_exit", "xy_entry_angle", "xz_entry_angle",
           "xy_exit_angle", "xz_exit_angle"]


@pytest.mark.timeout(900)
def test_large_segment_full_size_parity_and_invariants(oracle):
    from pct_reconstruction_amd import Reconstruction, configs, synth

    P, ph, bm, n = configs.large_200m(25_000_000)
    assert P.num_voxels > 1 << 24  # voxel indices wider than the head grid's
    data = synth.generate(ph, bm, 2024, n)
    print(f"[segment] {n} histories generated, grid {P.columns}x{P.rows}x{P.slices}", flush=True)
    with Reconstruction(P) as r:
        r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
        r.import_and_process_data()
        t, v, u, _ = oracle.convert_units(P, data["t"], data["v"], data["u"], data["wepl"])
        h, missed = oracle.intersections(P, t, v, u, data["angle_deg"])
        keep = ~missed
        bins, _ = oracle.binning(P, h)
        assert r.count("recon_volume") == int(keep.sum())
        assert np.array_equal(r.histories("bin_num"), bins[keep])
        sample = np.sort(np.random.default_rng(3).choice(int(keep.sum()), 100000, replace=False))
        for k, name in enumerate(H_NAMES):
            assert np.array_equal(r.histories(name)[sample], h[k][keep][sample]), name
        del h, bins, t, v, u, data
        print("[segment] intersections and bins identical", flush=True)

        r.hull_detection_finish()
        r.statistical_calculations_and_cuts()
        r.generate_preprocessing_images()
        r.reconstruction_cuts()
        r.build_mlp_lists()
        m = r.count("reconstruction")
        off, vox = r.lists()
        cnt = np.diff(off)
        first = r.histories("first_voxel")
        assert off[-1] == r.count("list_entries") == int(cnt.sum())
        assert int(vox.max()) < P.num_voxels and int(cnt.min()) >= 1
        assert np.array_equal(vox[off[:-1]], first)
        he = np.stack([r.histories(name) for name in H_NAMES])
        chord = r.histories("chord")
        idx = np.sort(np.random.default_rng(4).choice(m, 3000, replace=False))
        tabs = oracle.MlpTables(P)
        o_off, o_vox, o_chord = oracle.mlp_lists(P, tabs, np.ascontiguousarray(he[:, idx]), first[idx])
        for j, i in enumerate(idx):
            assert np.array_equal(vox[off[i]:off[i + 1]], o_vox[o_off[j]:o_off[j + 1]]), i
        assert np.array_equal(chord[idx], o_chord)
        print(f"[segment] {m} reconstruction histories, {off[-1]} entries; lists of 3000 sampled histories "
              f"identical to the oracle", flush=True)
        del off, vox, cnt, he

        res = [r.residual()]
        for _ in range(2):
            r.DROP(1)
            res.append(r.residual())
        x2 = r.image("x").reshape(-1)
        assert all(b < a for a, b in zip(res, res[1:])), res
        r.reset_iterate()
        r.DROP(2)
        assert np.array_equal(r.image("x").reshape(-1), x2)
        for launch in (dict(cols_sliced=False), dict(spatial_order=False, bricks=False), dict(buckets=True)):
            r.set_launch(**launch)
            r.reset_iterate()
            r.DROP(2)
            assert np.array_equal(r.image("x").reshape(-1), x2), launch
        print(f"[segment] DROP residuals {[round(x, 1) for x in res]}; reruns and schedules bit-identical",
              flush=True)
