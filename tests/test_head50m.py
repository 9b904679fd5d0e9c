"""Parity at the benchmark's full size (head-50M: 50M histories, 256x256x96, MSC+SM hulls, 20 DROP
iterations, block 60000), where the whole oracle pipeline would take too long on the CPU:

  - intersections and bin numbers of all 50M histories, oracle vs library, bit for bit;
  - MLP voxel lists and chords of 5000 sampled reconstruction histories, recomputed by the oracle from
    the library's own endpoints, bit for bit;
  - list invariants: counts sum to the entries, every index is a voxel, every list starts at its
    entry voxel;
  - DROP: two runs from x0 give the same bits, the visiting order, image layout and transposition
    (sort or buckets) change no bit, and the data residual falls with every iteration;
  - two DROP iterations within 1e-4 relative L2 of the oracle's FP64 DROP on the same lists.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H_NAMES = ["x_entry", "y_entry", "z_entry", "x_exit", "y_exit", "z_exit", "xy_entry_angle", "xz_entry_angle",
           "xy_exit_angle", "xz_exit_angle"]


@pytest.mark.timeout(900)
def test_head_50m_full_size_parity_and_invariants(oracle):
    from pct_reconstruction_amd import Reconstruction, configs, synth

    P, ph, bm, n = configs.head_50m()
    data = synth.generate(ph, bm, 2024, n)
    print(f"[head-50M] {n} histories generated", flush=True)
    with Reconstruction(P) as r:
        r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
        r.import_and_process_data()
        t, v, u, _ = oracle.convert_units(P, data["t"], data["v"], data["u"], data["wepl"])
        h, missed = oracle.intersections(P, t, v, u, data["angle_deg"])
        keep = ~missed
        bins, _ = oracle.binning(P, h)
        assert r.count("recon_volume") == int(keep.sum())
        assert np.array_equal(r.histories("bin_num"), bins[keep])
        sample = np.sort(np.random.default_rng(1).choice(int(keep.sum()), 100000, replace=False))
        for k, name in enumerate(H_NAMES):
            assert np.array_equal(r.histories(name)[sample], h[k][keep][sample]), name
        del h, bins, t, v, u
        print(f"[head-50M] intersections and bins of {n} histories identical", flush=True)

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
        idx = np.sort(np.random.default_rng(2).choice(m, 5000, replace=False))
        tabs = oracle.MlpTables(P)
        o_off, o_vox, o_chord = oracle.mlp_lists(P, tabs, np.ascontiguousarray(he[:, idx]), first[idx])
        for j, i in enumerate(idx):
            assert np.array_equal(vox[off[i]:off[i + 1]], o_vox[o_off[j]:o_off[j + 1]]), i
        assert np.array_equal(chord[idx], o_chord)
        print(f"[head-50M] {m} reconstruction histories, {off[-1]} entries; lists of 5000 sampled histories "
              f"identical to the oracle", flush=True)

        res = [r.residual()]
        for _ in range(3):
            r.DROP(1)
            res.append(r.residual())
        x3 = r.image("x").reshape(-1)
        assert all(b < a for a, b in zip(res, res[1:])), res
        r.reset_iterate()
        r.DROP(3)
        assert np.array_equal(r.image("x").reshape(-1), x3)
        for launch in (dict(spatial_order=False, bricks=False), dict(cols_sliced=False), dict(buckets=True)):
            r.set_launch(**launch)
            r.reset_iterate()
            r.DROP(3)
            assert np.array_equal(r.image("x").reshape(-1), x3), launch
            if launch.get("buckets"):
                assert r.metrics()["drop_plan"]["groups_sorted"] == 0
        print(f"[head-50M] DROP residuals {[round(x, 1) for x in res]}; reruns and schedules bit-identical",
              flush=True)
        # the image after two iterations against the oracle's FP64 DROP over the same lists (x, sums and
        # scalars in double): the FP64-oracle error of the benchmark's DROP
        r.set_launch()
        r.reset_iterate()
        r.DROP(2)
        x2 = r.image("x").reshape(-1).astype(np.float64)
        x64 = r.image("x0").reshape(-1).astype(np.float64)
        w = r.histories("wepl")
        for _ in range(2):
            x64 = oracle.drop_iteration_fp64(P, off, vox, chord, w, x64)
        rel = float(np.linalg.norm(x2 - x64) / np.linalg.norm(x64))
        print(f"[head-50M] two DROP iterations vs the FP64 oracle: rel L2 {rel:.3e}", flush=True)
        assert rel <= 1e-4
