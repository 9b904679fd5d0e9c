"""MLP lists and DROP plans built per segment of histories (LIST_SEGMENT_HISTORIES), the device-memory bound of
the large runs (the reference's partial-transfer DROP, DS.cu:9590-9742).

Segments are whole DROP blocks, so every block sees the same histories in the same order as in the resident run:
the DROP image is bit-identical, with and without TVS, and the residual agrees to rounding of its sum order.
"""
from __future__ import annotations

import numpy as np
import pytest

from cases import small_cylinder, small_head


def test_segment_parameter_parses_and_validates(tmp_path):
    from pct_reconstruction_amd.recon import default_params, derive, load_settings
    cfg = tmp_path / "s.cfg"
    cfg.write_text("LIST_SEGMENT_HISTORIES = 9600\nMLP_ALGORITHM = 0\n")
    p = load_settings(cfg)
    assert p.list_segment_histories == 9600 and p.mlp_algorithm == 0
    assert default_params().list_segment_histories == 0 and default_params().mlp_algorithm == 1
    with pytest.raises(ValueError):
        derive(default_params(list_segment_histories=-1))
    with pytest.raises(ValueError):
        derive(default_params(mlp_algorithm=2))


def _run(P, data, iterations, tvs=False):
    from pct_reconstruction_amd import Reconstruction
    with Reconstruction(P) as r:
        r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
        r.preprocessing()
        r.reconstruction_cuts()
        r.build_mlp_lists()
        r.DROP(iterations)
        out = dict(x=r.image("x").reshape(-1), residual=r.residual(), entries=r.count("list_entries"),
                   n=r.count("reconstruction"), metrics=r.metrics(), first=r.histories("first_voxel"),
                   tvs=r.tvs_log() if tvs else None)
        if P.list_segment_histories:
            with pytest.raises(RuntimeError):
                r.lists()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("tvs", [False, True])
def test_segmented_drop_is_bit_identical_to_the_resident_run(tvs):
    P, data = small_cylinder(n=60000, seed=41, drop_block_size=3200, histories_per_block=320, iterations=3,
                             tvs_on=int(tvs))
    res = _run(P, data, 3, tvs)
    Ps = P.copy()
    Ps.list_segment_histories = 3 * 3200 + 1000  # rounds down to 3 blocks
    seg = _run(Ps, data, 3, tvs)
    n = res["n"]
    n_seg = (n + 9599) // 9600
    print(f"[segments tvs={tvs}] {n} histories, {n_seg} segments, {res['entries']} entries; residual "
          f"{res['residual']:.9e} vs {seg['residual']:.9e}", flush=True)
    assert n_seg >= 3 and n % 3200 != 0  # several segments and a partial final block
    assert seg["metrics"]["counts"]["segments"]["count"] == n_seg
    assert seg["metrics"]["counts"]["segments"]["histories"] == 9600
    assert seg["entries"] == res["entries"] > 0
    assert np.array_equal(seg["first"], res["first"])  # the per-history arrays are whole again
    assert np.array_equal(seg["x"], res["x"])
    assert abs(seg["residual"] - res["residual"]) <= 1e-12 * res["residual"]
    if tvs:
        assert seg["tvs"] == res["tvs"]


@pytest.mark.gpu
def test_segmented_drop_matches_the_oracle_with_the_untabulated_mlp(oracle):
    """Segments of 2 blocks on the head case with MLP_ALGORITHM = STANDARD, against the oracle's DROP."""
    P, data = small_head(n=50000, seed=43, mlp_algorithm=0, drop_block_size=4000, histories_per_block=400,
                         iterations=2, list_segment_histories=8000)
    ref = oracle.run_pipeline(P, data, iterations=2, accum=P.drop_accum)
    got = _run(P, data, 2)
    rel = float(np.linalg.norm(got["x"] - ref["x"]) / np.linalg.norm(ref["x"]))
    print(f"[segments oracle] rel L2 {rel:.2e}", flush=True)
    assert got["n"] == ref["offsets"].shape[0] - 1
    assert got["entries"] == ref["voxels"].size
    assert rel <= 1e-6
