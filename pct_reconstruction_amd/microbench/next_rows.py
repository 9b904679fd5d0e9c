"""Timings of the stages on either side of the MLP + DROP path, on the head geometry (256x256x96 at
about 1 mm, MSC+SM hulls, block 60000) with a bounded number of histories (default 5M):

  - projection files: PCTN version-1 writing, reading (read_data_chunk_v0, DS.cu:4361) and upload;
  - the FBP median filter (radius 3, DS.h:480) inside generate_preprocessing_images (DS.cu:6150-6166);
  - preprocessed-data export and import (V.cu:6613-6658, V.cu:6775-6861);
  - one NTVS iteration (DS.cu:10583) between DROP iterations, against a DROP iteration alone;
  - the text, .dat and PNG image writers (DS.cu:10743, DS.cu:9939).

Every timing is host wall-clock around the library call (the calls synchronise their stream before
returning); file stages include the file system.  Prints one JSON object.

    python pct_reconstruction_amd/microbench/next_rows.py [--histories N] [--seed S]
"""
from __future__ import annotations

import argparse
import json
import shutil
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

from pct_reconstruction_amd import Reconstruction, configs, io, synth  # noqa: E402


def _timed(fn, *args):
    t0 = time.perf_counter()
    out = fn(*args)
    return out, time.perf_counter() - t0


def _dir_bytes(d: Path) -> int:
    return sum(f.stat().st_size for f in d.rglob("*") if f.is_file())


def _through_lists(r: Reconstruction) -> None:
    r.import_and_process_data()
    r.hull_detection_finish()
    r.statistical_calculations_and_cuts()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--histories", type=int, default=5_000_000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--files", type=int, default=10, help="projection files the histories are split over")
    args = ap.parse_args()

    P, ph, bm, n = configs.head_50m(args.histories)
    data = synth.generate(ph, bm, args.seed, n)
    work = Path(tempfile.mkdtemp(prefix="pct_next_rows_"))
    res: dict = {"workload": f"head geometry 256x256x96, {n} histories", "histories": n}
    try:
        # projection files (PCTN version 1: per-history angles)
        proj = work / "proj"
        proj.mkdir()
        parts = np.array_split(np.arange(n), args.files)
        t0 = time.perf_counter()
        for k, idx in enumerate(parts):
            io.write_projection(proj / f"projection_{k:03d}.bin", data["t"][:, idx], data["v"][:, idx],
                                data["u"][:, idx], data["wepl"][idx], float(k), angles=data["angle_deg"][idx])
        t_write = time.perf_counter() - t0
        nbytes = _dir_bytes(proj)
        t0 = time.perf_counter()
        chunks = [io.read_projection(p) for p in sorted(proj.iterdir())]
        t_read = time.perf_counter() - t0
        assert sum(c["wepl"].shape[0] for c in chunks) == n
        with Reconstruction(P) as r:
            t0 = time.perf_counter()
            for c in chunks:
                r.load_histories(c["t"], c["v"], c["u"], c["wepl"], c["angle_deg"])
            r.synchronize()
            t_load = time.perf_counter() - t0
        del chunks
        res["projection_files"] = {
            "files": args.files, "bytes": nbytes, "write_s": t_write, "read_s": t_read, "upload_s": t_load,
            "read_GBs": nbytes / t_read / 1e9, "read_histories_per_s": n / t_read,
            "upload_histories_per_s": n / t_load}
        shutil.rmtree(proj)

        # FBP median filter: preprocessing images with and without it
        Pm = P.copy()
        Pm.median_filter_fbp, Pm.fbp_med_filter_radius = 1, 3
        pre = {}
        for name, Q in (("plain", P), ("median_r3", Pm)):
            with Reconstruction(Q) as r:
                r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
                _through_lists(r)
                best = min(_timed(r.generate_preprocessing_images)[1] for _ in range(3))
                pre[name] = best
        res["fbp_median_filter"] = {"preprocessing_images_s": pre["plain"],
                                    "preprocessing_images_median_r3_s": pre["median_r3"],
                                    "median_filter_ms": (pre["median_r3"] - pre["plain"]) * 1e3}

        # lists, export, import
        exp = work / "preprocessed"
        with Reconstruction(P) as r:
            r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
            _through_lists(r)
            r.generate_preprocessing_images()
            r.reconstruction_cuts()
            r.build_mlp_lists()
            m, entries = r.count("reconstruction"), r.count("list_entries")
            _, t_exp = _timed(r.export_preprocessed, exp)
            x_img = r.image("x").reshape(P.slices, P.rows, P.columns)
        ebytes = _dir_bytes(exp)
        with Reconstruction(P) as q:
            _, t_imp = _timed(q.import_preprocessed, exp)
        res["preprocessed"] = {"reconstruction_histories": m, "list_entries": entries, "bytes": ebytes,
                               "export_s": t_exp, "import_s": t_imp, "export_GBs": ebytes / t_exp / 1e9,
                               "import_GBs": ebytes / t_imp / 1e9}

        # one DROP iteration with and without the NTVS iteration before it (DS.cu:9889-9915)
        Pt = P.copy()
        Pt.tvs_on = 1
        drop = {}
        for name, Q in (("drop", P), ("drop_tvs", Pt)):
            with Reconstruction(Q) as r:
                r.import_preprocessed(exp)
                times = []
                for _ in range(3):
                    times.append(_timed(r.DROP, 1)[1])
                drop[name] = min(times)
                if name == "drop_tvs":
                    res_log = r.tvs_log()
        res["tvs"] = {"drop_iteration_ms": drop["drop"] * 1e3, "drop_iteration_with_ntvs_ms": drop["drop_tvs"] * 1e3,
                      "ntvs_iteration_ms": (drop["drop_tvs"] - drop["drop"]) * 1e3,
                      "ntvs_L_after": [L for _, _, L in res_log]}
        shutil.rmtree(exp)

        # image writers
        wr = {}
        for name, fn, suffix in (("text", io.write_image_text, "txt"), ("dat", io.write_image_dat, "dat"),
                                 ("png", io.write_image_png, "png")):
            path = work / f"x.{suffix}"
            wr[name] = {"s": _timed(fn, path, x_img)[1], "bytes": path.stat().st_size}
        res["writers"] = {"voxels": int(x_img.size), **wr}
    finally:
        shutil.rmtree(work, ignore_errors=True)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
