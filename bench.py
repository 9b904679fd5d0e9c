"""MLP+DROP throughput on one MI355X: the head-50M workload.

Workload: 10-ellipsoid head phantom, 50M seeded synthetic histories with
continuous gantry angles, 256x256x96 grid at ~1 mm, MSC + SM hulls, 3-sigma cuts, FBP x0, MLP + DROP
for 20 iterations with block 60000.  Preprocessing (binning, cuts, hulls, FBP, endpoints) runs once
before the timed region; one timed step = building the MLP lists of all reconstruction histories
plus 20 DROP iterations from x0, all inputs resident in HBM.

metric: MLP+DROP proton histories/s = reconstruction histories x DROP iterations / step time.
The JSON line also carries the roofline of the dominant kernel (HIP-event timing of every launch of
that kernel during one profiled iteration) and the oracle's CPU rate on a bounded sample.

time_to_residual: on head-50M, the wall time from the FBP x0 to the residual of 20 DROP iterations against
the learned start (the shipped net on the FBP image and the hull, then DROP until the same residual), with
the residual checked after every iteration inside the timed region, so that the iteration counts are found
there; the protocol with the net on the iterate after one DROP iteration is reported beside it.

--workload large-200M-segment measures one GPU's share of the large-200M run instead: a 25M-history
segment on the 512x512x128 grid (not the default line).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(P, r, sample: int, threads: int, iterations: int, cpu_iterations: int) -> dict:
    """Oracle (C, OpenMP) MLP lists + `cpu_iterations` DROP iterations on the first `sample`
    reconstruction histories, extrapolated per history-iteration like the GPU metric."""
    sys.path.insert(0, str(ROOT / "oracle"))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    import pct_oracle as O
    names = ["x_entry", "y_entry", "z_entry", "x_exit", "y_exit", "z_exit", "xy_entry_angle", "xz_entry_angle",
             "xy_exit_angle", "xz_exit_angle"]
    h = np.ascontiguousarray(np.stack([r.histories(k)[:sample] for k in names]))
    first = np.ascontiguousarray(r.histories("first_voxel")[:sample])
    w = np.ascontiguousarray(r.histories("wepl")[:sample])
    x0 = r.image("x0").reshape(-1)
    tabs = O.MlpTables(P)
    t0 = time.perf_counter()
    off, vox, chord = O.mlp_lists(P, tabs, h, first)
    t_mlp = time.perf_counter() - t0
    x = x0.copy()
    t0 = time.perf_counter()
    for _ in range(cpu_iterations):
        x = O.drop_iteration(P, off, vox, chord, w, x, accum=0, nthreads=threads)
    t_drop = (time.perf_counter() - t0) / cpu_iterations
    n = h.shape[1]
    rate = n * iterations / (t_mlp + iterations * t_drop)
    return {"value": rate, "unit": "histories/s", "cores": threads, "kind": "port",
            "sample": f"first {n} reconstruction histories of the workload: oracle MLP lists ({t_mlp:.2f} s) + "
                      f"{cpu_iterations} DROP iterations ({t_drop:.2f} s each, FP32 reference-precision sums), "
                      f"rate = n*{iterations}/(t_mlp + {iterations}*t_drop)"}


def cpu_stage_rates(P, ph, bm, seed: int, n: int, threads: int) -> dict:
    """Every stage of the oracle's reference-precision pipeline (a)-(e) timed on `n` fresh histories of the
    workload on the host cores: wall time and histories/s of the histories each stage processes, one DROP
    iteration with the CPU baseline's float atomic sums."""
    sys.path.insert(0, str(ROOT / "oracle"))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    import pct_oracle as O
    from pct_reconstruction_amd import synth
    data = synth.generate(ph, bm, seed + 2, n)
    tm = {}
    out = O.run_pipeline(P, data, iterations=1, accum=0, timings=tm, nthreads=threads)
    volume, cut, recon = out["h_volume"].shape[1], out["h_cut"].shape[1], out["offsets"].size - 1
    work = {"intersections": n, "binning": volume, "hulls": volume, "stats_and_cuts": volume, "fbp": None,
            "hull_selection_x0": None, "endpoints": cut, "mlp_lists": recon, "drop": recon}
    stages = {}
    for k, w in work.items():
        sec = tm.get(k, 0.0)
        stages[k] = {"ms": 1e3 * sec, "histories": w, "histories_per_s": (w / sec) if (w and sec > 0) else None}
    return {"sample": f"{n} fresh histories of the workload (seed {seed + 2}): {volume} in the volume, {cut} after "
                      f"the cuts, {recon} reconstructed", "cores": threads, "stages": stages}


def stages_around(n: int, seed: int, files: int = 10) -> dict:
    """Host wall time of the stages on either side of MLP + DROP on the head geometry (256x256x96, MSC+SM,
    block 60000) with n histories: PCTN version-1 files written, read (read_data_chunk_v0, DS.cu:4361) and
    uploaded; the FBP median filter (radius 3, DS.cu:6150-6166) inside the preprocessing images; the
    preprocessed-data export and import (V.cu:6613-6658, :6775-6861); one DROP iteration with and without
    the NTVS iteration before it (DS.cu:9889-9915); the text, .dat and PNG image writers (DS.cu:10743,
    :9939).  Every call synchronises its stream before returning; file stages include the file system."""
    import shutil
    import tempfile
    from pct_reconstruction_amd import Reconstruction, configs, io, synth

    def timed(fn, *a):
        t0 = time.perf_counter()
        out = fn(*a)
        return out, time.perf_counter() - t0

    def dir_bytes(d):
        return sum(f.stat().st_size for f in Path(d).rglob("*") if f.is_file())

    P, ph, bm, n = configs.head_50m(n)
    data = synth.generate(ph, bm, seed, n)
    work = Path(tempfile.mkdtemp(prefix="pct_stages_"))
    res = {"workload": f"head geometry 256x256x96, {n} histories", "histories": n}
    try:
        proj = work / "proj"
        proj.mkdir()
        parts = np.array_split(np.arange(n), files)
        t0 = time.perf_counter()
        for k, idx in enumerate(parts):
            io.write_projection(proj / f"projection_{k:03d}.bin", data["t"][:, idx], data["v"][:, idx],
                                data["u"][:, idx], data["wepl"][idx], float(k), angles=data["angle_deg"][idx])
        t_write = time.perf_counter() - t0
        nbytes = dir_bytes(proj)
        chunks, t_read = timed(lambda: [io.read_projection(f) for f in sorted(proj.iterdir())])
        with Reconstruction(P) as r:
            t0 = time.perf_counter()
            for c in chunks:
                r.load_histories(c["t"], c["v"], c["u"], c["wepl"], c["angle_deg"])
            r.synchronize()
            t_load = time.perf_counter() - t0
        del chunks
        res["projection_files"] = {"files": files, "bytes": nbytes, "write_s": t_write, "read_s": t_read,
                                   "upload_s": t_load, "read_GBs": nbytes / t_read / 1e9,
                                   "read_histories_per_s": n / t_read, "upload_histories_per_s": n / t_load}
        shutil.rmtree(proj)
        Pm = P.copy()
        Pm.median_filter_fbp, Pm.fbp_med_filter_radius = 1, 3
        pre = {}
        for name, Q in (("plain", P), ("median_r3", Pm)):
            with Reconstruction(Q) as r:
                r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
                r.import_and_process_data()
                r.hull_detection_finish()
                r.statistical_calculations_and_cuts()
                pre[name] = min(timed(r.generate_preprocessing_images)[1] for _ in range(3))
        res["fbp_median_filter"] = {"preprocessing_images_s": pre["plain"],
                                    "preprocessing_images_median_r3_s": pre["median_r3"],
                                    "median_filter_ms": (pre["median_r3"] - pre["plain"]) * 1e3}
        exp = work / "preprocessed"
        with Reconstruction(P) as r:
            r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
            r.preprocessing()
            r.reconstruction_cuts()
            r.build_mlp_lists()
            m, entries = r.count("reconstruction"), r.count("list_entries")
            _, t_exp = timed(r.export_preprocessed, exp)
            x_img = r.image("x")
        ebytes = dir_bytes(exp)
        with Reconstruction(P) as q:
            _, t_imp = timed(q.import_preprocessed, exp)
        res["preprocessed"] = {"reconstruction_histories": m, "list_entries": entries, "bytes": ebytes,
                               "export_s": t_exp, "import_s": t_imp, "export_GBs": ebytes / t_exp / 1e9,
                               "import_GBs": ebytes / t_imp / 1e9}
        Pt = P.copy()
        Pt.tvs_on = 1
        drop, tvs_log = {}, []
        for name, Q in (("drop", P), ("drop_tvs", Pt)):
            with Reconstruction(Q) as r:
                r.import_preprocessed(exp)
                drop[name] = min(timed(r.DROP, 1)[1] for _ in range(3))
                if name == "drop_tvs":
                    tvs_log = r.tvs_log()
        res["tvs"] = {"drop_iteration_ms": drop["drop"] * 1e3, "drop_iteration_with_ntvs_ms": drop["drop_tvs"] * 1e3,
                      "ntvs_iteration_ms": (drop["drop_tvs"] - drop["drop"]) * 1e3,
                      "ntvs_L_after": [L for _, _, L in tvs_log]}
        shutil.rmtree(exp)
        wr = {}
        for name, fn, suffix in (("text", io.write_image_text, "txt"), ("dat", io.write_image_dat, "dat"),
                                 ("png", io.write_image_png, "png")):
            path = work / f"x.{suffix}"
            wr[name] = {"s": timed(fn, path, x_img)[1], "bytes": path.stat().st_size}
        res["writers"] = {"voxels": int(x_img.size), **wr}
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return res


def time_to_residual(r, P, weights, ref_iterations: int, fp64: bool) -> dict:
    """Time to the FBP-start residual from x0, for both starts, with the iteration counts found inside the timed
    region.  Target: ||Ax - b|| after `ref_iterations` DROP iterations from the FBP x0 (an untimed pass that
    defines it).  FBP start: DROP iterations, the residual after each, until it is at or below the target.
    Learned start: one DROP iteration, the net on that iterate (kept only if its residual is lower), then DROP
    iterations, the residual after each, until the target (at least one after the net, so that DROP defines
    the final image).  Both timings include every residual evaluation; replays with the iteration counts
    fixed (no residuals) are reported beside them.  With `fp64`, the learned start's final image is compared
    with the oracle's FP64 run of the same protocol (FP64 DROP, FP64 network) over the same lists."""
    from pct_reconstruction_amd.learned import WarmStartNet

    def clock():
        r.synchronize()
        return time.perf_counter()

    net = WarmStartNet(P.columns, P.rows, P.slices, seed=0)
    cap = 2 * ref_iterations
    try:
        net.load(str(weights))
        r.reset_iterate()
        r.DROP(ref_iterations)
        target = r.residual()
        hull = r.image("hull").reshape(-1) > 0
        x_fbp = r.image("x").reshape(-1)
        # FBP start, residual checked after every iteration
        r.reset_iterate()
        t0 = clock()
        k_fbp, res_fbp = 0, float("inf")
        while k_fbp < cap and res_fbp > target:
            r.DROP(1)
            k_fbp += 1
            res_fbp = r.residual()
        t_fbp = clock() - t0
        # learned start, residual checked after every iteration
        r.reset_iterate()
        t0 = clock()
        r.DROP(1)
        judge = net.apply_iterate(r)
        k_net = 1
        while True:
            r.DROP(1)
            k_net += 1
            res_net = r.residual()
            if res_net <= target or k_net >= cap:
                break
        t_net = clock() - t0
        x_net = r.image("x").reshape(-1)
        # replays with the counts fixed
        r.reset_iterate()
        t0 = clock()
        r.DROP(k_fbp)
        t_fbp_fixed = clock() - t0
        r.reset_iterate()
        t0 = clock()
        r.DROP(1)
        net.apply_iterate(r)
        r.DROP(k_net - 1)
        t_net_fixed = clock() - t0
        params = net.get_params()
    finally:
        net.close()
    d = x_net - x_fbp
    out = {"reference": f"FBP x0, {ref_iterations} DROP iterations", "residual_target": target,
           "residual_fbp": res_fbp, "residual_reached": res_net, "iterations_fbp": k_fbp,
           "iterations_learned": k_net, "net_used": judge["used_net"],
           "fbp_ms": 1e3 * t_fbp, "learned_ms": 1e3 * t_net, "speedup": t_fbp / t_net,
           "counts_found_in_timed_region": True,
           "fixed_counts": {"fbp_ms": 1e3 * t_fbp_fixed, "learned_ms": 1e3 * t_net_fixed,
                            "speedup": t_fbp_fixed / t_net_fixed},
           "learned_vs_fbp_start_image_rel_l2_hull": float(np.linalg.norm(d[hull]) / np.linalg.norm(x_fbp[hull])),
           "weights": weights.name}
    if fp64:
        out["fp64_oracle"] = learned_fp64_error(r, P, params, judge["used_net"], k_net, x_net)
    return out


def fbp_input_start(r, P, weights, warm: dict, fp64: bool) -> dict:
    """The network on the FBP image and the hull (pct_net_apply: the net's x0 replaces the FBP x0 when its
    residual is lower), then DROP iterations, the residual after each, until the FBP start's residual target
    (at least one).  Timed like the FBP start in `warm`, against which the speedup is taken.  The context's
    x0 becomes the net's, so this runs last.  With `fp64`, the final image is compared with the oracle's FP64
    run of the same protocol over the same lists."""
    from pct_reconstruction_amd.learned import WarmStartNet
    net = WarmStartNet(P.columns, P.rows, P.slices, seed=0)
    target, cap = warm["residual_target"], 2 * warm["iterations_fbp"]
    fbp = r.image("fbp")
    try:
        net.load(str(weights))
        r.reset_iterate()
        r.synchronize()
        t0 = time.perf_counter()
        judge = net.apply(r)
        k = 0
        while True:
            r.DROP(1)
            k += 1
            res = r.residual()
            if res <= target or k >= cap:
                break
        r.synchronize()
        t = time.perf_counter() - t0
        params = net.get_params()
    finally:
        net.close()
    x_net = r.image("x").reshape(-1)
    out = {"protocol": "net(FBP, hull) as x0 (pct_net_apply), then DROP", "net_used": judge["used_net"],
           "residual_fbp_x0": judge["residual_fbp"], "residual_net_x0": judge["residual_net"],
           "residual_reached": res, "iterations_learned": k, "learned_ms": 1e3 * t,
           "speedup": warm["fbp_ms"] / (1e3 * t), "weights": weights.name}
    if fp64:
        out["fp64_oracle"] = learned_fp64_error(r, P, params, judge["used_net"], k, x_net, fbp=fbp)
    return out


def learned_fp64_error(r, P, params, used_net: bool, k: int, x_gpu, fbp=None) -> dict:
    """The learned start's image against the oracle's FP64 run of the same protocol over the library's lists.
    Without `fbp` (network on the iterate after one DROP iteration): FP64 DROP iteration from x0, the FP64
    network on that iterate (rounded to float32, the image type the network reads), k - 1 FP64 DROP
    iterations.  With `fbp` (network on the FBP image): the FP64 network on the FBP image, k FP64 DROP
    iterations.  The GPU runs DROP with FP64 sums and the network in BF16 MFMA with FP32 accumulation."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pct_oracle as O
    t0 = time.perf_counter()
    off, vox = r.lists()
    chord, w = r.histories("chord"), r.histories("wepl")
    hull = r.image("hull").astype(np.uint8)
    shape = (P.slices, P.rows, P.columns)

    def net64(img):
        _, _, y = O.net_loss_grad(shape, params, np.asarray(img, np.float32).reshape(shape), hull,
                                  np.zeros(shape, np.float32), bf16=False, with_grad=False)
        return y.reshape(-1)

    if fbp is None:
        x64 = O.drop_iteration_fp64(P, off, vox, chord, w, r.image("x0").reshape(-1).astype(np.float64))
        if used_net:
            x64 = net64(x64)
        more = k - 1
        protocol = (f"FP64 DROP x1, FP64 network {'applied' if used_net else 'judged and not applied'}, "
                    f"{more} more FP64 DROP iterations")
    else:
        x64 = net64(fbp) if used_net else r.image("x0").reshape(-1).astype(np.float64)
        more = k
        protocol = f"FP64 network on the FBP image{'' if used_net else ' (not applied)'}, {k} FP64 DROP iterations"
    for _ in range(more):
        x64 = O.drop_iteration_fp64(P, off, vox, chord, w, x64)
    del off, vox
    x = np.asarray(x_gpu, np.float64).reshape(-1)
    rel = float(np.linalg.norm(x - x64) / np.linalg.norm(x64))
    return {"rel_l2": rel, "target": 1e-3, "meets_target": rel <= 1e-3, "max_abs": float(np.abs(x - x64).max()),
            "oracle_s": time.perf_counter() - t0, "protocol": protocol}


def reference_scheme_drop(r, P, iterations: int) -> dict:
    """The reference's DROP scheme on the same lists, through pct_drop_iteration_device (DROP_full_tx_iteration,
    DS.cu:9537-9589): per block, one wavefront per history scatters x_update and S with two atomics per list
    entry (calculate_x_update_GPU :9321-9419), then the touched voxels are applied (image_update_GPU
    :9522-9536).  Timed with HIP events on the default stream; the image after the first iteration is compared
    with the transposed DROP's from the same x0."""
    import ctypes as C
    import torch
    from pct_reconstruction_amd._lib import lib
    off, vox = r.lists()
    chord, wepl = r.histories("chord"), r.histories("wepl")
    x0 = r.image("x0").reshape(-1)
    n, nv, entries = off.shape[0] - 1, x0.size, int(off[-1])
    r.reset_iterate()
    r.DROP(1)
    x_transposed = r.image("x").reshape(-1)
    dev = torch.device("cuda:0")
    d_off, d_vox = torch.from_numpy(off).to(dev), torch.from_numpy(vox.view(np.int32)).to(dev)
    del vox
    d_chord, d_w = torch.from_numpy(chord).to(dev), torch.from_numpy(wepl).to(dev)
    d_x = torch.from_numpy(x0.copy()).to(dev)
    d_acc = torch.zeros(nv, dtype=torch.float64, device=dev)
    d_cnt = torch.zeros(nv, dtype=torch.int32, device=dev)
    f = lib().pct_drop_iteration_device
    args = (n, C.c_void_p(d_off.data_ptr()), C.c_void_p(d_vox.data_ptr()), C.c_void_p(d_chord.data_ptr()),
            C.c_void_p(d_w.data_ptr()), C.c_void_p(d_x.data_ptr()), C.c_void_p(d_acc.data_ptr()),
            C.c_void_p(d_cnt.data_ptr()), nv, P.drop_block_size, P.histories_per_block, P.quirk_a3_drop_floor,
            P.min_mlp_length, P.lambda_, None)
    torch.cuda.synchronize()
    times, rel = [], None
    for k in range(iterations):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        rc = f(*args)
        e1.record()
        torch.cuda.synchronize()
        if rc != 0:
            raise RuntimeError(f"pct_drop_iteration_device returned {rc}")
        times.append(e0.elapsed_time(e1))
        if k == 0:
            x1 = d_x.cpu().numpy()
            rel = float(np.linalg.norm(x1 - x_transposed) / np.linalg.norm(x_transposed))
    ms = float(np.median(times))
    return {"scheme": "two atomics per list entry (x_update as FP64 sums, S as u32), touched-voxel apply",
            "iterations": iterations, "ms_per_iteration": ms, "ms_each": times, "histories": n,
            "list_entries": entries, "entries_per_s": entries / (ms * 1e-3),
            "scatter_atomics_per_s": 2 * entries / (ms * 1e-3),
            "rel_l2_vs_transposed_iteration_1": rel}


def fp64_oracle_error(r, P, iterations: int) -> dict:
    """The image after `iterations` DROP iterations from x0 against the oracle's FP64 DROP (x, sums and scalars in
    double) over the library's lists, which the parity tests hold bit-identical to the oracle's own."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pct_oracle as O
    r.reset_iterate()
    r.DROP(iterations)
    x = r.image("x").reshape(-1).astype(np.float64)
    off, vox = r.lists()
    chord = r.histories("chord")
    w = r.histories("wepl")
    x64 = r.image("x0").reshape(-1).astype(np.float64)
    t0 = time.perf_counter()
    for _ in range(iterations):
        x64 = O.drop_iteration_fp64(P, off, vox, chord, w, x64)
    t = time.perf_counter() - t0
    del off, vox
    return {"iterations": iterations, "rel_l2": float(np.linalg.norm(x - x64) / np.linalg.norm(x64)),
            "max_abs": float(np.abs(x - x64).max()), "oracle_s": t,
            "sample": f"all {chord.size} reconstruction histories, DROP from x0 in FP64"}


def fp64_oracle_error_sampled(P, ph, bm, seed: int, n: int, iterations: int, blocks_per_segment: int, hull,
                              x0) -> dict:
    """The large-200M path on a sample of its workload: `n` fresh histories with the same geometry and beam,
    reconstructed on the full run's hull and x0 (HULL_TYPE = X_0_TYPE = IMPORT), lists built per segment of
    `blocks_per_segment` DROP blocks, `iterations` DROP iterations, against the oracle's FP64 DROP over the
    oracle's own lists from the same x0."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pct_oracle as O
    from pct_reconstruction_amd import Reconstruction, synth
    Q = P.copy()
    Q.list_segment_histories = blocks_per_segment * Q.drop_block_size
    Q.hull_type = 0
    Q.x_0_type = 0
    data = synth.generate(ph, bm, seed + 1, n)
    with Reconstruction(Q, device=0) as r:
        r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
        r.set_hull(hull)
        r.preprocessing(x0)
        r.reconstruction_cuts()
        r.build_mlp_lists()
        r.DROP(iterations)
        x = r.image("x").reshape(-1).astype(np.float64)
        segments = r.metrics()["counts"].get("segments", {}).get("count", 0)
        n_recon = r.count("reconstruction")
    ref = O.run_pipeline(Q, data, stop_after="lists", hull_import=hull, x0_import=x0)
    assert ref["offsets"].size - 1 == n_recon
    x64 = np.asarray(x0, np.float64).reshape(-1)
    for _ in range(iterations):
        x64 = O.drop_iteration_fp64(Q, ref["offsets"], ref["voxels"], ref["chord"], ref["wepl_recon"], x64)
    return {"iterations": iterations, "rel_l2": float(np.linalg.norm(x - x64) / np.linalg.norm(x64)),
            "max_abs": float(np.abs(x - x64).max()),
            "sample": f"{n} fresh histories of the same geometry and beam on the run's hull and x0 ({n_recon} "
                      f"reconstructed, {segments} list segments of {blocks_per_segment} blocks), DROP from x0, "
                      f"oracle FP64 DROP on its own lists"}


def fp64_pipeline_error(P, ph, bm, seed: int, n: int, iterations: int) -> dict:
    """Every image of the library's pipeline on `n` fresh histories of the workload against the FP64 oracle of
    the whole pipeline (pct_oracle_fp64.c: every stage in double), and against the reference-precision
    oracle.  Decisions (cuts, hull voxels, reconstructed histories) may differ between float and double
    arithmetic; the FP64 DROP on the reference-precision lists from the same x0 separates the DROP's own
    arithmetic error from them."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pct_oracle as O
    from pct_reconstruction_amd import Reconstruction, synth
    data = synth.generate(ph, bm, seed + 1, n)
    with Reconstruction(P, device=0) as r:
        r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
        r.preprocessing()
        r.image_reconstruction(iterations)
        img = {k: r.image(k).reshape(-1).astype(np.float64) for k in ("fbp", "x0", "x")}
        hull = r.image("hull").reshape(-1)
        counts = {k: r.count(k) for k in ("post_cuts", "reconstruction")}
    t0 = time.perf_counter()
    o64 = O.run_pipeline(P, data, iterations=iterations, fp64=True)
    o32 = O.run_pipeline(P, data, iterations=iterations, accum=P.drop_accum)
    t_oracle = time.perf_counter() - t0
    x64 = o32["x0"].astype(np.float64)
    for _ in range(iterations):
        x64 = O.drop_iteration_fp64(P, o32["offsets"], o32["voxels"], o32["chord"], o32["wepl_recon"], x64)

    def err(a, b):
        b = np.asarray(b, np.float64)
        return {"rel_l2": float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)),
                "max_abs": float(np.abs(a - b).max())}
    return {
        "histories": n, "iterations": iterations, "oracle_s": t_oracle,
        "vs_fp64_pipeline": {k: err(img[k], o64[k]) for k in img},
        "vs_reference_precision_oracle": {k: err(img[k], o32[k]) for k in img},
        "drop_vs_fp64_same_lists": err(img["x"], x64),
        "decisions": {"post_cuts": [counts["post_cuts"], int(o64["h_cut"].shape[1])],
                      "reconstruction": [counts["reconstruction"], int(o64["offsets"].size - 1)],
                      "hull_voxels_differing": int(np.sum(hull != o64["hull"]))},
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="head-50M", choices=["head-50M", "cylinder-2M", "large-200M", "large-200M-segment"],
                    help="head-50M (the benchmark); cylinder-2M: the CPU-runnable case; large-200M: 200M histories on 512x512x128 @ 0.5x0.5x1 mm, "
                         "230 MeV, SM hull, block 120000, 12 iterations, MLP lists and DROP plans built per "
                         "segment of 58 blocks inside every iteration, learned start; large-200M-segment: one "
                         "25M-history share of it with resident lists")
    ap.add_argument("--histories", type=int, default=None, help="default: 50M (head-50M), 200M, 25M (segment)")
    ap.add_argument("--segment-histories", type=int, default=6_960_000,
                    help="large-200M: reconstruction histories per list segment (whole DROP blocks)")
    ap.add_argument("--iterations", type=int, default=None, help="DROP iterations per step (default: config)")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="reconstruction histories timed on the CPU (0: all of them)")
    ap.add_argument("--cpu-iterations", type=int, default=2, help="DROP iterations timed on the CPU")
    ap.add_argument("--cpu-stages", type=int, default=2_000_000,
                    help="fresh histories on which every oracle stage is timed on the CPU (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the time-to-residual measurement")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-oracle comparison")
    ap.add_argument("--stages-around", type=int, default=0, metavar="N",
                    help="also time the stages around MLP + DROP (file I/O, median filter, preprocessed export and "
                         "import, TVS, image writers) on N head-geometry histories")
    ap.add_argument("--fp64-pipeline", type=int, default=0, metavar="N",
                    help="also compare every image with the FP64 oracle of the whole pipeline on N fresh histories "
                         "of the workload (minutes of CPU time for N = 2M)")
    ap.add_argument("--reference-scheme", type=int, default=0, metavar="K",
                    help="also time K iterations of the reference's atomic DROP scheme on the same lists")
    ap.add_argument("--seed", type=int, default=2024)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "runs on one GPU only: --gpus must be 1 (sharding over GPUs is a design, "
                                   "not implemented)"}))
        return 2

    from pct_reconstruction_amd import Reconstruction, configs, synth

    if args.workload == "head-50M":
        P, ph, bm, n = configs.head_50m(args.histories or 50_000_000)
        workload = "head-50M: 256x256x96 @ ~1 mm, MSC+SM hulls, 3-sigma cuts, FBP x0"
    elif args.workload == "cylinder-2M":
        P, ph, bm, n = configs.cylinder_2m(args.histories or 2_000_000)
        workload = ("cylinder-2M: water cylinder + 3 inserts, 200x200x32 @ 1x1x2.5 mm, 90 x 4 deg, SC hull, "
                    "Shepp-Logan FBP x0")
    elif args.workload == "large-200M":
        P, ph, bm, n = configs.large_200m(args.histories or 200_000_000)
        P.list_segment_histories = args.segment_histories
        workload = (f"large-200M: 512x512x128 @ 0.5x0.5x1 mm, 230 MeV, SM hull, 3-sigma cuts, learned x0 (net on the "
                    f"iterate after one DROP iteration, kept if the residual drops), MLP lists and DROP plans built "
                    f"per segment of {args.segment_histories} histories in every iteration")
    else:
        P, ph, bm, n = configs.large_200m(args.histories or 25_000_000)
        workload = ("large-200M segment: 25M of the 200M histories, 512x512x128 @ 0.5x0.5x1 mm, 230 MeV, SM hull, "
                    "3-sigma cuts, FBP x0")
    iters = P.iterations if args.iterations is None else args.iterations
    log(f"generating {n} histories (seed {args.seed})")
    t0 = time.perf_counter()
    data = synth.generate(ph, bm, args.seed, n)
    log(f"generated in {time.perf_counter() - t0:.1f} s")

    r = Reconstruction(P, device=0)
    r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
    del data
    log("loaded; preprocessing")
    t0 = time.perf_counter()
    r.preprocessing()
    r.reconstruction_cuts()
    r.synchronize()
    t_pre = time.perf_counter() - t0
    m = r.metrics()
    log(f"preprocessing {t_pre:.2f} s; counts {m['counts']}")
    n_recon = r.count("reconstruction")

    learned = None
    if args.workload == "large-200M":
        from pct_reconstruction_amd.learned import WarmStartNet
        learned = WarmStartNet(P.columns, P.rows, P.slices, seed=0)
        learned.load(str(ROOT / "pct_reconstruction_amd" / "weights" / "warmstart_head50m_x1.bin"))
    judges = []

    def step():
        r.reset_iterate()
        r.build_mlp_lists()
        if learned is None:
            r.DROP(iters)
            return
        # learned start: one DROP iteration, the net on that iterate (kept only if ||Ax - b|| drops), the rest
        r.DROP(1)
        judges.append(learned.apply_iterate(r))
        r.DROP(iters - 1)

    for k in range(args.warmup):
        step()
        r.synchronize()
        log(f"warmup step {k + 1}/{args.warmup} done")
    r.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        if learned is not None:
            r.synchronize()
            log(f"timed step {k + 1}/{args.steps} done")
    r.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"{args.steps} timed steps in {elapsed:.3f} s")
    ms_per_step = 1000.0 * elapsed / args.steps
    value = n_recon * iters * args.steps / elapsed
    m = r.metrics()
    entries = r.count("list_entries")
    segs = m["counts"].get("segments")

    # per-kernel timing: one profiled DROP iteration (every launch bracketed by HIP events)
    r.reset_iterate()
    prof = r.profile_drop(1)
    if segs:
        log(f"profiled iteration: {segs['count']} segments, list builds {segs['lists_ms']:.0f} ms, "
            f"DROP blocks {segs['drop_ms']:.0f} ms")
    mlp_ms = m["kernel_ms"]["mlp_lists"]
    nb = max(prof["blocks"], 1)
    # algorithmic bytes per launch of this implementation: rows = list index + x gather per entry, 32 B of
    # per-history header/result; cols = history index + r gather per summed entry, voxel id,
    # offset and x read-modify-write per touched voxel; MLP = 64 B in/out per history + 4 B per entry.
    rows_bytes = (8 * prof["gathered"] + 32 * prof["histories"]) / nb
    cols_bytes = (12 * prof["summed"] + 16 * prof["touched"]) / nb
    mlp_bytes = 4 * entries + 64 * n_recon

    def kern(ms, nbytes, launches_per_step):
        return {"avg_launch_ms": ms, "bytes_per_launch": nbytes, "GBs": nbytes / (ms * 1e-3) / 1e9,
                "launches_per_step": launches_per_step, "ms_per_step": ms * launches_per_step}

    kernels = {
        "drop_rows": kern(prof["rows_ms"], rows_bytes, iters * nb),
        "drop_cols": kern(prof["cols_ms"], cols_bytes, iters * nb),
    }
    if segs:  # one MLP launch per segment and iteration; the last segment's launch time
        ns = max(segs["count"], 1)
        kernels["mlp_lists"] = kern(mlp_ms, mlp_bytes / ns, iters * ns)
        kernels["list_build_ms_per_iteration"] = segs["lists_ms"]
        kernels["drop_blocks_ms_per_iteration"] = segs["drop_ms"]
    else:
        kernels["mlp_lists"] = kern(mlp_ms, mlp_bytes, 1)
        kernels["list_transpose_ms_per_step"] = max(m["stage_ms"]["mlp_build_lists"] - mlp_ms, 0.0)
    # The dominant work is the DROP block update, which the reference performs as
    # calculate_x_update_GPU + image_update_GPU and this build as the pair k_drop_rows + k_drop_cols.
    # Its algorithmic bytes per proton-iteration are B_alg = 18 n + 12 (list entry 6 B incl. chord,
    # x gather 4 B, x_update read-modify-write 8 B per voxel; 8 B header + 4 B WEPL per proton),
    # priced per block and divided by the block's average rows + cols duration.
    block_ms = prof["rows_ms"] + prof["cols_ms"]
    block_bytes = (18 * prof["gathered"] + 12 * prof["histories"]) / nb
    kernels["drop_block"] = kern(block_ms, block_bytes, iters * nb)
    # HBM traffic per block is not counted in this run: it comes from the committed rocprofv3 --pmc passes
    # (FETCH_SIZE, WRITE_SIZE) of the same workload and kernels, and is labelled as stored.
    traffic, traffic_src = None, None
    tfile = ROOT / "profiles" / "drop_block_traffic.json"
    if tfile.exists() and args.workload == "head-50M":
        try:
            t = json.loads(tfile.read_text())
            if t.get("workload_histories") == n:
                traffic = t.get("hbm_bytes_per_block")
                traffic_src = f"stored: {tfile.relative_to(ROOT)} (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)"
        except (OSError, ValueError):
            traffic = None
    roofline = {"bound": "hbm", "kernel": "drop_block (k_drop_rows + k_drop_cols per DROP block)",
                "achieved": block_bytes / (block_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": block_bytes / (block_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": traffic,
                "traffic_source": traffic_src, "bytes_per_launch": block_bytes, "avg_launch_ms": block_ms}
    # whole-step algorithmic HBM rate: list writes (4 B per entry + 64 B per history) and the DROP blocks
    step_bytes = iters * (block_bytes * nb) + (iters if segs else 1) * mlp_bytes
    hbm_step = {"GBs": step_bytes / (ms_per_step * 1e-3) / 1e9, "bytes_per_step": step_bytes}

    # time to a fixed residual, learned x0 vs FBP x0 (head-50M only: the shipped net is trained for that grid)
    warm = None
    weights = ROOT / "pct_reconstruction_amd" / "weights" / "warmstart_head50m_x1.bin"
    if not args.no_warm_start and args.workload == "head-50M" and weights.exists():
        warm = time_to_residual(r, P, weights, P.iterations, not args.no_fp64)
        log(f"time to residual: FBP start {warm['fbp_ms']:.1f} ms ({warm['iterations_fbp']} iterations), learned "
            f"start {warm['learned_ms']:.1f} ms ({warm['iterations_learned']} iterations), speedup "
            f"{warm['speedup']:.2f}")
        if "fp64_oracle" in warm:
            log(f"learned start vs FP64 oracle: rel L2 {warm['fp64_oracle']['rel_l2']:.3e}")

    fp64 = None
    if args.workload == "head-50M" and not args.no_fp64:
        log("FP64 oracle DROP over the library's lists")
        fp64 = fp64_oracle_error(r, r.params, 2)
        log(f"FP64 oracle: rel L2 {fp64['rel_l2']:.3e} after {fp64['iterations']} iterations")

    ref_scheme = None
    if args.reference_scheme > 0 and not segs:
        log(f"reference DROP scheme, {args.reference_scheme} iterations")
        ref_scheme = reference_scheme_drop(r, r.params, args.reference_scheme)
        log(f"reference scheme: {ref_scheme['ms_per_iteration']:.1f} ms per iteration, "
            f"{ref_scheme['scatter_atomics_per_s'] / 1e9:.1f} G atomics/s")

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu_default = 2_000_000 if args.workload == "large-200M" else 0
        want = args.cpu_sample if args.cpu_sample > 0 else cpu_default
        sample = n_recon if want <= 0 else min(want, n_recon)
        log(f"cpu baseline on {sample} histories, {threads} threads")
        cpu = cpu_baseline(r.params, r, sample, threads, iters, max(args.cpu_iterations, 1))
        if args.cpu_stages > 0:
            log(f"cpu stage rates on {args.cpu_stages} histories")
            cpu["per_stage"] = cpu_stage_rates(r.params, ph, bm, args.seed, args.cpu_stages, threads)

    stages = {k: round(v, 3) for k, v in m["stage_ms"].items()}
    cnt = m["counts"]
    per_stage_work = {"intersect_bin_carve": cnt["loaded"], "stats_and_cuts": cnt["recon_volume"],
                      "mlp_endpoints": cnt["post_cuts"], "mlp_build_lists": n_recon,
                      # the last DROP call of the timed step: all iterations, or those after the net's start
                      "drop": n_recon * (iters - 1 if learned is not None else iters)}
    stage_rates = {k: (w / (stages[k] * 1e-3) if stages.get(k) else None) for k, w in per_stage_work.items()}
    run_hull, run_x0 = r.image("hull").astype(np.uint8), r.image("x0")
    fbp_weights = ROOT / "pct_reconstruction_amd" / "weights" / "warmstart_head50m_fbp.bin"
    if warm is not None and fbp_weights.exists():
        # the network on FBP + hull is the primary learned start; the iterate protocol is reported beside it
        fi = fbp_input_start(r, P, fbp_weights, warm, not args.no_fp64)
        x1 = {k: warm.pop(k) for k in ("residual_reached", "iterations_learned", "net_used", "learned_ms",
                                       "speedup", "fixed_counts", "learned_vs_fbp_start_image_rel_l2_hull",
                                       "weights", "fp64_oracle") if k in warm}
        x1["protocol"] = "one DROP iteration, net(iterate, hull) (pct_net_apply_iterate), then DROP"
        warm.update(fi)
        warm["iterate_protocol"] = x1
        log(f"FBP-input start: {fi['learned_ms']:.1f} ms ({fi['iterations_learned']} iterations), "
            f"speedup {fi['speedup']:.2f}" + (f", FP64 oracle rel L2 {fi['fp64_oracle']['rel_l2']:.3e}"
                                               if "fp64_oracle" in fi else ""))
    if learned is not None:
        learned.close()
    r.close()
    around = None
    if args.stages_around > 0:
        log(f"stages around the path on {args.stages_around} histories")
        around = stages_around(args.stages_around, args.seed)
    fp64_pipe = None
    if args.fp64_pipeline > 0:
        log(f"FP64 oracle of the whole pipeline on {args.fp64_pipeline} histories")
        fp64_pipe = fp64_pipeline_error(P, ph, bm, args.seed, args.fp64_pipeline, iters)
        log(f"FP64 pipeline: x rel L2 {fp64_pipe['vs_fp64_pipeline']['x']['rel_l2']:.3e}, same lists "
            f"{fp64_pipe['drop_vs_fp64_same_lists']['rel_l2']:.3e}")
    if args.workload == "large-200M" and not args.no_fp64:
        log("FP64 oracle on a 1M-history sample of the workload")
        fp64 = fp64_oracle_error_sampled(P, ph, bm, args.seed, 1_000_000, iters, 2, run_hull, run_x0)
        log(f"FP64 oracle (sample): rel L2 {fp64['rel_l2']:.3e}")
    out = {
        "metric": "MLP+DROP proton histories/s; time to fixed residual: learned x0 vs FBP x0",
        "value": value,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (MLP, accumulators) / f32 (image)",
        "data": ("synthetic (seeded water cylinder with 3 inserts, Highland scatter, WEPL noise 3 mm, 1% outliers)"
                 if args.workload == "cylinder-2M" else
                 "synthetic (seeded 10-ellipsoid head phantom, Highland scatter, WEPL noise 2.5 mm, 1% outliers)"),
        "config": {"workload": f"{workload}, MLP + DROP {iters} iterations, block {P.drop_block_size}",
                   "histories_loaded": n, "histories_reconstructed": n_recon, "list_entries": entries,
                   "drop_iterations": iters, "drop_block_size": P.drop_block_size},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp64_oracle": fp64,
        "fp64_pipeline": fp64_pipe,
        "stages_around": around,
        "hbm_step": hbm_step,
        "time_to_residual": warm,
        "reference_scheme_drop": ref_scheme,
        "learned_start": judges[-1] if judges else None,
        "kernels": kernels,
        "stages_ms": stages,
        "stage_histories_per_s": stage_rates,
        "preprocessing_s": t_pre,
        "counts": m["counts"],
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
