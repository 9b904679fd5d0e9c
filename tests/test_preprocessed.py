"""Preprocessed-data export/import (accumulate_and_export_preprocessed_data, V.cu:6613-6658;
init_import_preprocessed_data, V.cu:6775-6861).

CPU: the oracle's writer against hand-computed bytes of the reference layout, and its reader on
both header forms.  GPU: the library's files equal the oracle's byte for byte for the same run; an
import reproduces DROP (bit for bit with the FP64 chords) from either header form; malformed files are
refused before anything reaches the device.
"""
from __future__ import annotations

import shutil
import struct

import numpy as np
import pytest

from cases import small_head
from pct_reconstruction_amd import derive
from pct_reconstruction_amd.recon import default_params

FILES = ["MLP.bin", "voxels_per_path.bin", "avg_chord_lengths.bin", "avg_chord_lengths_f64.bin", "WEPL.bin",
         "histories.bin", "hull.txt", "x_0.txt"]


def _tiny_out():
    P = derive(default_params(recon_cyl_radius=10.0, columns=4, rows=3, slice_thickness=1.0, ssd_v_size=3.0))
    out = dict(offsets=np.array([0, 2, 2, 5], np.int64), voxels=np.array([7, 1, 23, 4, 0], np.uint32),
               chord=np.array([0.1, 0.25, 1.0 / 3.0]), wepl_recon=np.array([1.5, 2.5, 3.5], np.float32),
               first_voxel=np.array([7, 1, 23], np.uint32),
               h_recon=np.arange(30, dtype=np.float32).reshape(10, 3), bin_recon=np.array([5, 6, 7], np.int32),
               gantry_recon=np.array([0.0, 90.0, 359.9], np.float32), hull=np.array([1, 0] * 12, np.uint8),
               x0=np.linspace(0, 1, 24).astype(np.float32))
    return P, out


def test_oracle_writer_layout_is_the_reference_layout(oracle, tmp_path):
    P, out = _tiny_out()
    assert P.num_voxels == 24
    oracle.write_preprocessed(P, out, tmp_path)
    u32 = lambda *v: struct.pack(f"<{len(v)}I", *v)
    assert (tmp_path / "MLP.bin").read_bytes() == u32(3) + u32(7, 1, 23, 4, 0)
    assert (tmp_path / "voxels_per_path.bin").read_bytes() == u32(3) + u32(2, 0, 3)
    assert (tmp_path / "avg_chord_lengths.bin").read_bytes() == u32(3) + struct.pack("<3f", 0.1, 0.25, 1.0 / 3.0)
    assert (tmp_path / "avg_chord_lengths_f64.bin").read_bytes() == u32(3) + struct.pack("<3d", 0.1, 0.25, 1.0 / 3.0)
    assert (tmp_path / "WEPL.bin").read_bytes() == u32(3) + struct.pack("<3f", 1.5, 2.5, 3.5)
    hb = (tmp_path / "histories.bin").read_bytes()
    # voxel 7 = (x 3, y 1, z 0), 1 = (1, 0, 0), 23 = (3, 2, 1) on a 4 x 3 x 2 grid
    expect = u32(3) + struct.pack("<3i", 3, 1, 3) + struct.pack("<3i", 1, 0, 2) + struct.pack("<3i", 0, 0, 1)
    expect += np.arange(30, dtype="<f4").tobytes() + struct.pack("<3i", 5, 6, 7) + struct.pack("<3i", 0, 90, 359)
    assert hb == expect
    assert (tmp_path / "hull.txt").read_bytes() == bytes([1, 0] * 12)
    assert (tmp_path / "x_0.txt").read_bytes() == np.linspace(0, 1, 24).astype("<f4").tobytes()
    # the exporter's own text header (V.cu:6628-6629) on MLP.bin and WEPL.bin
    oracle.write_preprocessed(P, out, tmp_path / "text", text_header=True)
    assert (tmp_path / "text" / "MLP.bin").read_bytes() == b"3\n" + u32(7, 1, 23, 4, 0)
    assert (tmp_path / "text" / "WEPL.bin").read_bytes()[:2] == b"3\n"
    for d in (tmp_path, tmp_path / "text"):
        r = oracle.read_preprocessed(d)
        assert r["n"] == 3 and list(r["voxels"]) == [7, 1, 23, 4, 0] and list(r["counts"]) == [2, 0, 3]
        assert list(r["wepl"]) == [1.5, 2.5, 3.5] and list(r["gantry"]) == [0, 90, 359]


# ---------------------------------------------------------------- GPU
def _build(P, data):
    from pct_reconstruction_amd import Reconstruction
    r = Reconstruction(P)
    r.load_histories(data["t"], data["v"], data["u"], data["wepl"], data["angle_deg"])
    r.preprocessing()
    r.reconstruction_cuts()
    r.build_mlp_lists()
    return r


@pytest.mark.gpu
def test_gpu_export_equals_the_oracle_files(oracle, tmp_path):
    P, data = small_head(n=40000, seed=13)
    ref = oracle.run_pipeline(P, data, stop_after="lists")
    oracle.write_preprocessed(P, ref, tmp_path / "oracle")
    with _build(P, data) as r:
        r.export_preprocessed(tmp_path / "lib")
        assert list(r.histories("gantry_angle")) == list(ref["gantry_recon"])
    for name in FILES:
        a = (tmp_path / "lib" / name).read_bytes()
        b = (tmp_path / "oracle" / name).read_bytes()
        print(f"[export] {name}: {len(a)} bytes", flush=True)
        assert a == b, name


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["f64", "float-chords", "text-header"])
def test_gpu_import_resumes_drop(tmp_path, form):
    from pct_reconstruction_amd import Reconstruction
    P, data = small_head(n=40000, seed=13, iterations=3, tvs_on=1, tvs_seed=5)
    with _build(P, data) as r:
        r.export_preprocessed(tmp_path)
        r.DROP()
        x_mem = r.image("x").reshape(-1)
        log_mem = r.tvs_log()
    if form == "float-chords":
        (tmp_path / "avg_chord_lengths_f64.bin").unlink()
    if form == "text-header":
        for name in ("MLP.bin", "WEPL.bin"):
            b = (tmp_path / name).read_bytes()
            n = struct.unpack_from("<I", b)[0]
            (tmp_path / name).write_bytes(f"{n}\n".encode() + b[4:])
    with Reconstruction(P) as q:
        q.import_preprocessed(tmp_path)
        q.DROP()
        x_imp = q.image("x").reshape(-1)
        log_imp = q.tvs_log()
    rel = np.linalg.norm(x_imp - x_mem) / np.linalg.norm(x_mem)
    print(f"[import {form}] rel L2 {rel:.3e}, L {[e[2] for e in log_imp]} vs {[e[2] for e in log_mem]}", flush=True)
    if form == "float-chords":
        assert rel <= 1e-6
    else:
        assert np.array_equal(x_imp, x_mem)
        assert log_imp == log_mem


@pytest.mark.gpu
def test_gpu_import_refuses_malformed_files(tmp_path):
    from pct_reconstruction_amd import Reconstruction
    P, data = small_head(n=20000, seed=13)
    with _build(P, data) as r:
        r.export_preprocessed(tmp_path / "good")

    def variant(name, mutate):
        d = tmp_path / name
        shutil.copytree(tmp_path / "good", d)
        mutate(d)
        return d

    def bad_voxel(d):
        b = bytearray((d / "MLP.bin").read_bytes())
        b[4:8] = struct.pack("<I", P.num_voxels)  # first entry: one past the last voxel
        (d / "MLP.bin").write_bytes(bytes(b))

    def short_wepl(d):
        (d / "WEPL.bin").write_bytes((d / "WEPL.bin").read_bytes()[:-4])

    def long_path(d):
        b = bytearray((d / "voxels_per_path.bin").read_bytes())
        b[4:8] = struct.pack("<I", 10 ** 6)
        (d / "voxels_per_path.bin").write_bytes(bytes(b))

    cases = [variant("bad_voxel", bad_voxel), variant("short_wepl", short_wepl), variant("long_path", long_path),
             variant("no_x0", lambda d: (d / "x_0.txt").unlink())]
    for d in cases:
        with Reconstruction(P) as q:
            with pytest.raises(RuntimeError) as e:
                q.import_preprocessed(d)
            print(f"[import {d.name}] {e.value}", flush=True)
    with Reconstruction(P) as q:  # the untouched files still import
        q.import_preprocessed(tmp_path / "good")
        q.DROP(1)
