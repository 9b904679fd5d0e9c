"""Host-side behaviour of the library that needs no GPU: ABI, parameters, I/O, synthetic data."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from pct_reconstruction_amd import _lib, derive, io, load_settings, set_key, synth
from pct_reconstruction_amd.recon import default_params

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = _lib.exported_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []


def test_params_struct_layout_matches_c():
    assert _lib.lib().pct_params_size() == C.sizeof(_lib.Params)


def test_settings_cfg_parsing_and_derivation():
    p = derive(load_settings(GOLDEN / "settings_head.cfg"))
    assert (p.columns, p.rows, p.slices) == (256, 256, 96)
    assert p.recon_cyl_height == pytest.approx(9.61)
    assert p.voxel_thickness == pytest.approx(9.61 / 96)
    assert (p.t_bins, p.v_bins, p.angular_bins) == (350, 42, 90)
    assert p.hull_type == 2 and p.x_0_type == 2 and p.sm_on == 1 and p.sc_on == 0
    assert p.lambda_ == pytest.approx(1e-4) and p.drop_block_size == 60000


def test_settings_errors_are_reported(tmp_path):
    bad = tmp_path / "bad.cfg"
    bad.write_text("COLUMNS = 200\nNOT_A_KEY = 3\n")
    with pytest.raises(ValueError, match="NOT_A_KEY"):
        load_settings(bad)
    p = default_params()
    with pytest.raises(KeyError):
        set_key(p, "NOPE", 1)
    with pytest.raises(ValueError):
        derive(default_params(columns=0))
    with pytest.raises(ValueError):
        derive(default_params(recon_cyl_radius=25.0))   # tables cover only 40 cm of depth


def test_projection_file_round_trip(tmp_path):
    ph = synth.phantom_cylinder()
    d = synth.generate(ph, synth.beam(), 3, 1000)
    f = tmp_path / "projection_000.bin"
    io.write_projection(f, d["t"] * 10, d["v"] * 10, d["u"] * 10, d["wepl"] * 10, angle_deg=36.0)
    r = io.read_projection(f)
    assert r["header_angle_deg"] == 36.0
    assert np.array_equal(r["t"], d["t"] * 10) and np.array_equal(r["wepl"], d["wepl"] * 10)
    assert (r["angle_deg"] == 36.0).all()
    g = tmp_path / "projection_cont.bin"
    io.write_projection(g, d["t"], d["v"], d["u"], d["wepl"], angle_deg=0.0, angles=d["angle_deg"])
    r2 = io.read_projection(g)
    assert np.array_equal(r2["angle_deg"], d["angle_deg"])
    (tmp_path / "junk.bin").write_bytes(b"XXXX" + bytes(40))
    with pytest.raises(RuntimeError):
        io.read_projection(tmp_path / "junk.bin")
    with pytest.raises(RuntimeError):
        io.read_projection(tmp_path / "missing.bin")


def test_image_writers(tmp_path):
    img = (np.arange(2 * 3 * 4, dtype=np.float32) / 7).reshape(2, 3, 4)
    io.write_image_text(tmp_path / "x.txt", img)
    lines = (tmp_path / "x.txt").read_text().splitlines()
    assert len(lines) == 6 and len(lines[0].split()) == 4
    assert float(lines[5].split()[3]) == pytest.approx(img[1, 2, 3], rel=1e-5)
    io.write_image_dat(tmp_path / "x.dat", img)
    assert np.array_equal(np.fromfile(tmp_path / "x.dat", np.float32), (img / 2).reshape(-1))


def test_text_writer_bytes_match_the_default_stream_format(tmp_path):
    """array_2_disk writes `value << " "` per voxel and std::endl per row with the default stream
    format, which is C's %g (precision 6).  20 slices: more than the writer formats at once."""
    rng = np.random.default_rng(12)
    img = (rng.standard_normal((20, 7, 9)) * np.exp(rng.uniform(-25, 25, (20, 7, 9)))).astype(np.float32)
    img[0, 0, :6] = [0.0, -0.0, np.inf, -np.inf, 1e-45, 3.4e38]
    io.write_image_text(tmp_path / "x.txt", img)
    want = "".join("".join(f"{float(v):g} " for v in row) + "\n" for row in img.reshape(-1, img.shape[2]))
    assert want.startswith("0 -0 inf -inf 1.4013e-45 3.4e+38 ")
    assert (tmp_path / "x.txt").read_bytes() == want.encode()


def test_png_writer_decodes_to_the_quantized_image(tmp_path):
    from PIL import Image
    rng = np.random.default_rng(4)
    img = (rng.random((3, 70, 300)) * 2.4 - 0.2).astype(np.float32)  # values below 0 and above 2 clamp
    img[0, 0, :4] = [0.0, 2.0, 1.0, np.nan]
    io.write_image_png(tmp_path / "x.png", img)
    with Image.open(tmp_path / "x.png") as im:
        assert im.size == (300, 3 * 70)  # COLUMNS x ROWS*SLICES
        got = np.asarray(im).astype(np.int64)
    v = np.nan_to_num(img.reshape(3 * 70, 300) / 2, nan=0.0)
    expect = np.floor(np.clip(v, 0.0, 1.0).astype(np.float32) * np.float32(65535) + np.float32(0.5)).astype(np.int64)
    assert np.array_equal(got, expect)
    assert list(got[0, :4]) == [0, 65535, 32768, 0]
    # more than one 65535-byte stored deflate block: 210 rows of 601 bytes
    assert (tmp_path / "x.png").stat().st_size > 65535


def test_png_writer_splits_large_images_into_idat_chunks(tmp_path):
    """A PNG chunk holds at most 2^31 - 1 bytes; the writer splits the zlib stream into IDAT chunks of
    at most 1 MiB, which decoders concatenate."""
    import struct
    from PIL import Image
    rng = np.random.default_rng(5)
    img = (rng.random((4, 200, 700)) * 2.0).astype(np.float32)   # 800 scanlines of 1401 bytes: 1.12 MB
    io.write_image_png(tmp_path / "big.png", img)
    b = (tmp_path / "big.png").read_bytes()
    pos, sizes = 8, []
    while pos < len(b):
        n, kind = struct.unpack(">I4s", b[pos:pos + 8])
        if kind == b"IDAT":
            sizes.append(n)
        pos += 12 + n
    assert len(sizes) == 2 and sizes[0] == 1 << 20
    with Image.open(tmp_path / "big.png") as im:
        got = np.asarray(im).astype(np.int64)
    v = img.reshape(800, 700) / 2
    expect = np.floor(np.clip(v, 0.0, 1.0).astype(np.float32) * np.float32(65535) + np.float32(0.5)).astype(np.int64)
    assert np.array_equal(got, expect)


def test_launch_sizing_beyond_one_dispatch(tmp_path):
    """Kernels with one work-item per element refuse (zero grid) more work-items than one dispatch holds;
    grid-stride kernels are capped.  The residual kernels take 16 lanes per history, so above 2^28
    histories they need the grid-stride form."""
    import ctypes
    import subprocess
    src = tmp_path / "sizing.cpp"
    src.write_text('#include "launch.h"\n'
                   'extern "C" unsigned grid_for(long long n, int b) { return pct_grid_for(n, b); }\n'
                   'extern "C" unsigned grid_capped(long long n, int b, unsigned c) { return pct_grid_capped(n, b, c); }\n')
    lib = tmp_path / "libsizing.so"
    csrc = Path(__file__).resolve().parent.parent / "pct_reconstruction_amd" / "csrc"
    subprocess.run(["g++", "-std=c++17", "-shared", "-fPIC", "-I", str(csrc), "-o", str(lib), str(src)], check=True)
    L = ctypes.CDLL(str(lib))
    L.grid_for.argtypes = [ctypes.c_longlong, ctypes.c_int]
    L.grid_for.restype = ctypes.c_uint
    L.grid_capped.argtypes = [ctypes.c_longlong, ctypes.c_int, ctypes.c_uint]
    L.grid_capped.restype = ctypes.c_uint
    assert L.grid_for(0, 256) == 1 and L.grid_for(257, 256) == 2
    assert L.grid_for((1 << 28) * 16 - 256, 256) == (1 << 24) - 1     # just fits
    for n in (1 << 28, 300_000_000, 1 << 31):                        # histories x 16 lanes: too many
        assert L.grid_for(n * 16, 256) == 0
        g = L.grid_capped(n * 16, 256, 1 << 20)
        assert g == 1 << 20 and g * 256 < 1 << 32
    assert L.grid_capped(1000, 256, 1 << 20) == 4


def test_generator_is_deterministic_and_physical():
    ph = synth.phantom_cylinder()
    a = synth.generate(ph, synth.beam(), 9, 5000)
    b = synth.generate(ph, synth.beam(), 9, 5000)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # noiseless, no scatter: WEPL of a central proton equals the water chord 16 cm
    bm = synth.beam(wepl_noise_sigma=0.0, outlier_fraction=0.0, scatter=False, t_half_range=1e-6,
                    v_half_range=1e-6)
    c = synth.generate(ph, bm, 1, 4)
    expect = []
    for ang in np.deg2rad(c["angle_deg"].astype(np.float64)):
        w = 16.0
        for phi, rsp in zip(np.deg2rad([0.0, 120.0, 240.0]), (0.3, 1.1, 1.7)):
            d = 4.0 * abs(np.sin(phi - ang))
            if d < 1.0:
                w += (rsp - 1.0) * 2 * np.sqrt(1 - d * d)
        expect.append(w)
    assert np.allclose(c["wepl"], expect, atol=1e-4)
    # integer angles come in angle-major order
    assert (np.diff(a["angle_deg"]) >= 0).all()


def test_rasterize_cylinder_phantom():
    p = derive(default_params(columns=100, rows=100, slice_thickness=1.0))
    img = synth.rasterize(synth.phantom_cylinder(), p)
    assert img.shape == (p.slices, 100, 100)
    vals = np.unique(img)
    assert all(np.isclose(v, [0.0, 0.3, 1.0, 1.1, 1.7], atol=1e-6).any() for v in vals)
    assert img[4, 50, 50] == pytest.approx(1.0)


def _pct_header(magic, version, n, angle, energy, strings):
    """A projection-file header laid out as count_histories_v0 reads it (DS.cu:3590-3643): magic, version,
    count, angle, energy, two dates, then three length-prefixed strings."""
    import struct
    b = magic + struct.pack("<iiffii", version, n, angle, energy, 1234567, 7654321)
    for s in strings:
        b += struct.pack("<i", len(s)) + s
    return b


def test_projection_reader_follows_the_reference_header(tmp_path):
    """Files built byte by byte from the reference's header description, not by the library's writer: odd
    string lengths (the data starts at 4 + 7*4 + 2*4 + the three lengths, DS.cu:3643), the 13 arrays in
    t0..t3, v0..v3, u0..u3, WEPL order (DS.cu:4404-4416), the per-history angles of PCTN / version 1
    (:4446) and int(header angle) for PCTD version 0 (:4432)."""
    rng = np.random.default_rng(5)
    n = 37
    arrays = rng.normal(size=(13, n)).astype(np.float32)
    angles = rng.uniform(0, 360, n).astype(np.float32)
    strings = [b"phantom X", b"src", b"prepared by someone"]
    f0 = tmp_path / "v0.bin"
    f0.write_bytes(_pct_header(b"PCTD", 0, n, 36.7, 200.0, strings) + arrays.tobytes())
    r = io.read_projection(f0)
    assert r["header_angle_deg"] == np.float32(36.7)
    for k, name in enumerate(["t", "v", "u"]):
        assert np.array_equal(r[name], arrays[4 * k:4 * k + 4])
    assert np.array_equal(r["wepl"], arrays[12])
    assert (r["angle_deg"] == 36.0).all()
    f1 = tmp_path / "v1.bin"
    f1.write_bytes(_pct_header(b"PCTN", 1, n, 0.0, 230.0, strings) + arrays.tobytes() + angles.tobytes())
    r1 = io.read_projection(f1)
    assert np.array_equal(r1["t"], arrays[0:4]) and np.array_equal(r1["angle_deg"], angles)
    f2 = tmp_path / "short.bin"
    f2.write_bytes(_pct_header(b"PCTD", 0, n, 4.0, 200.0, strings) + arrays.tobytes()[:-4])
    with pytest.raises(RuntimeError):
        io.read_projection(f2)
    f3 = tmp_path / "v2.bin"
    f3.write_bytes(_pct_header(b"PCTD", 2, n, 4.0, 200.0, strings) + arrays.tobytes())
    with pytest.raises(RuntimeError):
        io.read_projection(f3)
