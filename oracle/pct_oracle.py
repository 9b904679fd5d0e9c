"""Python binding of the CPU oracle (oracle/_build/libpct_oracle.so) and its pipeline driver.

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Imported by tests/, by
__graft_entry__.smoke() and by the cpu_baseline leg of bench.py; the product never imports it.

`run_pipeline` strings the oracle's stage functions together in the reference's order
(preprocessing DS.cu:7394 -> image_reconstruction DS.cu:10650), doing the host compactions of
the reference (DS.cu:5429-5449, 5892-5911, 8286-8294) with numpy.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ORACLE_LIB = Path(__file__).resolve().parent / "_build" / "libpct_oracle.so"
_L = None


def lib() -> C.CDLL:
    global _L
    if _L is not None:
        return _L
    if not ORACLE_LIB.exists():
        raise RuntimeError(f"{ORACLE_LIB} missing: run pct_reconstruction_amd/build.py")
    L = C.CDLL(str(ORACLE_LIB))
    vp, i64, P = C.c_void_p, C.c_int64, C.POINTER
    f3 = C.c_void_p * 3
    f4 = C.c_void_p * 4
    f10 = C.c_void_p * 10
    from pct_reconstruction_amd._lib import Params
    sig = {
        "o_derive": ([P(Params)], None),
        "o_intersections": ([P(Params), i64, f4, f4, f4, vp, f10, vp], None),
        "o_binning": ([P(Params), i64, f10, vp, vp], None),
        "o_stats_and_cuts": ([P(Params), i64, vp, vp, vp, f10, vp, f3, f3, vp, vp, vp], None),
        "o_construct_sinogram": ([P(Params), vp, vp], None),
        "o_init_hull_u8": ([P(Params), vp], None),
        "o_init_hull_i32": ([P(Params), vp], None),
        "o_sc": ([P(Params), i64, f10, vp, vp], None),
        "o_msc": ([P(Params), i64, f10, vp, vp], None),
        "o_sm": ([P(Params), i64, f10, vp, vp], None),
        "o_msc_edge": ([P(Params), vp, vp], None),
        "o_sm_edge": ([P(Params), vp, vp], None),
        "o_filter": ([P(Params), vp, vp], None),
        "o_backproject": ([P(Params), vp, vp], None),
        "o_fbp_hull": ([P(Params), vp, vp], None),
        "o_median_filter_2d": ([P(Params), vp, C.c_int, vp], None),
        "o_avg_filter_hull": ([P(Params), vp, C.c_int, C.c_double, vp], None),
        "o_tables_create": ([P(Params), vp], C.c_int),
        "o_tables_destroy": ([vp], None),
        "o_endpoints": ([P(Params), vp, i64, f10, vp], None),
        "o_chord_length": ([P(Params), C.c_double, C.c_double], C.c_double),
        "o_mlp_lists": ([P(Params), vp, i64, f10, vp, vp, vp, i64, vp], i64),
        "o_mlp_steps": ([P(Params), vp, vp, C.c_uint32, vp, i64, vp, i64, vp, vp], i64),
        "o_drop_iteration": ([P(Params), i64, vp, vp, vp, vp, vp, C.c_int, C.c_int], None),
        "o_drop_iteration_fp64": ([P(Params), i64, vp, vp, vp, vp, vp], None),
        "o_residual": ([i64, vp, vp, vp, vp, vp, C.c_int], C.c_double),
        "o_tile_sum": ([vp, i64], C.c_double),
        "o_tvs_eta_len": ([P(Params)], C.c_int),
        "o_tvs_eta": ([P(Params), C.c_int], C.c_float),
        "o_tvs_draw": ([C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "o_tv": ([P(Params), vp, C.c_int], C.c_float),
        "o_tvs_perturbation": ([P(Params), vp, vp, vp, C.c_int], C.c_float),
        "o_tvs_iteration": ([P(Params), vp, vp, C.c_int, P(C.c_int), C.c_int, P(C.c_float), P(C.c_float)], None),
        "on_num_params": ([], i64),
        "on_loss_grad": ([C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp], C.c_double),
    }
    twins = ["o_intersections", "o_binning", "o_stats_and_cuts", "o_construct_sinogram", "o_sc", "o_msc", "o_sm",
             "o_filter", "o_backproject", "o_fbp_hull", "o_median_filter_2d", "o_endpoints", "o_chord_length",
             "o_mlp_lists", "o_mlp_steps", "o_drop_iteration_fp64", "o_residual", "o_tables_create",
             "o_tables_destroy"]
    for name in twins:  # the FP64 twin (pct_oracle_fp64.c): same arguments, double arrays
        sig[name + "_f64"] = sig[name]
    for name, (a, r) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = a
        fn.restype = r
    _L = L
    return L


class Tables(C.Structure):
    _fields_ = [("n_trig", C.c_int64), ("n_coeff", C.c_int64), ("n_poly", C.c_int64)] + [
        (k, C.POINTER(C.c_double)) for k in ("sin_t", "cos_t", "scat", "p12", "p23", "p34", "p26", "p312")]


def _p(a: np.ndarray) -> C.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def _arr(arrs, n):
    return (C.c_void_p * n)(*[a.ctypes.data for a in arrs])


def _fn(name, fp64):
    """Entry point `name` of the reference-precision oracle, or of its FP64 twin."""
    return getattr(lib(), name + ("_f64" if fp64 else ""))


def _real(fp64):
    return np.float64 if fp64 else np.float32


H_NAMES = ["x_entry", "y_entry", "z_entry", "x_exit", "y_exit", "z_exit", "xy_entry_angle", "xz_entry_angle",
           "xy_exit_angle", "xz_exit_angle"]


def derive(P):
    q = P.copy()
    lib().o_derive(C.byref(q))
    return q


def intersections(P, t, v, u, angle, fp64=False):
    n = angle.shape[0]
    R = _real(fp64)
    h = np.zeros((10, n), R)
    missed = np.zeros(n, np.uint8)
    tt = [np.ascontiguousarray(x, R) for x in t]
    vv = [np.ascontiguousarray(x, R) for x in v]
    uu = [np.ascontiguousarray(x, R) for x in u]
    ang = np.ascontiguousarray(angle, np.float32)
    _fn("o_intersections", fp64)(C.byref(P), n, _arr(tt, 4), _arr(vv, 4), _arr(uu, 4), _p(ang), _arr(list(h), 10),
                                 _p(missed))
    return h, missed.astype(bool)


def binning(P, h, fp64=False):
    n = h.shape[1]
    b = np.zeros(n, np.int32)
    r = np.zeros(n, np.uint8)
    _fn("o_binning", fp64)(C.byref(P), n, _arr(list(h), 10), _p(b), _p(r))
    return b, r


def stats_and_cuts(P, bin_num, in_range, wepl, h, fp64=False):
    n = wepl.shape[0]
    NB = P.num_bins
    R = _real(fp64)
    counts_pre = np.zeros(NB, np.int32)
    mean = np.zeros((3, NB), R)
    std = np.zeros((3, NB), R)
    counts_post = np.zeros(NB, np.int32)
    sino = np.zeros(NB, R)
    failed = np.zeros(n, np.uint8)
    _fn("o_stats_and_cuts", fp64)(C.byref(P), n, _p(bin_num), _p(in_range), _p(wepl), _arr(list(h), 10),
                                  _p(counts_pre), _arr(list(mean), 3), _arr(list(std), 3), _p(counts_post), _p(sino),
                                  _p(failed))
    return dict(counts_pre=counts_pre, mean=mean, std=std, counts_post=counts_post, sinogram_sum=sino,
                failed=failed.astype(bool))


def init_hull(P, dtype):
    out = np.zeros(P.num_voxels, dtype)
    if dtype == np.uint8:
        lib().o_init_hull_u8(C.byref(P), _p(out))
    else:
        lib().o_init_hull_i32(C.byref(P), _p(out))
    return out


def carve(P, method, h, wepl, fp64=False):
    n = wepl.shape[0]
    out = init_hull(P, np.uint8 if method == "sc" else np.int32)
    _fn("o_" + method, fp64)(C.byref(P), n, _arr(list(h), 10), _p(wepl), _p(out))
    return out


def msc_edge(P, counts):
    out = np.zeros(P.num_voxels, np.int32)
    lib().o_msc_edge(C.byref(P), _p(np.ascontiguousarray(counts, np.int32)), _p(out))
    return out


def sm_edge(P, counts):
    out = np.zeros(P.num_voxels, np.int32)
    lib().o_sm_edge(C.byref(P), _p(np.ascontiguousarray(counts, np.int32)), _p(out))
    return out


def construct_sinogram(P, counts_post, sino_sum, fp64=False):
    s = sino_sum.copy()
    _fn("o_construct_sinogram", fp64)(C.byref(P), _p(counts_post), _p(s))
    return s


def filter_sinogram(P, sino, fp64=False):
    R = _real(fp64)
    out = np.zeros(P.num_bins, R)
    _fn("o_filter", fp64)(C.byref(P), _p(np.ascontiguousarray(sino, R)), _p(out))
    return out


def backproject(P, filtered, fp64=False):
    R = _real(fp64)
    out = np.zeros(P.num_voxels, R)
    _fn("o_backproject", fp64)(C.byref(P), _p(np.ascontiguousarray(filtered, R)), _p(out))
    return out


def fbp_hull(P, image, fp64=False):
    out = np.zeros(P.num_voxels, np.uint8)
    _fn("o_fbp_hull", fp64)(C.byref(P), _p(np.ascontiguousarray(image, _real(fp64))), _p(out))
    return out


def median_filter_2d(P, image, radius, fp64=False):
    R = _real(fp64)
    out = np.zeros(P.num_voxels, R)
    _fn("o_median_filter_2d", fp64)(C.byref(P), _p(np.ascontiguousarray(image, R).reshape(-1)), int(radius), _p(out))
    return out


def avg_filter_hull(P, hull, radius, threshold):
    out = np.zeros(P.num_voxels, np.uint8)
    lib().o_avg_filter_hull(C.byref(P), _p(np.ascontiguousarray(hull, np.uint8).reshape(-1)), int(radius),
                            float(threshold), _p(out))
    return out


def endpoints(P, hull_u8, h, fp64=False):
    h = np.ascontiguousarray(h.copy(), _real(fp64))
    n = h.shape[1]
    first = np.zeros(n, np.uint32)
    _fn("o_endpoints", fp64)(C.byref(P), _p(np.ascontiguousarray(hull_u8, np.uint8)), n, _arr(list(h), 10),
                             _p(first))
    return h, first


def chord_length(P, angle_t, angle_v):
    return lib().o_chord_length(C.byref(P), float(angle_t), float(angle_v))


class MlpTables:
    def __init__(self, P):
        self.t = Tables()
        if lib().o_tables_create(C.byref(P), C.byref(self.t)) != 0:
            raise MemoryError("oracle tables")

    def __del__(self):
        try:
            lib().o_tables_destroy(C.byref(self.t))
        except Exception:
            pass

    def array(self, name):
        n = {"sin_t": self.t.n_trig, "cos_t": self.t.n_trig, "scat": self.t.n_coeff}.get(name, self.t.n_poly)
        return np.ctypeslib.as_array(getattr(self.t, name), shape=(n,)).copy()


def mlp_lists(P, tables: MlpTables, h, first, fp64=False):
    h = np.ascontiguousarray(h, _real(fp64))
    n = h.shape[1]
    off = np.zeros(n + 1, np.int64)
    chord = np.zeros(n, np.float64)
    cap = max(16, int(n) * (int(2 * P.recon_cyl_radius / P.mlp_u_step) + 4))
    vox = np.zeros(cap, np.uint32)
    total = _fn("o_mlp_lists", fp64)(C.byref(P), C.byref(tables.t), n, _arr(list(h), 10),
                                     _p(np.ascontiguousarray(first, np.uint32)), _p(off), _p(vox), cap, _p(chord))
    if total < 0:
        raise RuntimeError("oracle MLP capacity")
    return off, vox[:total].copy(), chord


def mlp_steps(P, tables: MlpTables, hist10, first, fp64=False):
    """Every u-step of the MLP loop for one history: returns (steps, path, chord) where steps is a
    (k, 6) array of u, t_1, v_1, x_1, y_1, voxel."""
    h = np.ascontiguousarray(hist10, _real(fp64)).reshape(10)
    cap = 8 * (int(2 * P.recon_cyl_radius / P.mlp_u_step) + 8)
    rec = np.zeros((cap, 6), np.float64)
    path = np.zeros(cap, np.uint32)
    plen = C.c_int64()
    chord = C.c_double()
    k = _fn("o_mlp_steps", fp64)(C.byref(P), C.byref(tables.t), _p(h), int(first), _p(rec), cap, _p(path), cap,
                                 C.byref(plen), C.byref(chord))
    if k > cap or plen.value < 0:
        raise RuntimeError("oracle MLP step capacity")
    return rec[:k].copy(), path[:plen.value].copy(), chord.value


def drop_iteration(P, off, vox, chord, wepl, x, accum=1, nthreads=1):
    x = np.ascontiguousarray(x, np.float32).copy()
    lib().o_drop_iteration(C.byref(P), off.shape[0] - 1, _p(off), _p(vox), _p(chord),
                           _p(np.ascontiguousarray(wepl, np.float32)), _p(x), int(accum), int(nthreads))
    return x


def drop_iteration_fp64(P, off, vox, chord, wepl, x64):
    """DROP with x, the sums and the scalars in double; WEPL of the float pipeline (float32) or of the FP64
    pipeline (float64)."""
    x = np.ascontiguousarray(x64, np.float64).copy()
    fp64 = np.asarray(wepl).dtype == np.float64
    _fn("o_drop_iteration_fp64", fp64)(C.byref(P), off.shape[0] - 1, _p(off), _p(vox), _p(chord),
                                       _p(np.ascontiguousarray(wepl, _real(fp64))), _p(x))
    return x


def residual(off, vox, chord, wepl, x, min_len):
    return lib().o_residual(off.shape[0] - 1, _p(off), _p(vox), _p(chord), _p(np.ascontiguousarray(wepl, np.float32)),
                            _p(np.ascontiguousarray(x, np.float32).reshape(-1)), int(min_len))


TVS_SUM_SERIAL_FLOAT, TVS_SUM_TILED_FP64 = 0, 1


def write_preprocessed(P, out, directory, text_header=False, f64=True):
    """The preprocessed-data files of accumulate_and_export_preprocessed_data (V.cu:6613-6658) and the
    exports it calls (V.cu:6540-6612), from run_pipeline's results.  Every file starts with the
    history count as uint32; text_header=True writes MLP.bin and WEPL.bin with the exporter's "%u\n"
    header instead (V.cu:6628-6629)."""
    import struct
    from pathlib import Path
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    n = int(out["chord"].shape[0])
    cnt = np.diff(out["offsets"]).astype("<u4")
    u32 = struct.pack("<I", n)
    hdr = f"{n}\n".encode() if text_header else u32
    (d / "MLP.bin").write_bytes(hdr + np.asarray(out["voxels"], "<u4").tobytes())
    (d / "voxels_per_path.bin").write_bytes(u32 + cnt.tobytes())
    (d / "avg_chord_lengths.bin").write_bytes(u32 + np.asarray(out["chord"], "<f8").astype("<f4").tobytes())
    if f64:
        (d / "avg_chord_lengths_f64.bin").write_bytes(u32 + np.asarray(out["chord"], "<f8").tobytes())
    (d / "WEPL.bin").write_bytes(hdr + np.asarray(out["wepl_recon"], "<f4").tobytes())
    fv = np.asarray(out["first_voxel"], np.int64)
    C, R = P.columns, P.rows
    parts = [u32] + [np.asarray(a, "<i4").tobytes() for a in (fv % C, (fv // C) % R, fv // (C * R))]
    parts += [np.asarray(out["h_recon"][k], "<f4").tobytes() for k in range(10)]
    parts += [np.asarray(out["bin_recon"], "<i4").tobytes(),
              np.asarray(out["gantry_recon"], np.float32).astype("<i4").tobytes()]
    (d / "histories.bin").write_bytes(b"".join(parts))
    (d / "hull.txt").write_bytes(np.asarray(out["hull"], np.uint8).tobytes())
    (d / "x_0.txt").write_bytes(np.asarray(out["x0"], "<f4").tobytes())


def read_preprocessed(directory):
    """Reads the files of write_preprocessed (either header form).  Returns a dict of arrays."""
    import struct
    from pathlib import Path
    d = Path(directory)
    raw = (d / "voxels_per_path.bin").read_bytes()
    n = struct.unpack_from("<I", raw)[0]
    cnt = np.frombuffer(raw, "<u4", count=n, offset=4)

    def payload(name, nbytes):
        b = (d / name).read_bytes()
        if len(b) == 4 + nbytes and struct.unpack_from("<I", b)[0] == n:
            return b[4:]
        k = b.index(b"\n")
        assert int(b[:k]) == n and len(b) == k + 1 + nbytes, name
        return b[k + 1:]

    total = int(cnt.sum(dtype=np.int64))
    res = dict(n=n, counts=cnt, voxels=np.frombuffer(payload("MLP.bin", 4 * total), "<u4"),
               wepl=np.frombuffer(payload("WEPL.bin", 4 * n), "<f4"),
               chord32=np.frombuffer((d / "avg_chord_lengths.bin").read_bytes(), "<f4", count=n, offset=4))
    hb = (d / "histories.bin").read_bytes()
    assert len(hb) == 4 + 15 * 4 * n
    arr = [np.frombuffer(hb, "<i4" if k in (0, 1, 2, 13, 14) else "<f4", count=n, offset=4 + 4 * n * k)
           for k in range(15)]
    res.update(voxel_xyz=arr[:3], h=arr[3:13], bin=arr[13], gantry=arr[14])
    return res


def tile_sum(terms):
    """FP64 sum in the library's fixed order (tiles of 1024, halving trees)."""
    t = np.ascontiguousarray(terms, np.float64)
    return lib().o_tile_sum(_p(t), t.size)


def tvs_eta(P, L):
    """TVS_eta_sequence_h[L - L_0] (DS.cu:10149-10156), 0 past the table."""
    return lib().o_tvs_eta(C.byref(P), int(L))


def tvs_eta_len(P):
    return lib().o_tvs_eta_len(C.byref(P))


def tvs_draw(seed, k, lo, hi):
    return lib().o_tvs_draw(int(seed), int(k), int(lo), int(hi))


def total_variation(P, img, mode=TVS_SUM_TILED_FP64):
    """calculate_total_variation (DS.cu:10165) with the given summation order."""
    img = np.ascontiguousarray(img, np.float32)
    return float(lib().o_tv(C.byref(P), _p(img), int(mode)))


def tvs_perturbation(P, img, mode=TVS_SUM_TILED_FP64):
    """generate_perturbation_array (DS.cu:10262): returns (G, v, norm)."""
    img = np.ascontiguousarray(img, np.float32)
    G = np.empty(P.num_voxels, np.float32)
    v = np.empty(P.num_voxels, np.float32)
    n = lib().o_tvs_perturbation(C.byref(P), _p(img), _p(G), _p(v), int(mode))
    return G, v, float(n)


def tvs_iteration(P, hull, x, k, L, mode=TVS_SUM_TILED_FP64):
    """NTVS_iteration (DS.cu:10583) for DROP iteration k (1-based).  Returns (x, L, tv_before, tv_after)."""
    x = np.ascontiguousarray(x, np.float32).copy()
    hull = np.ascontiguousarray(hull, np.uint8)
    Lc = C.c_int(int(L))
    tb, ta = C.c_float(), C.c_float()
    lib().o_tvs_iteration(C.byref(P), _p(hull), _p(x), int(k), C.byref(Lc), int(mode), C.byref(tb), C.byref(ta))
    return x, Lc.value, float(tb.value), float(ta.value)


def convert_units(P, t, v, u, wepl, fp64=False):
    """convert_mm_2_cm (DS.cu:4070) and apply_tuv_shifts (DS.cu:4090) in float as the reference does (in double
    for the FP64 pipeline)."""
    R = _real(fp64)

    def mm(a):
        return (a.astype(np.float64) * 0.1).astype(R) if P.data_in_mm else a.astype(R)

    def sh(a, s):
        return (a.astype(np.float64) + s).astype(R) if (P.t_shift or P.u_shift or P.v_shift) else a
    t = sh(mm(np.asarray(t, np.float32)), P.t_shift)
    u = sh(mm(np.asarray(u, np.float32)), P.u_shift)
    v = sh(mm(np.asarray(v, np.float32)), P.v_shift)
    return t, v, u, mm(np.asarray(wepl, np.float32))


def run_pipeline(P, data, iterations=None, stop_after=None, accum=1, with_fp64=False, tvs_sum=TVS_SUM_TILED_FP64,
                 hull_import=None, x0_import=None, fp64=False, timings=None, nthreads=1):
    """The reference pipeline on the CPU.  Returns a dict with every intermediate result.  hull_import and
    x0_import are the images HULL_TYPE = IMPORT and X_0_TYPE = IMPORT read (DS.cu:7358, :8614).

    fp64=True runs the FP64 twin of every stage (pct_oracle_fp64.c): per-history geometry, bin statistics,
    sinogram, FBP, x0 and the DROP iterate in double, the float rounding points of the reference taken
    exactly; the integer decisions (bins, cuts, hulls, endpoints, lists) follow from those values.

    timings: a dict that receives the wall time of each stage in seconds; nthreads: OpenMP threads of the DROP
    iterations (the CPU baseline's float atomic sums, order-dependent; 1 = history order)."""
    import time
    P = derive(P)
    out = {}
    clock = [time.perf_counter()]

    def lap(name):
        t = time.perf_counter()
        if timings is not None:
            timings[name] = timings.get(name, 0.0) + (t - clock[0])
        clock[0] = t
    R = _real(fp64)
    if fp64 and P.tvs_on:
        raise ValueError("the FP64 pipeline has no TVS twin")
    t, v, u, w = convert_units(P, data["t"], data["v"], data["u"], data["wepl"], fp64)
    h, missed = intersections(P, t, v, u, data["angle_deg"], fp64)
    lap("intersections")
    bin_num, in_range = binning(P, h, fp64)
    lap("binning")
    keep = ~missed
    ids = np.nonzero(keep)[0]  # loaded-history index of each surviving history
    h = np.ascontiguousarray(h[:, keep])
    w = np.ascontiguousarray(w[keep])
    bin_num = np.ascontiguousarray(bin_num[keep])
    in_range = np.ascontiguousarray(in_range[keep])
    out.update(h_volume=h, wepl_volume=w, bin_num=bin_num, in_range=in_range, missed=missed)
    if P.sc_on:
        out["sc_hull"] = carve(P, "sc", h, w, fp64)
    if P.msc_on:
        out["msc_counts"] = carve(P, "msc", h, w, fp64)
        out["msc_hull"] = msc_edge(P, out["msc_counts"])
    if P.sm_on:
        out["sm_counts"] = carve(P, "sm", h, w, fp64)
        out["sm_hull"] = sm_edge(P, out["sm_counts"])
    lap("hulls")
    if stop_after == "hulls":
        return out
    st = stats_and_cuts(P, bin_num, in_range, w, h, fp64)
    lap("stats_and_cuts")
    out.update(st)
    keep = ~st["failed"]
    ids = ids[keep]
    bin_cut = bin_num[keep]
    h = np.ascontiguousarray(h[:, keep])
    w = np.ascontiguousarray(w[keep])
    out.update(h_cut=h, wepl_cut=w)
    if stop_after == "cuts":
        return out
    sino = construct_sinogram(P, st["counts_post"], st["sinogram_sum"], fp64)
    out["sinogram"] = sino
    if P.fbp_on:
        out["sinogram_filtered"] = filter_sinogram(P, sino, fp64)
        out["fbp"] = backproject(P, out["sinogram_filtered"], fp64)
        fbp_x0 = out["fbp"]
        if P.median_filter_fbp and P.fbp_med_filter_radius > 0 and not P.avg_filter_fbp:  # DS.cu:6150
            fbp_x0 = out["fbp_median"] = median_filter_2d(P, out["fbp"], P.fbp_med_filter_radius, fp64)
        if P.hull_type == 4:
            out["fbp_hull"] = fbp_hull(P, out["fbp"], fp64)
    lap("fbp")
    hull = {0: lambda: np.asarray(hull_import).reshape(-1) != 0,
            1: lambda: out["sc_hull"], 2: lambda: (out["msc_hull"] == 1).astype(np.uint8),
            3: lambda: (out["sm_hull"] == 1).astype(np.uint8), 4: lambda: out["fbp_hull"]}[P.hull_type]()
    hull = np.ascontiguousarray(hull, np.uint8)
    out["hull"] = hull
    hull_ep = hull  # the device hull used by the endpoints (DS.cu:7375-7386)
    if P.avg_filter_hull and P.hull_avg_filter_radius > 0:
        hull_ep = out["hull_filtered"] = avg_filter_hull(P, hull, P.hull_avg_filter_radius,
                                                         P.hull_avg_filter_threshold)
    if P.x_0_type == 0:
        x0 = np.ascontiguousarray(x0_import, R).reshape(-1).copy()
    elif P.x_0_type == 1:
        x0 = hull.astype(R)
    elif P.x_0_type == 2:
        x0 = fbp_x0.copy()
    elif P.x_0_type == 3:
        x0 = np.where(hull != 0, fbp_x0, R(0)).astype(R)
    else:
        x0 = np.zeros(P.num_voxels, R)
    out["x0"] = x0
    if stop_after == "fbp":
        return out
    lap("hull_selection_x0")
    h2, first = endpoints(P, hull_ep, h, fp64)
    lap("endpoints")
    keep = first != P.num_voxels
    ids = ids[keep]
    out.update(bin_recon=np.ascontiguousarray(bin_cut[keep]),
               gantry_recon=np.ascontiguousarray(np.asarray(data["angle_deg"], np.float32)[ids]))
    h2 = np.ascontiguousarray(h2[:, keep])
    w2 = np.ascontiguousarray(w[keep])
    first = np.ascontiguousarray(first[keep])
    out.update(h_recon=h2, wepl_recon=w2, first_voxel=first)
    if stop_after == "endpoints":
        return out
    tabs = MlpTables(P)
    off, vox, chord = mlp_lists(P, tabs, h2, first, fp64)
    lap("mlp_lists")
    out.update(offsets=off, voxels=vox, chord=chord)
    if stop_after == "lists":
        return out
    its = P.iterations if iterations is None else iterations
    x = x0.copy()
    x64 = x0.astype(np.float64)
    xs = []
    L, tvs_log = 0, []
    for k in range(1, its + 1):  # DROP_GPU, DS.cu:9885-9915
        if fp64:
            x = drop_iteration_fp64(P, off, vox, chord, w2, x)
            xs.append(x.copy())
            continue
        if P.tvs_on and P.tvs_first:
            x, L, tb, ta = tvs_iteration(P, hull, x, k, L, tvs_sum)
            tvs_log.append((tb, ta, L))
        lap("tvs")
        x = drop_iteration(P, off, vox, chord, w2, x, accum=accum, nthreads=nthreads)
        lap("drop")
        if P.tvs_on and not P.tvs_first:
            x, L, tb, ta = tvs_iteration(P, hull, x, k, L, tvs_sum)
            tvs_log.append((tb, ta, L))
        if with_fp64:
            x64 = drop_iteration_fp64(P, off, vox, chord, w2, x64)
        xs.append(x.copy())
    out.update(x=x, x_iterations=xs, tvs_log=tvs_log)
    if with_fp64:
        out["x_fp64"] = x64
    return out


# ------------------------------------------------------------------ learned warm-start network
def net_num_params() -> int:
    return int(lib().on_num_params())


def net_loss_grad(shape, params, fbp, hull, target, bf16=False, with_grad=True):
    """FP64 loss, gradient and x0 of the warm-start network (pct_net_oracle.c); shape = (D, H, W).
    bf16=True rounds to bfloat16 where the GPU stores bfloat16."""
    D, H, W = shape
    p = np.ascontiguousarray(params, np.float32)
    f = np.ascontiguousarray(fbp, np.float32).reshape(-1)
    h = np.ascontiguousarray(hull, np.uint8).reshape(-1)
    t = np.ascontiguousarray(target, np.float32).reshape(-1)
    g = np.zeros(p.size, np.float64) if with_grad else None
    x0 = np.zeros(f.size, np.float64)
    loss = lib().on_loss_grad(D, H, W, _p(p), _p(f), _p(h), _p(t), int(bool(bf16)), _p(g) if g is not None else None,
                              _p(x0))
    return loss, g, x0.reshape(shape)
