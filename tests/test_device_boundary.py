"""The kernel-level DROP boundary pct_drop_iteration_device (DROP_full_tx_iteration, DS.cu:9537-9589,
on caller-owned device buffers).

CPU: argument checks return PCT_E_ARG before any device call.  GPU: buffers allocated with the HIP
runtime through ctypes; two iterations against the oracle's FP64-sum iteration, with the scratch sums
and counts returned to zero.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from cases import small_cylinder


def test_device_boundary_rejects_bad_arguments():
    from pct_reconstruction_amd._lib import lib
    f = lib().pct_drop_iteration_device
    null = None
    assert f(-1, null, null, null, null, null, null, null, 10, 100, 10, 1, 1, 1e-4, null) == -1
    assert f(10, null, null, null, null, null, null, null, 10, 100, 10, 1, 1, 1e-4, null) == -1
    p = C.c_void_p(16)  # never dereferenced: the size checks fail first
    assert f(10, p, p, p, p, p, p, p, 0, 100, 10, 1, 1, 1e-4, null) == -1
    assert f(10, p, p, p, p, p, p, p, 10, 0, 10, 1, 1, 1e-4, null) == -1


def _hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.fail("HIP runtime library not found")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMalloc.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipFree.restype = C.c_int
    h.hipDeviceSynchronize.restype = C.c_int
    return h


class _Dev:
    """Device copies of numpy arrays (hipMemcpyHostToDevice = 1, DeviceToHost = 2)."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)


@pytest.mark.gpu
def test_device_boundary_matches_the_oracle(oracle):
    from pct_reconstruction_amd._lib import lib
    P, data = small_cylinder(n=30000, seed=4)
    ref = oracle.run_pipeline(P, data, stop_after="lists")
    off, vox, chord, w, x0 = ref["offsets"], ref["voxels"], ref["chord"], ref["wepl_recon"], ref["x0"]
    n, NV = off.shape[0] - 1, P.num_voxels
    print(f"[device boundary] {n} histories, {vox.size} entries", flush=True)
    hip = _hip()
    d = _Dev(hip)
    try:
        d_off, d_vox, d_chord, d_w = d.put(off), d.put(vox), d.put(chord), d.put(w)
        d_x = d.put(x0)
        d_acc = d.put(np.zeros(NV, np.float64))
        d_cnt = d.put(np.zeros(NV, np.uint32))
        x_ref = x0.copy()
        for it in range(2):
            rc = lib().pct_drop_iteration_device(n, d_off, d_vox, d_chord, d_w, d_x, d_acc, d_cnt, NV,
                                                 P.drop_block_size, P.histories_per_block, P.quirk_a3_drop_floor,
                                                 P.min_mlp_length, P.lambda_, None)
            assert rc == 0
            assert hip.hipDeviceSynchronize() == 0
            x_ref = oracle.drop_iteration(P, off, vox, chord, w, x_ref, accum=1)
            x = d.get(d_x, x0)
            rel = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
            print(f"[device boundary] iteration {it + 1}: rel L2 vs oracle {rel:.2e}", flush=True)
            assert rel <= 1e-6
            assert not d.get(d_acc, np.zeros(NV, np.float64)).any()
            assert not d.get(d_cnt, np.zeros(NV, np.uint32)).any()
    finally:
        d.free()
