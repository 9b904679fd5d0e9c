"""pct_drop_iteration_device_noalloc: the DROP iteration of pct_drop_iteration_device (DROP_full_tx_iteration,
DS.cu:9537-9589, on caller-owned device buffers) with the per-voxel sums and counts in the library's own
device array instead of caller workspace.

CPU: argument and size-limit checks return before any device call.  GPU: two calls in a row (the second on
a smaller input, a smaller grid and fewer histories) against the workspace entry point on the same inputs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from cases import small_cylinder

NOALLOC_MAX_VOXELS = 1 << 23


def test_noalloc_rejects_bad_arguments_and_oversize_grids():
    from pct_reconstruction_amd._lib import lib
    f = lib().pct_drop_iteration_device_noalloc
    null = None
    assert f(-1, null, null, null, null, null, 10, 100, 10, 1, 1, 1e-4, null) == -1
    p = C.c_void_p(16)  # never dereferenced: the checks fail first
    assert f(10, p, p, p, p, p, 0, 100, 10, 1, 1, 1e-4, null) == -1
    assert f(10, p, p, p, p, p, 10, 0, 10, 1, 1, 1e-4, null) == -1
    assert f(10, p, p, p, p, p, NOALLOC_MAX_VOXELS + 1, 100, 10, 1, 1, 1e-4, null) == -6


@pytest.mark.gpu
def test_noalloc_matches_the_workspace_entry_point(oracle):
    from pct_reconstruction_amd._lib import lib
    from test_device_boundary import _Dev, _hip
    hip = _hip()
    cases = [small_cylinder(n=30000, seed=4), small_cylinder(n=9000, seed=8, columns=48, rows=48)]
    for call, (P, data) in enumerate(cases):
        ref = oracle.run_pipeline(P, data, stop_after="lists")
        off, vox, chord, w, x0 = ref["offsets"], ref["voxels"], ref["chord"], ref["wepl_recon"], ref["x0"]
        n, NV = off.shape[0] - 1, P.num_voxels
        print(f"[noalloc] call {call + 1}: {n} histories, {vox.size} entries, {NV} voxels", flush=True)
        d = _Dev(hip)
        try:
            d_off, d_vox, d_chord, d_w = d.put(off), d.put(vox), d.put(chord), d.put(w)
            d_x_ws, d_x_na = d.put(x0), d.put(x0)
            d_acc = d.put(np.zeros(NV, np.float64))
            d_cnt = d.put(np.zeros(NV, np.uint32))
            args = (NV, P.drop_block_size, P.histories_per_block, P.quirk_a3_drop_floor, P.min_mlp_length, P.lambda_,
                    None)
            assert lib().pct_drop_iteration_device(n, d_off, d_vox, d_chord, d_w, d_x_ws, d_acc, d_cnt, *args) == 0
            assert lib().pct_drop_iteration_device_noalloc(n, d_off, d_vox, d_chord, d_w, d_x_na, *args) == 0
            assert hip.hipDeviceSynchronize() == 0
            x_ws, x_na = d.get(d_x_ws, x0), d.get(d_x_na, x0)
            rel = float(np.linalg.norm(x_na - x_ws) / np.linalg.norm(x_ws))
            print(f"[noalloc] call {call + 1}: rel L2 vs workspace entry {rel:.2e}", flush=True)
            assert np.isfinite(x_na).all()
            assert rel <= 1e-6  # FP64 per-voxel sums: only the order of the atomic adds differs
            assert not np.array_equal(x_na, x0)
        finally:
            d.free()
