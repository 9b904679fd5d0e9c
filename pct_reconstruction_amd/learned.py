"""Learned warm start for DROP: a 3D U-Net from (FBP image, hull) to x0, trained on the GPU.

The reference's initial iterate is the FBP image, the hull or their product
(define_initial_iterate, src/pCT_Reconstruction_Data_Segments_Blake.cu:8612-8646).  `WarmStartNet`
wraps the pct_net_* entry points of lib/libpct.so (kernels in csrc/k_net.hip): training data are
(FBP, hull, DROP image after many iterations) triples that the pipeline itself produces from seeded
phantoms; `apply` installs the net's x0 in a reconstruction only when its data residual ||Ax - b||
is below the FBP start's, so DROP still defines the final image.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import lib, ptr


class WarmStartNet:
    def __init__(self, columns: int, rows: int, slices: int, seed: int = 0, device: int = 0):
        self.shape = (slices, rows, columns)
        self._net = C.c_void_p()
        rc = lib().pct_net_create(columns, rows, slices, seed, device, C.byref(self._net))
        if rc != _lib.PCT_OK:
            raise RuntimeError(f"pct_net_create failed: {_lib.ERRORS.get(rc, rc)} (grid must be multiples of 8; "
                               "is a HIP device visible?)")

    def close(self) -> None:
        if self._net:
            lib().pct_net_destroy(self._net)
            self._net = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc != _lib.PCT_OK:
            detail = lib().pct_net_last_error(self._net)
            raise RuntimeError(f"{what}: {_lib.ERRORS.get(rc, rc)}: {detail.decode(errors='replace') if detail else ''}")

    def _img(self, a, dtype) -> np.ndarray:
        b = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
        if b.size != int(np.prod(self.shape)):
            raise ValueError(f"image must have {self.shape} voxels")
        return b

    @property
    def num_params(self) -> int:
        return int(lib().pct_net_num_params(self._net))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.num_params, np.float32)
        self._check(lib().pct_net_get_params(self._net, ptr(out), out.size), "get_params")
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        self._check(lib().pct_net_set_params(self._net, ptr(p), p.size), "set_params")

    def loss_grad(self, fbp, hull, target, with_grad: bool = True):
        f, h, t = self._img(fbp, np.float32), self._img(hull, np.uint8), self._img(target, np.float32)
        loss = C.c_double()
        g = np.empty(self.num_params, np.float32) if with_grad else None
        self._check(lib().pct_net_loss_grad(self._net, ptr(f), ptr(h), ptr(t), C.byref(loss),
                                            ptr(g) if g is not None else None), "loss_grad")
        return loss.value, g

    def predict(self, fbp, hull) -> np.ndarray:
        f, h = self._img(fbp, np.float32), self._img(hull, np.uint8)
        out = np.empty(f.size, np.float32)
        self._check(lib().pct_net_predict(self._net, ptr(f), ptr(h), ptr(out)), "predict")
        return out.reshape(self.shape)

    def add_sample(self, fbp, hull, target) -> None:
        f, h, t = self._img(fbp, np.float32), self._img(hull, np.uint8), self._img(target, np.float32)
        self._check(lib().pct_net_add_sample(self._net, ptr(f), ptr(h), ptr(t)), "add_sample")

    @property
    def num_samples(self) -> int:
        return int(lib().pct_net_num_samples(self._net))

    def train(self, steps: int, lr: float = 1e-3, seed: int = 0) -> float:
        loss = C.c_double()
        self._check(lib().pct_net_train(self._net, int(steps), float(lr), int(seed), C.byref(loss)), "train")
        return loss.value

    def save(self, path: str | Path) -> None:
        self._check(lib().pct_net_save(self._net, str(path).encode()), "save")

    def load(self, path: str | Path) -> None:
        self._check(lib().pct_net_load(self._net, str(path).encode()), "load")

    def apply_iterate(self, recon) -> dict:
        """Replace the current iterate of `recon` by net(x, hull) when that lowers ||Ax - b||."""
        used, rx, rn = C.c_int(), C.c_double(), C.c_double()
        self._check(lib().pct_net_apply_iterate(self._net, recon._ctx, C.byref(used), C.byref(rx), C.byref(rn)),
                    "apply_iterate")
        return {"used_net": bool(used.value), "residual_x": rx.value, "residual_net": rn.value}

    def apply(self, recon) -> dict:
        """Install the net's x0 in `recon` (after build_mlp_lists) when it lowers ||Ax - b||."""
        used, rf, rn = C.c_int(), C.c_double(), C.c_double()
        self._check(lib().pct_net_apply(self._net, recon._ctx, C.byref(used), C.byref(rf), C.byref(rn)), "apply")
        return {"used_net": bool(used.value), "residual_fbp": rf.value, "residual_net": rn.value}
