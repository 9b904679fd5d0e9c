"""MI355X-native proton-CT reconstruction path (binning/cuts, hulls, FBP, MLP, DROP).

The compute lives in lib/libpct.so (hand-written gfx950 HIP kernels behind the C-ABI of
include/pct.h); this package is its host-side mirror of the reference's stage functions.
"""
from ._lib import LIB_PATH, Params, lib  # noqa: F401
from .recon import Reconstruction, default_params, derive, load_settings, set_key  # noqa: F401

__all__ = ["Reconstruction", "Params", "default_params", "derive", "load_settings", "set_key", "lib", "LIB_PATH"]
