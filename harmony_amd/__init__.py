"""harmony_amd: Harmony's per-node IVF-flat search path, native on one AMD MI355X (gfx950).

The product is libharmony_amd.so (C-ABI in include/harmony_ivf.h): host C++ plus hand-written
HIP kernels. This package holds the sources, the in-tree build, and a Python mirror of the
reference's harmony::Index class over the C-ABI.
"""
from ._lib import HV_PRECISION_BF16, HV_PRECISION_FP8, HV_PRECISION_FP32, HipExtensionMissing, HvError, lib  # noqa: F401
from .index import DeviceArray, Index  # noqa: F401
from . import datagen, learned, metrics  # noqa: F401
