"""heliosk -- MI355X-native line-by-line opacity engine (HELIOS-K hot path).

Host-side Python over the C ABI of lib/libheliosk.so (include/heliosk.h).
The heavy work runs in hand-written HIP kernels for gfx950.
"""
from ._lib import HKError, HK_VOIGT_FP32C, HK_VOIGT_FP64, lib  # noqa: F401

__version__ = "0.1"
