"""MI355X-native 3+1D glasma engine (SU(3), FP64) — HIP kernels behind a C ABI (libglasma.so).

The device library is the only execution path: importing works without a GPU, but every compute call
goes through libglasma.so and raises GlasmaError if it is not built or no device is present.
"""
from ._lib import GlasmaError, header_functions, lib  # noqa: F401
from .engine import Glasma, GlasmaBatch, bulk_from_raw, fft_host, su3_batch  # noqa: F401

__all__ = ["Glasma", "GlasmaBatch", "GlasmaError", "bulk_from_raw", "fft_host", "su3_batch", "lib", "header_functions"]
