"""doma_amd -- MI355X-native discrete-ordinates sweep and scattering source
iteration (DOMA fixed-source path) on hand-written HIP kernels for gfx950.

The compute path is the native library ``libdoma_amd.so`` (C ABI in
``include/doma_c.h``); this package only binds it.  There is no CPU fallback:
every entry point raises when the library or a GPU is unavailable.
"""
from .capi import (  # noqa: F401
    FP32_SWEEP, FP64, SCHEME_DD, SCHEME_TW_DD, Batch, Context, DomaError, MediaSpec, Problem,
    deck_read, gc_quadrature, gl_quadrature, lib, media_generate, run_deck,
)

__all__ = [
    "Batch", "Context", "DomaError", "MediaSpec", "Problem", "gc_quadrature", "gl_quadrature", "lib",
    "media_generate", "deck_read", "run_deck", "FP64", "FP32_SWEEP", "SCHEME_DD", "SCHEME_TW_DD",
]
