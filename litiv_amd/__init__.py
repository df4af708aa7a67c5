"""litiv_amd: MI355X-native (gfx950) SuBSENSE / LOBSTER background subtraction.

The compute path is the HIP library liblitiv_bgs.so behind the C ABI of include/litiv_bgs.h;
this package is its Python mirror of litiv's IBackgroundSubtractor interface. Importing it
loads the library and fails loudly if it has not been built (there is no CPU fallback).
"""
from .bgs import (BackgroundSubtractorLOBSTER, BackgroundSubtractorSuBSENSE, Batch, BinClassif,  # noqa: F401
                  Server, lbsp_compute_abs, library_path)

__all__ = ["BackgroundSubtractorSuBSENSE", "BackgroundSubtractorLOBSTER", "Batch", "BinClassif", "Server",
           "lbsp_compute_abs", "library_path"]
