"""monofonic_amd -- MI355X-native (gfx950) nLPT field pipeline of monofonIC.

The compute path is the shared library ``libmfx.so`` (HIP kernels behind the C
ABI of ``include/mfx.h``); this package is its Python host mirror.
"""
from ._lib import LIB_PATH, MfxError, load  # noqa: F401
from .core import (  # noqa: F401
    PLT, Batch, Context, CosmologyCalculator, GENERATORS, Grid_FFT, ICGenerator, OrszagConvolver, RNG_music, RNG_ngenic,
    RNG_panphasia, RNG_threefry, assemble_kspace, kspace_id, mt19937_raw, noise_normalise, op, panphasia_info,
    panphasia_tables, phi1, ranlxd1_raw, rspace_id, select_RNG_plugin, threefry_uniforms)

__all__ = [
    "Context", "CosmologyCalculator", "Grid_FFT", "ICGenerator", "OrszagConvolver", "RNG_music", "RNG_ngenic",
    "RNG_threefry", "RNG_panphasia", "panphasia_info", "op", "kspace_id", "rspace_id", "load", "MfxError", "PLT", "Batch",
]
