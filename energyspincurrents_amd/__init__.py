"""energyspincurrents_amd -- MI355X-native Lanczos dynamical-correlation engine.

Compute lives in libesc.so (HIP kernels for gfx950 behind the C ABI of
include/esc.h); this package is the Python binding of that ABI.
"""
from ._capi import EscError, lib  # noqa: F401
from .grad import SpectrumLoss, TableLoss, spectrum_loss  # noqa: F401
from .engine import (Batch, Chain, Engine, Honeycomb, LanczosResult, SpinModel, cf_eval, cf_spectrum,  # noqa: F401
                     field111, kernel_plan, lehmann_spectrum, load_params, precompile_kernel, reference_omega,
                     tridiag_eig, tridiag_eigenpairs, tridiag_lowest)

__all__ = ["SpectrumLoss", "TableLoss", "spectrum_loss", "Batch", "Chain", "Engine", "Honeycomb", "SpinModel", "LanczosResult", "EscError", "lib", "cf_spectrum",
           "kernel_plan", "lehmann_spectrum", "precompile_kernel", "tridiag_eig", "load_params", "field111", "reference_omega",
           "tridiag_lowest", "tridiag_eigenpairs", "cf_eval"]
