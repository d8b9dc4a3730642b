"""MI355X-native two-scale WCSPH solver (hot path of segfault11/TwoScaleSplit's
WCSPHSolver): hand-written HIP kernels for gfx950 behind the C ABI of
include/tss.h (libtss.so), with the reference's class API kept by
twoscalesplit_amd/host/WCSPHSolver.h and this thin Python binding.
"""
from ._native import (DT_CFL, DT_FIXED, FIELD, FIELDS, MODE_BAND, MODE_PLAIN,
                      MODE_REFERENCE, TssConfig, TssError, lib)
from .solver import Batch, Solver, config_init

__all__ = ["Batch", "Solver", "config_init", "lib", "TssConfig", "TssError",
           "FIELD", "FIELDS", "MODE_PLAIN", "MODE_REFERENCE", "MODE_BAND",
           "DT_FIXED", "DT_CFL"]
