"""cuda_fot_amd -- MI355X-native edge-centred indirect-increment loop engine.

The compute path is libfot.so (hand-written HIP kernels for gfx950 plus the
native host planner) behind the C-ABI in include/fot_capi.h.  This package
is a thin Python binding over that library; it has no CPU fallback.
"""
from . import capi, meshgen, plan  # noqa: F401
from .capi import FotError, InvalidInputFile, ColourOverflow  # noqa: F401
from .device import Batch, DeviceArray, Loop, Stream, Timer, synchronize  # noqa: F401
from .problem import (Problem, getacc_problem, gpu_batch, lulesh_problem, mini_aero_problem,  # noqa: F401
                      res_calc_problem, volna_problem)

__all__ = ["capi", "meshgen", "plan", "FotError", "InvalidInputFile", "ColourOverflow",
           "Batch", "DeviceArray", "Loop", "Stream", "Timer", "synchronize", "Problem",
           "getacc_problem", "gpu_batch", "lulesh_problem", "mini_aero_problem",
           "res_calc_problem",
           "volna_problem"]
