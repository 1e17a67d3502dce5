"""MI355X-native MHIT36 timestep: HIP kernels behind the C ABI of include/mhit.h.

Importing the package does not touch the GPU; `Solver` loads libmhit.so and
creates a device context.
"""
from ._lib import (MHIT_OK, MHIT_ERR_CFL, MHIT_ERR_NAN, MHIT_FP64, MHIT_MIXED,
                   MHIT_K_REFERENCE, MHIT_K_MODIFIED, MhitError, Params)
from .solver import Solver, default_params, read_input, derive, step_batch, FIELD_NAMES
from .ensemble import Ensemble

__all__ = ["Solver", "Ensemble", "step_batch", "default_params", "read_input", "derive", "FIELD_NAMES", "MhitError",
           "Params", "MHIT_OK", "MHIT_ERR_CFL", "MHIT_ERR_NAN", "MHIT_FP64", "MHIT_MIXED",
           "MHIT_K_REFERENCE", "MHIT_K_MODIFIED"]
