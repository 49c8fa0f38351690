"""lbfast_amd -- MI355X-native moment-represented lattice-Boltzmann time step.

The compute path is liblbfast_hip.so (hand-written HIP kernels for gfx950)
behind the C-ABI in include/lbfast_hip.h; this package is the host-side
mirror of the reference driver interface (lb_cuda_driver_module.f90) over
that C-ABI.  Importing the package does not load the library; the first
use of :mod:`lbfast_amd.hip` does, and raises if it is missing.
"""
from .params import (LbfParams, make_params, derive, PREC_FP64, PREC_FP32, PREC_MIXED,  # noqa: F401
                     PREC_FP16S, PREC_BF16S)

__all__ = ["LbfParams", "make_params", "derive"]
