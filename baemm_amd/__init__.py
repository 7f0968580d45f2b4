"""baemm_amd -- MI355X-native brute-force boundary-element hot path of BAEMM.

The numerics live in ``libbaemm_hip.so`` (HIP kernels for gfx950 behind the C
ABI in ``include/baemm.h``).  This package holds its Python binding
(:class:`Helmholtz_HIP`, a mirror of ``BAEMM::Helmholtz_OpenCL``) and mesh
utilities.  Importing the package does not load the GPU library; creating a
:class:`Helmholtz_HIP` does, and fails loudly if it is missing.
"""
from ._lib import BaemmError, LIB_PATH
from .helmholtz import (ApplyBoundaryOperatorsBatch, BatchAutotune, BatchStats, BatchTuning, Helmholtz_HIP,
                        SolveLoadedBatch, WaveType)
from . import meshes
from .serving import GNServer

__all__ = ["Helmholtz_HIP", "WaveType", "BaemmError", "LIB_PATH", "meshes", "ApplyBoundaryOperatorsBatch",
           "BatchAutotune", "BatchTuning", "BatchStats", "GNServer", "SolveLoadedBatch"]
