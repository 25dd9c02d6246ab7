"""openmoc_amd: the OpenMOC method-of-characteristics source iteration
(transport sweep, FSR sources, scalar-flux update, keff, normalisation,
residual) on one AMD MI355X (gfx950) in hand-written HIP kernels, behind the
reference Solver interface.

The HIP library is built in-tree (``python -m openmoc_amd.build``); there is
no CPU fallback for the solver.
"""
from .workloads import Problem, get as get_workload  # noqa: F401
from .solver import (ADJOINT, DIAGONAL, EXP_HW, EXP_RATIONAL, FISSION_SOURCE, FORWARD,  # noqa: F401
                     GLOBAL,
                     SCALAR_FLUX, SCATTER_AUTO, SCATTER_FP32, SCATTER_FP64, SCATTER_MFMA_F16,
                     TOTAL_SOURCE, YAMAMOTO, GPUSolver, HipLSSolver, HipSolver, MOCError,
                     TrackGenerator, fsr_centroids, match_centroids, read_flux_file,
                     write_flux_file)
from . import process  # noqa: F401
