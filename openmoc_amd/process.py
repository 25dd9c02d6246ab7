"""Post-processing helpers mirroring openmoc.process for the hot path's
outputs (reference openmoc/process.py:48-97)."""
import numpy as np


def get_scalar_fluxes(solver, fsrs="all", groups="all"):
    """(num_fsrs, num_groups) array of FSR scalar fluxes, as the reference's
    ``get_scalar_fluxes`` returns from ``solver.getFluxes`` (process.py:90-97)."""
    n, g = solver._num_fsrs, solver._num_groups
    fluxes = solver.getFluxes(n * g).reshape(n, g)
    if not (isinstance(fsrs, str) and fsrs == "all"):
        fluxes = fluxes[np.asarray(fsrs)]
    if not (isinstance(groups, str) and groups == "all"):
        fluxes = fluxes[:, np.asarray(groups) - 1]
    return fluxes
