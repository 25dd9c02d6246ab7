"""Adjoint solves (Solver::setSolverMode(ADJOINT), src/Solver.cpp:597-602,
applied in initializeMaterials, src/Solver.cpp:792-812): the sources use the
transposed scattering and fission matrices (Material::
transposeProductionMatrices, src/Material.cpp:1091-1103), while keff, the
normalisation and the fission-source residual keep nu-sigma-f.  Compared with
the compiled reference's adjoint solves on the LDS sweep (7 groups) and with
the matrix-core contraction (70 groups)."""
import numpy as np
import pytest

import openmoc_amd as om

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", ["pin_cell_7g", "lattice_3x3_vac_7g", "pin_cell_70g"])
def test_adjoint_matches_reference(tracks, golden, name):
    ref = golden(f"{name}_adjoint")
    s = om.HipSolver(tracks(name))
    s.setSolverMode(om.ADJOINT)
    s.computeEigenvalue(1000)
    assert abs(s.getKeff() - ref["keff"]) < 1e-5                  # 1 pcm
    assert rel_l2(s.getFluxes(), ref["flux"]) < 1e-4
    assert abs(s.getNumIterations() - int(ref["num_iterations"])) <= 1
    # and back: the forward solve of the same solver
    fwd = golden(name)
    s.setSolverMode(om.FORWARD)
    s.computeEigenvalue(1000)
    assert abs(s.getKeff() - fwd["keff"]) < 1e-5
    assert rel_l2(s.getFluxes(), fwd["flux"]) < 1e-4


def test_restart_across_a_solver_mode_switch(tracks, golden):
    """setSolverMode only records the mode (src/Solver.cpp:597-602): after a
    forward solve, setSolverMode(ADJOINT) with setRestartStatus(True) starts
    the adjoint solve from the forward fluxes and reaches the adjoint
    solution; switching back and restarting reaches the forward one."""
    name = "lattice_3x3_vac_7g"
    adj, fwd = golden(f"{name}_adjoint"), golden(name)
    s = om.HipSolver(tracks(name))
    s.computeEigenvalue(1000)
    s.setSolverMode(om.ADJOINT)
    s.setRestartStatus(True)
    s.computeEigenvalue(1000)
    assert abs(s.getKeff() - adj["keff"]) < 1e-5
    assert rel_l2(s.getFluxes(), adj["flux"]) < 1e-4
    s.setSolverMode(om.FORWARD)
    s.computeEigenvalue(1000)
    assert abs(s.getKeff() - fwd["keff"]) < 1e-5
    assert rel_l2(s.getFluxes(), fwd["flux"]) < 1e-4


def test_adjoint_scatter_precisions_agree(tracks):
    tg = tracks("pin_cell_70g")
    out = []
    for mode in (om.SCATTER_FP64, om.SCATTER_MFMA_F16, om.SCATTER_FP32):
        s = om.HipSolver(tg)
        s.setScatterPrecision(mode)
        s.setSolverMode(om.ADJOINT)
        s.computeEigenvalue(1000)
        out.append((s.getKeff(), s.getFluxes()))
    for k, f in out[1:]:
        assert abs(k - out[0][0]) < 1e-6
        assert rel_l2(f, out[0][1]) < 1e-5


def test_unknown_solver_mode(tracks):
    s = om.HipSolver(tracks("pin_cell_7g"))
    with pytest.raises(om.MOCError):
        s.setSolverMode(2)


def test_reference_getters(tracks):
    tg = tracks("pin_cell_7g")
    s = om.HipSolver(tg)
    s.setConvergenceThreshold(3e-6)
    assert s.getConvergenceThreshold() == 3e-6
    assert s.getTrackGenerator() is tg
    assert s.getNumPolarAngles() == tg.info.num_polar
    assert not s.is3D() and om.HipSolver(tracks("pin_cell_3d_7g")).is3D()
    assert s.getFSRVolume(0) == tg.fsr_volume[0]
    with pytest.raises(om.MOCError):
        s.getFSRVolume(tg.getNumFSRs())
    assert s.getMaxOpticalLength() == tg.getMaxOpticalLength()
    assert not s.isUsingDoublePrecision() and not s.isUsingExponentialInterpolation()
    s.setNumThreads(8)
    s.useExponentialIntrinsic()
    s.computeEigenvalue(1000)
    assert s.getTotalTime() > 0
    ref = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "pin_cell_7g_ref.npz")
    assert abs(s.getKeff() - float(ref["keff"])) < 1e-5      # hardware exp within 1 pcm


def test_chi_spectrum_initial_guess_matches_reference(tracks, golden):
    """setChiSpectrumMaterial (src/Solver.cpp:708-710, 1722-1725): the solve
    starts from the material's chi in every FSR."""
    ref = golden("lattice_3x3_vac_7g_chi_spectrum")
    tg = tracks("lattice_3x3_vac_7g")
    s = om.HipSolver(tg)
    s.setChiSpectrumMaterial(int(ref["material"]))
    s.computeEigenvalue(1000)
    assert abs(s.getNumIterations() - int(ref["num_iterations"])) <= 1
    assert abs(s.getKeff() - ref["keff"]) < 1e-6
    assert rel_l2(s.getFluxes(), ref["flux"]) < 1e-4
    # the same material by name gives the same solve, bit for bit
    k_idx, f_idx = s.getKeff(), s.getFluxes()
    s.setChiSpectrumMaterial("fuel")
    s.computeEigenvalue(1000)
    assert s.getKeff() == k_idx
    np.testing.assert_array_equal(s.getFluxes(), f_idx)
    s.setChiSpectrumMaterial(None)         # flat guess again
    s.computeEigenvalue(1000)
    assert abs(s.getKeff() - golden("lattice_3x3_vac_7g")["keff"]) < 1e-6
    with pytest.raises(om.MOCError):
        s.setChiSpectrumMaterial(99)
