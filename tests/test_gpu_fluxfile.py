"""Flux files on the GPU solver: dumpFSRFluxes / loadFSRFluxes round trips,
restarts from a flux file (loadInitialFSRFluxes: normalised loaded fluxes,
30 sweeps on their sources, then the source iteration; src/Solver.cpp:
1710-1751) and residuals against a reference flux (setResidualByReference,
src/CPUSolver.cpp:2113-2252).

The expected values are the compiled reference's (double build) on the same
3x3 leaking lattice: its flux files and its keff / fluxes / iteration counts
after the same calls (tests/golden/lattice_3x3_vac_7g_fluxfile_ref.npz).
Five iterations after a restart from a 10-iteration file are not converged,
so they depend on every step of the restart: the loaded k_eff, the
normalisation, the startup sweeps and the first sources."""
from pathlib import Path

import numpy as np
import pytest

import openmoc_amd as om

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
NAME = "lattice_3x3_vac_7g"
IT10 = GOLDEN / f"{NAME}_it10.fluxes"
CONV = GOLDEN / f"{NAME}_ref.fluxes"


@pytest.fixture(scope="module")
def ref():
    return dict(np.load(GOLDEN / f"{NAME}_fluxfile_ref.npz"))


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_dump_holds_keff_centroids_and_fluxes(tracks, tmp_path):
    tg = tracks(NAME)
    s = om.HipSolver(tg)
    s.computeEigenvalue(1000)
    f = tmp_path / "out.fluxes"
    s.dumpFSRFluxes(f)
    k, xyz, flux = om.read_flux_file(f)
    assert k == s.getKeff()
    np.testing.assert_array_equal(flux.ravel(), s.getFluxes())
    np.testing.assert_array_equal(xyz, om.fsr_centroids(tg))


def test_load_into_a_fresh_solver(tracks, tmp_path, ref):
    tg = tracks(NAME)
    s = om.HipSolver(tg)
    s.computeEigenvalue(1000)
    f = tmp_path / "out.fluxes"
    s.dumpFSRFluxes(f)
    t = om.HipSolver(tg)
    assert t.loadFSRFluxes(f, assign_k_eff=True) == 0.0
    assert t.getKeff() == s.getKeff()
    np.testing.assert_array_equal(t.getFluxes(), s.getFluxes())
    # the reference's file: fluxes onto this solver's FSRs, k_eff unchanged
    u = om.HipSolver(tg)
    err = u.loadFSRFluxes(CONV)
    assert err < 1e-7
    np.testing.assert_array_equal(u.getFluxes(), ref["flux"])


def test_restart_five_iterations_matches_reference(tracks, ref):
    s = om.HipSolver(tracks(NAME))
    s.loadInitialFSRFluxes(IT10)
    s.computeEigenvalue(5)
    assert s.getNumIterations() == int(ref["restart5_num_iterations"]) == 5
    assert abs(s.getKeff() - ref["restart5_keff"]) < 1e-7
    assert rel_l2(s.getFluxes(), ref["restart5_flux"]) < 1e-5


def test_restart_to_convergence_matches_reference(tracks, ref):
    s = om.HipSolver(tracks(NAME))
    s.loadInitialFSRFluxes(IT10)
    s.computeEigenvalue(1000)
    assert abs(s.getNumIterations() - int(ref["restart_num_iterations"])) <= 1
    assert abs(s.getKeff() - ref["restart_keff"]) < 1e-6
    assert rel_l2(s.getFluxes(), ref["restart_flux"]) < 1e-4
    # switched off again: a plain solve
    s.loadInitialFSRFluxes(None)
    s.computeEigenvalue(1000)
    assert abs(s.getNumIterations() - int(ref["num_iterations"])) <= 1
    assert abs(s.getKeff() - ref["keff"]) < 1e-6


def test_residual_by_reference_matches_reference(tracks, ref):
    s = om.HipSolver(tracks(NAME))
    s.setResidualByReference(CONV)
    s.computeEigenvalue(1000)
    assert abs(s.getNumIterations() - int(ref["residual_ref_num_iterations"])) <= 1
    assert abs(s.getKeff() - ref["residual_ref_keff"]) < 1e-6
    assert rel_l2(s.getFluxes(), ref["residual_ref_flux"]) < 1e-4
    s.setResidualByReference(None)
    s.computeEigenvalue(1000)
    assert abs(s.getNumIterations() - int(ref["num_iterations"])) <= 1


def test_restart_with_linear_source_reaches_the_same_solution(tracks):
    tg = tracks(NAME)
    plain = om.HipLSSolver(tg)
    plain.computeEigenvalue(1000)
    s = om.HipLSSolver(tg)
    s.loadInitialFSRFluxes(IT10)
    s.computeEigenvalue(1000)
    assert abs(s.getKeff() - plain.getKeff()) < 1e-5
    assert rel_l2(s.getFluxes(), plain.getFluxes()) < 1e-3


def test_restart_errors(tracks, tmp_path):
    s = om.HipSolver(tracks(NAME))
    s.loadInitialFSRFluxes(tmp_path / "absent.fluxes")
    with pytest.raises(om.MOCError) as e:
        s.computeEigenvalue(10)
    assert e.value.code == -6
    # another geometry's file: FSR count mismatch
    s.loadInitialFSRFluxes(None)
    other = om.HipSolver(tracks("pin_cell_7g"))
    other.computeEigenvalue(1000)
    f = tmp_path / "pin.fluxes"
    other.dumpFSRFluxes(f)
    with pytest.raises(om.MOCError) as e:
        s.loadFSRFluxes(f)
    assert e.value.code == -1
    s.computeEigenvalue(1000)        # still usable
    assert s.getNumIterations() > 0


@pytest.mark.parametrize("name", ["pin_cell_3d_7g", "pin_cell_70g"])
def test_restart_and_reference_residual_on_other_sweep_paths(tracks, name):
    """3D (z-stack sweep, int64 global tally) and 70 groups (global-atomic
    sweep, source residual terms per FSR), from the reference's converged
    file.  A restart from a converged file stops after one or two
    iterations, a few pcm from the converged eigenvalue (the 30 startup
    sweeps leave the angular fluxes short of convergence); the reference
    does the same, and the values are compared with its own."""
    ref = dict(np.load(GOLDEN / f"{name}_fluxfile_ref.npz"))
    f = GOLDEN / f"{name}_ref.fluxes"
    tg = tracks(name)
    r = om.HipSolver(tg)
    r.loadInitialFSRFluxes(f)
    r.computeEigenvalue(1000)
    assert abs(r.getNumIterations() - int(ref["restart_num_iterations"])) <= 1
    assert abs(r.getKeff() - ref["restart_keff"]) < 1e-6
    assert rel_l2(r.getFluxes(), ref["restart_flux"]) < 1e-4
    q = om.HipSolver(tg)
    q.setResidualByReference(f)
    q.computeEigenvalue(1000)
    assert abs(q.getNumIterations() - int(ref["residual_ref_num_iterations"])) <= 1
    assert abs(q.getKeff() - ref["residual_ref_keff"]) < 1e-6
    assert rel_l2(q.getFluxes(), ref["residual_ref_flux"]) < 1e-4


def test_restart_status_continues_the_previous_solve(tracks, ref):
    """setRestartStatus (src/Solver.cpp:610-615): 10 iterations, then a solve
    that keeps k_eff, the angular and the scalar fluxes.  The reference stops
    after 1 + 8 iterations with the eigenvalue of the uninterrupted 18; after
    a converged solve it runs 1 + 1 more."""
    tg = tracks(NAME)
    full = om.HipSolver(tg)
    full.computeEigenvalue(1000)
    s = om.HipSolver(tg)
    s.computeEigenvalue(10)
    s.setRestartStatus(True)
    s.computeEigenvalue(1000)
    assert s.getNumIterations() == int(ref["restart10_num_iterations"]) == 9
    assert abs(s.getKeff() - ref["restart10_keff"]) < 1e-7
    assert rel_l2(s.getFluxes(), ref["restart10_flux"]) < 1e-5
    # the continued iteration is the uninterrupted one
    assert abs(s.getKeff() - full.getKeff()) < 1e-12
    np.testing.assert_allclose(s.getFluxes(), full.getFluxes(), rtol=1e-12, atol=0)
    # continuing a converged solve
    full.setRestartStatus(True)
    full.computeEigenvalue(1000)
    assert full.getNumIterations() == int(ref["restart40_num_iterations"])
    assert abs(full.getKeff() - ref["restart40_keff"]) < 1e-7
    assert rel_l2(full.getFluxes(), ref["restart40_flux"]) < 1e-5
    # off again: a fresh solve
    full.setRestartStatus(False)
    full.computeEigenvalue(1000)
    assert abs(full.getNumIterations() - int(ref["num_iterations"])) <= 1


def test_restart_status_needs_an_earlier_solve(tracks):
    s = om.HipSolver(tracks(NAME))
    s.setRestartStatus(True)
    with pytest.raises(om.MOCError) as e:
        s.computeEigenvalue(10)
    assert e.value.code == -2
    s.setRestartStatus(False)
    s.computeEigenvalue(10)
    assert s.getNumIterations() == 10


def test_restart_status_with_the_linear_source(tracks):
    """setRestartStatus on CPULSSolver semantics: the flux moments carry over
    too; 1 + 8 iterations after 10, the eigenvalue of the uninterrupted solve."""
    ref = dict(np.load(GOLDEN / f"{NAME}_ls_restart_ref.npz"))
    tg = tracks(NAME)
    full = om.HipLSSolver(tg)
    full.computeEigenvalue(1000)
    s = om.HipLSSolver(tg)
    s.computeEigenvalue(10)
    s.setRestartStatus(True)
    s.computeEigenvalue(1000)
    assert s.getNumIterations() == int(ref["num_iterations"]) == 9
    assert abs(s.getKeff() - ref["keff"]) < 1e-6
    assert rel_l2(s.getFluxes(), ref["flux"]) < 1e-4
    assert abs(s.getKeff() - full.getKeff()) < 1e-9
    m, mr = s.getFluxMoments(), ref["flux_moments"]
    np.testing.assert_allclose(m.reshape(mr.shape), mr, atol=1e-4 * np.abs(mr).max())
