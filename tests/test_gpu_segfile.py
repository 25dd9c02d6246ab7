"""Solves on tracks read from segment files (readSegmentsFromFile,
src/TrackGenerator.cpp:1518-1638): files the reference itself wrote give the
reference's results, and files this library wrote give bitwise the results
of the generated tracks."""
import numpy as np
import pytest

import openmoc_amd as om

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["lattice_3x3_vac_7g", "pin_cell_3d_7g"])
def test_solve_on_tracks_from_a_segment_file(tracks, tmp_path, name):
    tg = tracks(name)
    f = tmp_path / "t.tracks"
    tg.dumpSegmentsToFile(f)
    t2 = om.TrackGenerator(om.get_workload(name))
    assert t2.readSegmentsFromFile(f)
    a, b = om.HipSolver(tg), om.HipSolver(t2)
    a.computeEigenvalue(1000)
    b.computeEigenvalue(1000)
    assert b.getNumIterations() == a.getNumIterations()
    assert b.getKeff() == a.getKeff()
    np.testing.assert_array_equal(b.getFluxes(), a.getFluxes())


@pytest.mark.parametrize("name", ["c5g7_pin_cell", "pin_cell_7g"])
def test_solve_on_segment_files_the_reference_wrote(golden, name):
    """Tracks and segments from the reference's own dumpSegmentsToFile output
    (tests/golden/*_ref_segments.data): keff within 1 pcm and FSR flux
    relative L2 <= 1e-4 of the reference's converged solve."""
    from conftest import GOLDEN
    rd = om.TrackGenerator(om.get_workload(name))
    assert rd.readSegmentsFromFile(GOLDEN / f"{name}_ref_segments.data")
    g = golden(name)
    s = om.HipSolver(rd)
    s.computeEigenvalue(1000)
    ref = np.asarray(g["flux"], np.float64).ravel()
    rl2 = float(np.linalg.norm(s.getFluxes() - ref) / np.linalg.norm(ref))
    print(f"{name} on the reference's segment file: keff {s.getKeff():.8f} "
          f"(reference {float(g['keff']):.8f}), flux rel L2 {rl2:.2e}")
    assert abs(s.getKeff() - float(g["keff"])) < 1e-5
    assert rl2 < 1e-4
