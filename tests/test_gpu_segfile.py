"""Solves on tracks read back from a segment file (readSegmentsFromFile,
src/TrackGenerator.cpp:1518-1638) give bitwise the results of the generated
tracks: the file restores every array the solver is given."""
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
