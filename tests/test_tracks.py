"""The CPU producer (openmoc_amd/csrc/tracks.cpp) against the compiled
reference TrackGenerator's output (golden fixtures) and invariants."""
import math

import numpy as np
import pytest


def test_pin_cell_7g_tracks_identical_to_reference(tracks, golden):
    g = golden("pin_cell_7g")
    tg = tracks("pin_cell_7g")
    assert tg.getNumTracks() == int(g["num_tracks"]) == 532
    assert tg.getNumSegments() == int(g["num_segments"])
    assert tg.getNumFSRs() == int(g["num_fsrs"]) == 32
    np.testing.assert_array_equal(np.diff(tg.seg_offset), g["trk_num_segments"])
    np.testing.assert_array_equal(tg.seg_fsr, g["seg_fsr"])          # FSR numbering
    np.testing.assert_allclose(tg.seg_length, g["seg_length"], rtol=0, atol=1e-12)
    for k in ("next_fwd", "next_bwd", "next_fwd_is_fwd", "next_bwd_is_fwd", "bc_fwd", "bc_bwd"):
        np.testing.assert_array_equal(getattr(tg, k), g["trk_" + k], err_msg=k)
    np.testing.assert_array_equal(tg.azim, g["trk_azim"])
    np.testing.assert_allclose(tg.start.ravel(), g["trk_start"], atol=1e-14)
    np.testing.assert_allclose(tg.total_weight.ravel(), g["quad_total_weight"], rtol=1e-15)
    np.testing.assert_allclose(tg.phi, g["quad_phi"], rtol=0, atol=0)
    np.testing.assert_allclose(tg.fsr_volume, g["fsr_volume"], rtol=1e-13)
    np.testing.assert_array_equal(tg.fsr_material, g["fsr_material"])


def test_lattice_17x17_tracks_match_reference(tracks, golden):
    g = golden("lattice_17x17_7g")
    tg = tracks("lattice_17x17_7g")
    assert tg.getNumTracks() == int(g["num_tracks"]) == 29144
    assert tg.getNumSegments() == int(g["num_segments"])
    assert tg.getNumFSRs() == int(g["num_fsrs"]) == 9248
    ns = np.diff(tg.seg_offset)
    np.testing.assert_array_equal(ns, g["trk_num_segments"])
    np.testing.assert_allclose(np.add.reduceat(tg.seg_length, tg.seg_offset[:-1]),
                               g["trk_length_sum"], rtol=1e-12)
    np.testing.assert_array_equal(np.add.reduceat(tg.seg_fsr.astype(np.int64), tg.seg_offset[:-1]),
                                  g["trk_fsr_sum"])
    np.testing.assert_array_equal(tg.next_fwd, g["trk_next_fwd"])
    np.testing.assert_array_equal(tg.next_bwd, g["trk_next_bwd"])
    np.testing.assert_allclose(tg.fsr_volume, g["fsr_volume"], rtol=1e-11)
    np.testing.assert_array_equal(tg.fsr_material, g["fsr_material"])


@pytest.mark.parametrize("name", ["c5g7_pin_cell", "pin_cell_7g", "hom_inf_medium"])
def test_quadrature_weight_invariant(tracks, name):
    # sum_{a,p} w(a,p) / (sin(theta_p) * spacing_a) = 4 pi
    # (tests/test_quadrature/test_quadrature.py:44-52 with unit spacing)
    tg = tracks(name)
    s = (tg.total_weight / tg.sin_theta[None, :] / tg.azim_spacing[:, None]).sum()
    assert abs(s - 4 * math.pi) < 1e-10


@pytest.mark.parametrize("name", ["c5g7_pin_cell", "pin_cell_7g", "lattice_17x17_7g"])
def test_volumes_sum_to_domain_area(tracks, name):
    tg = tracks(name)
    p = tg.problem
    assert abs(tg.fsr_volume.sum() / (p.width_x * p.width_y) - 1) < 1e-12


@pytest.mark.parametrize("name", ["pin_cell_7g", "lattice_17x17_7g", "minicore_70g"])
def test_links_form_a_permutation(tracks, name):
    tg = tracks(name)
    T = tg.getNumTracks()
    slots = []
    for d, (nxt, isf, bc) in enumerate([(tg.next_fwd, tg.next_fwd_is_fwd, tg.bc_fwd),
                                        (tg.next_bwd, tg.next_bwd_is_fwd, tg.bc_bwd)]):
        m = bc != 0
        slots.append(nxt[m] * 2 + (1 - isf[m]))
    s = np.concatenate(slots)
    assert len(np.unique(s)) == len(s)
    # reflective cycles close: following links from any track returns to it
    if tg.problem.bc == ("R", "R", "R", "R"):
        assert len(s) == 2 * T


def test_fsr_centroids_inside_domain(tracks):
    tg = tracks("pin_cell_7g")
    assert np.all(np.abs(tg.fsr_centroid) < 0.63)


def test_minicore_counts(tracks):
    tg = tracks("minicore_70g")
    assert tg.getNumEnergyGroups() == 70
    # 4 fuel assemblies x 289 pins x 32 FSRs + 5 reflector assemblies x 289 blocks
    assert tg.getNumFSRs() == 4 * 289 * 32 + 5 * 289
    assert (tg.bc_fwd == 0).any() and (tg.bc_fwd == 1).any()
