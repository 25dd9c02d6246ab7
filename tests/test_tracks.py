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


# --- explicit 3D tracks (TrackGenerator3D, EXPLICIT_3D) ---

@pytest.mark.parametrize("name", ["pin_cell_3d_7g", "pin_cell_3d_vac_7g", "lattice_3x3_3d_7g"])
def test_3d_tracks_identical_to_reference(tracks, golden, name):
    g = golden(name)
    tg = tracks(name)
    assert tg.info.is_3d == 1
    assert tg.getNumTracks() == int(g["num_tracks"])
    assert tg.getNumSegments() == int(g["num_segments"])
    assert tg.getNumFSRs() == int(g["num_fsrs"])
    np.testing.assert_array_equal(tg.azim, g["trk_azim"])
    np.testing.assert_array_equal(tg.xy, g["trk_xy"])
    np.testing.assert_array_equal(tg.polar, g["trk_polar"])
    np.testing.assert_allclose(np.c_[tg.start, tg.start_z].ravel(), g["trk_start"], atol=1e-13)
    np.testing.assert_allclose(np.c_[tg.end, tg.end_z].ravel(), g["trk_end"], atol=1e-13)
    # links, including the double reflections at edges, and vacuum ends
    for k in ("next_fwd", "next_bwd", "next_fwd_is_fwd", "next_bwd_is_fwd", "bc_fwd", "bc_bwd"):
        np.testing.assert_array_equal(getattr(tg, k), g["trk_" + k], err_msg=k)
    np.testing.assert_array_equal(np.diff(tg.seg_offset), g["trk_num_segments"])
    np.testing.assert_array_equal(tg.seg_fsr, g["seg_fsr"])          # 3D FSR numbering
    np.testing.assert_allclose(tg.seg_length, g["seg_length"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(tg.theta.ravel(), g["quad_theta"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(tg.total_weight.ravel(), g["quad_total_weight"], rtol=1e-13)
    np.testing.assert_allclose(tg.fsr_volume, g["fsr_volume"], rtol=1e-12)
    np.testing.assert_array_equal(tg.fsr_material, g["fsr_material"])


@pytest.mark.parametrize("name", ["pin_cell_3d_7g", "lattice_3x3_3d_7g"])
def test_3d_volumes_and_layers(tracks, name):
    from openmoc_amd import workloads
    p = workloads.get(name)
    tg = tracks(name)
    assert tg.fsr_volume.sum() == pytest.approx(p.width_x * p.width_y * p.height, rel=1e-12)
    # every 2D region has all its axial layers; each FSR's track-estimated
    # centroid lies inside its layer and the layers of a region add up to
    # the region's 2D area times the height (tracks sample each z-stack
    # uniformly over the height)
    key2 = p.nx * p.ny * 4096
    layer = tg.fsr_key // key2
    assert sorted(set(layer.tolist())) == list(range(p.layers))
    dz = p.height / p.layers
    lo = -p.height / 2 + layer * dz
    assert np.all(tg.fsr_centroid_z > lo) and np.all(tg.fsr_centroid_z < lo + dz)
    by2 = {}
    for f, k in enumerate(tg.fsr_key % key2):
        by2.setdefault(int(k), []).append(tg.fsr_volume[f])
    for v in by2.values():
        assert len(v) == p.layers
        np.testing.assert_allclose(v, np.mean(v), rtol=0.1)


def test_3d_links_are_reciprocal_and_polar_weights_integrate(tracks):
    tg = tracks("pin_cell_3d_7g")
    for t in range(tg.getNumTracks()):
        n, f = tg.next_fwd[t], tg.next_fwd_is_fwd[t]
        back = tg.next_bwd[n] if f else tg.next_fwd[n]
        back_f = tg.next_bwd_is_fwd[n] if f else tg.next_fwd_is_fwd[n]
        assert back == t and back_f == 0
    # Gauss-Legendre: the polar weights of one hemisphere sum to 1/2 each way
    w = tg.polar_weight
    assert w.sum() == pytest.approx(1.0, rel=1e-14)
