"""FSR flux files on the host: the binary layout of Solver::dumpFSRFluxes
(src/Solver.cpp:2066-2112) and the centroid matching of
Solver::loadFSRFluxes (src/Solver.cpp:2135-2318).

Pinned against files the compiled reference wrote (tests/golden/*.fluxes,
its double build) and against what its loadFSRFluxes read back from a file
this library wrote with the FSR order shuffled.  The matching rule is also
checked against brute-force nearest neighbours."""
import os
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

import openmoc_amd as om

GOLDEN = Path(__file__).resolve().parent / "golden"
REF_F64 = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "ref_driver_f64"
NAME = "lattice_3x3_vac_7g"


def parse(path):
    """The layout written out by hand: f64 k, i32 G, i64 N, N x (3 + G) f64."""
    b = Path(path).read_bytes()
    k, G, N = struct.unpack_from("<diq", b, 0)
    rec = np.frombuffer(b, "<f8", offset=20).reshape(N, 3 + G)
    assert len(b) == 20 + N * (3 + G) * 8
    return k, rec[:, :3], rec[:, 3:]


@pytest.fixture(scope="module")
def fixture():
    return dict(np.load(GOLDEN / f"{NAME}_fluxfile_ref.npz"))


def test_write_layout_and_read_back(tmp_path):
    rng = np.random.default_rng(1)
    xyz = rng.uniform(-3, 3, (17, 3))
    flux = rng.uniform(0, 1, (17, 5))
    f = tmp_path / "a.fluxes"
    om.write_flux_file(f, 0.987654321, xyz, flux)
    k, x2, f2 = parse(f)
    assert k == 0.987654321
    np.testing.assert_array_equal(x2, xyz)
    np.testing.assert_array_equal(f2, flux)
    k3, x3, f3 = om.read_flux_file(f)
    assert k3 == k
    np.testing.assert_array_equal(x3, xyz)
    np.testing.assert_array_equal(f3, flux)


def test_reads_files_the_reference_wrote(fixture):
    for tag, kkey, fkey in (("ref", "keff", "flux"), ("it10", "it10_keff", "it10_flux")):
        k, xyz, flux = om.read_flux_file(GOLDEN / f"{NAME}_{tag}.fluxes")
        N = len(xyz)
        assert k == fixture[kkey]
        np.testing.assert_array_equal(flux, fixture[fkey].reshape(N, -1))
        np.testing.assert_array_equal(xyz[:, :2], fixture["fsr_centroid"])
        assert np.all(xyz[:, 2] == 0.0)
        # the hand parser agrees
        k2, xyz2, flux2 = parse(GOLDEN / f"{NAME}_{tag}.fluxes")
        assert k2 == k
        np.testing.assert_array_equal(flux2, flux)


def test_track_generator_centroids_match_the_reference_file(tracks):
    tg = tracks(NAME)
    _, ref_xyz, _ = om.read_flux_file(GOLDEN / f"{NAME}_ref.fluxes")
    own = om.fsr_centroids(tg)
    # track-estimated centroids; segment start points differ from the
    # reference's by up to its 1e-8 cm nudges across surfaces
    np.testing.assert_allclose(own, ref_xyz, rtol=0, atol=1e-7)
    idx, err = om.match_centroids(ref_xyz, own)
    np.testing.assert_array_equal(idx, np.arange(len(own)))
    assert err < 1e-7


def test_reference_loads_a_shuffled_file_this_library_wrote(fixture):
    # the reference's loadFSRFluxes put every written flux back on its FSR
    np.testing.assert_array_equal(fixture["loaded_flux"], fixture["written_flux"])


def test_own_matching_undoes_a_shuffle(tracks, fixture):
    tg = tracks(NAME)
    own = om.fsr_centroids(tg)
    perm = fixture["written_perm"]
    idx, err = om.match_centroids(own[perm], own)
    np.testing.assert_array_equal(perm[idx], np.arange(len(own)))
    assert err == 0.0


def brute_nearest(file_xyz, own, tol):
    cell = lambda p: np.trunc(p / tol).astype(np.int64)  # noqa: E731
    fc, oc = cell(file_xyz), cell(own)
    out, worst = [], 0.0
    for i, p in enumerate(own):
        near = np.all(np.abs(fc - oc[i]) <= 1, axis=1)
        cand = np.flatnonzero(near)
        d = np.linalg.norm(file_xyz[cand] - p, axis=1)
        j = cand[np.argmin(d)]      # argmin takes the lowest index among ties
        out.append(j)
        worst = max(worst, d.min())
    return np.array(out), worst


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_matching_is_nearest_in_the_27_cells(seed):
    rng = np.random.default_rng(seed)
    tol = 0.05
    own = rng.uniform(-1, 1, (400, 3))
    own[:, 2] = np.round(own[:, 2], 1)               # shared z planes, signs on both sides of 0
    file_xyz = np.concatenate([own + rng.normal(0, 0.004, own.shape),
                               own + rng.normal(0, 0.02, own.shape)])
    idx, err = om.match_centroids(file_xyz, own, tol)
    ref, worst = brute_nearest(file_xyz, own, tol)
    np.testing.assert_array_equal(idx, ref)
    assert err == pytest.approx(worst, rel=1e-12)


def test_ties_go_to_the_lower_file_index():
    own = np.array([[0.5, 0.5, 0.0]])
    file_xyz = np.array([[0.5, 0.503, 0.0], [0.503, 0.5, 0.0], [0.5, 0.497, 0.0]])
    idx, _ = om.match_centroids(file_xyz, own, 0.01)
    assert idx[0] == 0


def test_unmatched_centroid_is_an_error():
    with pytest.raises(om.MOCError) as e:
        om.match_centroids(np.array([[5.0, 5.0, 0.0]]), np.array([[0.0, 0.0, 0.0]]), 0.01)
    assert e.value.code == -1


def test_missing_and_malformed_files(tmp_path):
    with pytest.raises(om.MOCError) as e:
        om.read_flux_file(tmp_path / "absent.fluxes")
    assert e.value.code == -6
    good = tmp_path / "g.fluxes"
    om.write_flux_file(good, 1.0, np.zeros((4, 3)), np.ones((4, 2)))
    b = good.read_bytes()
    (tmp_path / "short.fluxes").write_bytes(b[:-8])
    (tmp_path / "long.fluxes").write_bytes(b + b"\0" * 8)
    (tmp_path / "head.fluxes").write_bytes(b[:10])
    for bad in ("short", "long", "head"):
        with pytest.raises(om.MOCError) as e:
            om.read_flux_file(tmp_path / f"{bad}.fluxes")
        assert e.value.code == -6, bad


@pytest.mark.skipif(not REF_F64.exists() or not Path("/root/reference").exists(),
                    reason="needs the compiled reference (make -C oracle ref)")
def test_reference_reads_this_writer_live(tmp_path, tracks):
    """Writer -> the reference's loadFSRFluxes, run now (the compiled
    reference, double build), FSR order reversed."""
    from oracle import mocdump
    tg = tracks(NAME)
    own = om.fsr_centroids(tg)
    N, G = len(own), tg.getNumEnergyGroups()
    flux = np.arange(N * G, dtype=np.float64).reshape(N, G) + 1.0
    rev = np.arange(N)[::-1]
    f = tmp_path / "w.fluxes"
    om.write_flux_file(f, 1.0, own[rev], flux[rev])
    out = tmp_path / "d.bin"
    env = dict(os.environ, OMP_NUM_THREADS="1")
    subprocess.run([str(REF_F64), str(GOLDEN / f"{NAME}.txt"), str(out), "--threads", "1",
                    "--max-iters", "2", "--load-fluxes", str(f)],
                   check=True, capture_output=True, env=env, timeout=300)
    d = mocdump.read(str(out))
    np.testing.assert_array_equal(d["loaded_flux"].reshape(N, G), flux)
