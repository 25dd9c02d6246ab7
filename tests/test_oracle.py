"""Pins the oracle (oracle/moc_oracle.cpp) against golden vectors, closed
forms and the compiled reference's outputs.  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

from conftest import GOLDEN

# Reference unit-test golden values of expF1_fractional
# (tests/unit_tests/test_exponentials.py:11,75-78).
TAU = [1e-8, 1e-7, 1e-6, 1e-5, 1e-4, 1e-3, 1e-2, 1e-1, 1, 10, 100]
F1_GOLDEN = [0.9999999950003422, 0.9999999500034228, 0.9999995000343784,
             0.9999950003587617, 0.9999500050851808, 0.9995002005718668,
             0.9950169413610763, 0.9516272442309586, 0.6321211831479621,
             0.09999547551656497, 0.010000005017389789]


def test_exp_f1_golden_values():
    for x, ref in zip(TAU, F1_GOLDEN):
        assert abs(orc.exp_f1(x) - ref) < 1e-12


def test_exp_f1_against_closed_form():
    # the fit's stated accuracy: 2e-6 over [0, 1e5] (src/exponentials.h:150-152)
    for x in np.concatenate([np.logspace(-8, 2, 400), [0.3, 3.7, 42.0]]):
        exact = -math.expm1(-x) / x
        assert abs(orc.exp_f1(x) - exact) < 2e-6, x


def test_pairwise_sum_matches_exact_integer_sums():
    rng = np.random.default_rng(0)
    for n in [0, 1, 2, 3, 4, 5, 7, 8, 63, 64, 65, 1000, 12345]:
        v = rng.integers(-1000, 1000, n).astype(np.float64)
        assert orc.pairwise_sum(v) == float(v.sum())


def test_single_segment_attenuation(tracks):
    """One sweep on a 1-cell homogeneous problem: every track's outgoing flux
    obeys psi_out = psi_in e^-tau + q (1 - e^-tau) per segment chain (with the
    fit's error), checked on the CPU restatement."""
    tg = tracks("hom_inf_medium")
    op = orc.OracleProblem.from_track_generator(tg)
    G, P = op.G, op.P
    Q = np.full(op.N * G, 0.3)
    psi = np.zeros(op.T * 2 * P * G, np.float32)
    psi[:] = 1.0
    tally, psi_out = op.transport_sweep(Q, psi, threads=1)
    sig = tg.sigma_t[0]
    st = tg.sin_theta[:P]
    for t in [0, 5, 17]:
        L = tg.seg_length[tg.seg_offset[t]:tg.seg_offset[t + 1]].sum()
        nxt, nd = tg.next_fwd[t], 1 - tg.next_fwd_is_fwd[t]
        got = psi_out[(nxt * 2 + nd) * P * G:(nxt * 2 + nd + 1) * P * G].reshape(P, G)
        for p in range(P):
            tau = sig * L / st[p]
            qs = 0.3 / sig
            exp = 1.0 * np.exp(-tau) + qs * (1 - np.exp(-tau))
            np.testing.assert_allclose(got[p], exp, rtol=2e-5)


def test_infinite_medium_k_matches_closed_form(tracks):
    # k_inf = (nsf1 + nsf2 * s12 / sa2) / (sa1 + s12) for the 2-group medium
    # of tests/input_set.py:32-92 (reference result 1.72307).
    tg = tracks("hom_inf_medium")
    sol = orc.OracleProblem.from_track_generator(tg).eigenvalue(max_iters=500, tol=1e-10, threads=2)
    s11, s12, s22 = 0.1, 0.117, 1.42
    sa1 = 0.2208 - s11 - s12
    sa2 = 1.604 - s22
    k_inf = (0.0015 + 0.325 * s12 / sa2) / (sa1 + s12)
    assert abs(sol.keff - k_inf) < 2e-5
    # flat flux with ratio phi1/phi2 = sa2 / s12
    ratio = sol.flux[:, 0] / sol.flux[:, 1]
    np.testing.assert_allclose(ratio, sa2 / s12, rtol=1e-4)


def test_reference_fixture_hom_inf_medium(tracks, golden):
    # tests/test_forward_hom_inf_medium/results_true.dat: 108 iterations, 1.72307
    g = golden("hom_inf_medium")
    tg = tracks("hom_inf_medium")
    sol = orc.OracleProblem.from_track_generator(tg).eigenvalue(max_iters=500, tol=1e-5, threads=1)
    assert sol.num_iterations == 108 == int(g["num_iterations"])
    assert f"{sol.keff:12.5E}" == " 1.72307E+00"
    np.testing.assert_allclose(sol.flux.ravel(), g["flux"], rtol=1e-9)


def test_reference_fixture_c5g7_pin_cell(tracks, golden):
    # tests/test_forward_pin_cell/results_true.dat (261 iterations, 1.04666,
    # 14 fluxes printed with %12.6E)
    truth = [6.428760E-01, 1.099273E+00, 5.710249E-01, 2.507108E-01, 1.962766E-01,
             4.993489E-01, 1.287533E+00, 1.104038E+00, 1.411933E+00, 5.559559E-01,
             2.268887E-01, 1.900581E-01, 4.444841E-01, 9.440460E-01]
    tg = tracks("c5g7_pin_cell")
    sol = orc.OracleProblem.from_track_generator(tg).eigenvalue(max_iters=500, tol=1e-5, threads=1)
    assert sol.num_iterations == 261
    assert f"{sol.keff:12.5E}" == " 1.04666E+00"
    assert [f"{v:12.6E}" for v in sol.flux.ravel()] == [f"{v:12.6E}" for v in truth]
    np.testing.assert_allclose(sol.flux.ravel(), golden("c5g7_pin_cell")["flux"], rtol=1e-9)


def test_reference_fixture_pin_cell_7g(tracks, golden):
    g = golden("pin_cell_7g")
    tg = tracks("pin_cell_7g")
    sol = orc.OracleProblem.from_track_generator(tg).eigenvalue(max_iters=1000, tol=1e-5, threads=2)
    assert sol.num_iterations == int(g["num_iterations"])
    assert abs(sol.keff - float(g["keff"])) < 1e-11
    np.testing.assert_allclose(sol.flux.ravel(), g["flux"], rtol=1e-9)


def test_oracle_on_reference_tracks(golden):
    """The oracle run on the reference's own tracks (the golden fixture)
    reproduces the reference CPUSolver's k and fluxes."""
    from openmoc_amd import workloads
    g = golden("pin_cell_7g")
    p = workloads.get("pin_cell_7g")
    d = {"trk_azim": g["trk_azim"], "trk_num_segments": g["trk_num_segments"],
         "seg_length": g["seg_length"], "seg_fsr": g["seg_fsr"],
         "trk_next_fwd": g["trk_next_fwd"], "trk_next_bwd": g["trk_next_bwd"],
         "trk_next_fwd_is_fwd": g["trk_next_fwd_is_fwd"],
         "trk_next_bwd_is_fwd": g["trk_next_bwd_is_fwd"], "trk_bc_fwd": g["trk_bc_fwd"],
         "trk_bc_bwd": g["trk_bc_bwd"], "num_azim": 32, "num_polar": 6,
         "quad_total_weight": g["quad_total_weight"],
         "quad_sin_theta": np.tile([math.sin(math.asin(v)) for v in
                                    (0.166648, 0.537707, 0.932954, 0.932954, 0.537707, 0.166648)],
                                   16),
         "fsr_material": g["fsr_material"], "fsr_volume": g["fsr_volume"]}
    sol = orc.from_ref_dump(d, p).eigenvalue(max_iters=1000, tol=1e-5, threads=2)
    assert sol.num_iterations == int(g["num_iterations"])
    assert abs(sol.keff - float(g["keff"])) < 1e-12
    np.testing.assert_allclose(sol.flux.ravel(), g["flux"], rtol=1e-11)


def test_particle_balance_reflective(tracks):
    """Reflective pin cell, fixed source only, no fission: absorption equals
    the total source (no leakage) once converged (computeSource with k=1 and a
    non-fissile copy)."""
    tg = tracks("c5g7_pin_cell")
    N, G = tg.getNumFSRs(), tg.getNumEnergyGroups()
    fixed = np.zeros((N, G))
    fixed[:, 0] = 1.0
    op = orc.OracleProblem(
        num_azim=tg.info.num_azim, num_polar=tg.info.num_polar, seg_offset=tg.seg_offset,
        seg_length=tg.seg_length, seg_fsr=tg.seg_fsr, azim=tg.azim, next_fwd=tg.next_fwd,
        next_bwd=tg.next_bwd, next_fwd_is_fwd=tg.next_fwd_is_fwd,
        next_bwd_is_fwd=tg.next_bwd_is_fwd, bc_fwd=tg.bc_fwd, bc_bwd=tg.bc_bwd,
        total_weight=tg.total_weight, sin_theta=tg.sin_theta, sigma_t=tg.sigma_t,
        sigma_s=tg.sigma_s, nu_sigma_f=np.zeros_like(tg.nu_sigma_f), chi=tg.chi,
        fsr_material=tg.fsr_material, fsr_volume=tg.fsr_volume, fixed_source=fixed)
    sol = op.source(1.0, max_iters=2000, tol=1e-9, res_type=orc.SCALAR_FLUX, threads=2)
    sa = tg.sigma_t - tg.sigma_s.reshape(-1, G, G).sum(axis=1)
    absorption = sum(tg.fsr_volume[r] * (sa[tg.fsr_material[r]] * sol.flux[r]).sum()
                     for r in range(N))
    source = (fixed * tg.fsr_volume[:, None]).sum()
    assert abs(absorption / source - 1) < 2e-3


# --- neutron-balance keff, transport stabilisation, Krylov operator sweeps ---

def test_tracks_of_new_fixtures_match_reference(tracks, golden):
    for name in ("lattice_3x3_vac_7g", "pin_cell_7g_tc"):
        g = golden(name)
        tg = tracks(name)
        assert tg.getNumTracks() == int(g["num_tracks"])
        assert tg.getNumSegments() == int(g["num_segments"])
        assert tg.getNumFSRs() == int(g["num_fsrs"])
    g = golden("lattice_3x3_vac_7g")
    tg = tracks("lattice_3x3_vac_7g")
    np.testing.assert_array_equal(np.diff(tg.seg_offset), g["trk_num_segments"])


def test_oracle_balance_keff_matches_reference(tracks, golden):
    g = golden("lattice_3x3_vac_7g")
    op = orc.OracleProblem.from_track_generator(tracks("lattice_3x3_vac_7g"))
    sol = op.eigenvalue(max_iters=1000, tol=1e-5, balance=True)
    assert abs(sol.keff - float(g["balance_keff"])) < 1e-8
    assert sol.num_iterations == int(g["balance_num_iterations"])
    assert np.linalg.norm(sol.flux.ravel() - g["balance_flux"]) / np.linalg.norm(g["balance_flux"]) < 1e-7
    # the vacuum leakage is what separates it from the fission-rate keff
    assert abs(float(g["balance_keff"]) - float(g["keff"])) > 0.1


def test_oracle_balance_keff_equals_fission_rate_keff_without_leakage(tracks):
    # reflective infinite medium: no leakage, so at convergence
    # fission / absorption equals the fission-rate eigenvalue k_inf
    # (FP64 angular fluxes, so that only the iteration error remains)
    op = orc.OracleProblem.from_track_generator(tracks("hom_inf_medium"))
    kb = op.eigenvalue(max_iters=5000, tol=1e-12, balance=True, psi_fp64=True).keff
    kf = op.eigenvalue(max_iters=5000, tol=1e-12, psi_fp64=True).keff
    s11, s12, s22 = 0.1, 0.117, 1.42
    k_inf = (0.0015 + 0.325 * s12 / (1.604 - s22)) / (0.2208 - s11 - s12 + s12)
    assert abs(kb - k_inf) < 1e-9
    assert abs(kf - k_inf) < 1e-8


@pytest.mark.parametrize("tag,typ", [("diagonal", 0), ("yamamoto", 1), ("global", 2)])
def test_oracle_stabilization_matches_reference(tracks, golden, tag, typ):
    g = golden("pin_cell_7g_tc")
    op = orc.OracleProblem.from_track_generator(tracks("pin_cell_7g_tc"))
    sol = op.eigenvalue(max_iters=1000, tol=1e-5,
                        stabilization=(float(g[f"{tag}_factor"]), typ))
    assert abs(sol.keff - float(g[f"{tag}_keff"])) < 1e-9
    assert sol.num_iterations == int(g[f"{tag}_num_iterations"])
    ref = g[f"{tag}_flux"]
    assert np.linalg.norm(sol.flux.ravel() - ref) / np.linalg.norm(ref) < 1e-9


def test_oracle_stabilization_changes_the_iteration(golden):
    # damping changes the path (iteration counts differ from the plain solve),
    # DIAGONAL only acts through the negative in-group scattering
    g = golden("pin_cell_7g_tc")
    assert int(g["yamamoto_num_iterations"]) != int(g["num_iterations"])
    assert int(g["global_num_iterations"]) != int(g["num_iterations"])
    assert abs(float(g["diagonal_keff"]) - float(g["keff"])) < 1e-5


def test_oracle_krylov_sweeps_match_reference(tracks, golden):
    g = golden("lattice_3x3_vac_7g")
    op = orc.OracleProblem.from_track_generator(tracks("lattice_3x3_vac_7g"))
    psi = np.zeros(2 * op.T * op.P * op.G, np.float32)
    phi0 = g["krylov_phi0"]
    out = {}
    out["krylov_fission_1"], psi = op.krylov_sweep(1, phi0, psi)
    out["krylov_scatter"], psi = op.krylov_sweep(2, phi0, psi)
    out["krylov_fission_2"], psi = op.krylov_sweep(1, phi0, psi)
    for k, v in out.items():
        assert np.linalg.norm(v - g[k]) / np.linalg.norm(g[k]) < 1e-12, k


def test_oracle_krylov_sweeps_are_linear(tracks, golden):
    # all-vacuum problem: start fluxes stay zero, so the sweep is linear in
    # the source and fission-only + scatter-only equals one sweep on the
    # total (k = 1, no fixed source) source
    g = golden("lattice_3x3_vac_7g")
    op = orc.OracleProblem.from_track_generator(tracks("lattice_3x3_vac_7g"))
    phi0 = g["krylov_phi0"]
    zero = np.zeros(2 * op.T * op.P * op.G, np.float32)
    f, _ = op.krylov_sweep(1, phi0, zero)
    s, _ = op.krylov_sweep(2, phi0, zero)
    Q = op.fsr_sources(phi0, 1.0)
    tally, _ = op.transport_sweep(Q, zero)
    total = op.add_source(tally, Q)
    np.testing.assert_allclose(f + s, total, rtol=1e-5)


# --- linear source (CPULSSolver) ---

@pytest.mark.parametrize("name", ["pin_cell_7g", "lattice_3x3_vac_7g"])
def test_oracle_linear_source_matches_reference(tracks, golden, name):
    g = golden(f"{name}_ls")
    tg = tracks(name)
    op = orc.OracleProblem.from_track_generator(tg)
    sol, mom, cen = op.eigenvalue_ls(max_iters=1000, tol=1e-5)
    assert abs(sol.keff - float(g["keff"])) < 1e-8
    assert sol.num_iterations == int(g["num_iterations"])
    assert np.linalg.norm(sol.flux.ravel() - g["flux"]) / np.linalg.norm(g["flux"]) < 1e-7
    m = g["flux_moments"]
    # the reference's segment starts carry its 1e-8 cm surface nudges
    assert np.linalg.norm(mom - m) / np.linalg.norm(m) < 1e-4
    np.testing.assert_allclose(cen, g["fsr_centroid"], atol=1e-7)
    # the linear source changes the answer (it is not the flat source)
    flat = op.eigenvalue(max_iters=1000, tol=1e-5)
    assert abs(flat.keff - sol.keff) > 1e-6


def test_oracle_linear_source_reduces_to_flat_in_infinite_medium(tracks):
    # homogeneous reflective medium: the flux has no gradient, so the
    # moments vanish and the linear-source k equals the flat-source k
    op = orc.OracleProblem.from_track_generator(tracks("hom_inf_medium"))
    sol, mom, _ = op.eigenvalue_ls(max_iters=1000, tol=1e-8, psi_fp64=True)
    flat = op.eigenvalue(max_iters=1000, tol=1e-8, psi_fp64=True)
    s11, s12, s22 = 0.1, 0.117, 1.42
    k_inf = (0.0015 + 0.325 * s12 / (1.604 - s22)) / (0.2208 - s11 - s12 + s12)
    # both stop once |dk| < 1 pcm; within that they agree with k_inf
    assert abs(sol.keff - flat.keff) < 1e-6
    assert abs(sol.keff - k_inf) < 1e-5
    assert np.abs(mom).max() < 1e-6 * np.abs(sol.flux).max()


def test_oracle_centroids_are_volume_weighted_means(tracks):
    # centroid of the whole pin-cell geometry (volume-weighted mean of the
    # FSR centroids) is the cell centre, and every centroid lies in the cell
    tg = tracks("pin_cell_7g")
    op = orc.OracleProblem.from_track_generator(tg)
    _, _, cen = op.eigenvalue_ls(max_iters=2, tol=1e-5)
    v = tg.fsr_volume
    np.testing.assert_allclose((cen * v[:, None]).sum(0) / v.sum(), [0.0, 0.0], atol=1e-7)
    assert np.all(np.abs(cen) < 0.63)


def test_oracle_70_groups_matches_reference(tracks, golden):
    g = golden("pin_cell_70g")
    op = orc.OracleProblem.from_track_generator(tracks("pin_cell_70g"))
    sol = op.eigenvalue(max_iters=1000, tol=1e-5)
    assert abs(sol.keff - float(g["keff"])) < 1e-8
    assert sol.num_iterations == int(g["num_iterations"])
    assert np.linalg.norm(sol.flux.ravel() - g["flux"]) / np.linalg.norm(g["flux"]) < 1e-7


@pytest.mark.parametrize("name", ["pin_cell_3d_7g", "pin_cell_3d_vac_7g", "lattice_3x3_3d_7g"])
def test_oracle_3d_matches_reference(tracks, golden, name):
    # the 3D sweep (3D segment lengths, F1(tau), w(a, p) per track at the
    # FSR flush) on this package's tracks, which equal the reference's
    g = golden(name)
    sol = orc.OracleProblem.from_track_generator(tracks(name)).eigenvalue(max_iters=1000, tol=1e-5)
    assert abs(sol.keff - float(g["keff"])) < 1e-12
    assert sol.num_iterations == int(g["num_iterations"])
    f, fr = sol.flux.ravel(), g["flux"]
    assert np.linalg.norm(f - fr) / np.linalg.norm(fr) < 1e-12


def test_oracle_3d_axially_uniform_problem_has_uniform_layers(tracks):
    # reflective top and bottom with axially uniform materials: every axial
    # layer of a 2D region carries the same flux (up to the axial laydown)
    tg = tracks("pin_cell_3d_7g")
    sol = orc.OracleProblem.from_track_generator(tg).eigenvalue(max_iters=1000, tol=1e-7)
    key2 = 4096
    by2 = {}
    for f, k in enumerate(tg.fsr_key % key2):
        by2.setdefault(int(k), []).append(sol.flux[f])
    for v in by2.values():
        v = np.array(v)
        assert np.abs(v / v.mean(0) - 1).max() < 2e-3


# ---------------------------------------------------------------------------
# computeFlux / computeSource (src/Solver.cpp:1352-1516) against the compiled
# reference and the reference's own regression files
# ---------------------------------------------------------------------------
WATER_BOX_FLUX_SOURCES = ((1, 1.0), (2, 0.5), (3, 0.25))   # test_compute_flux.py:86-88
WATER_BOX_SOURCE_SOURCE = ((1, 1.0),)                       # test_compute_source.py:84
SOURCE_CELL = (1, 0)   # the water box's source universe, its only cell


def _results_string(num_iterations, flux):
    """The regression-file text of tests/testing_harness.py:158-207 for a
    fixed-source solve: iterations, then every flux as {:12.6E}."""
    return (f"# Iterations: {num_iterations}\nfluxes:\n"
            + "\n".join("{0:12.6E}".format(f) for f in np.ravel(flux)) + "\n")


def _cell_sources(tg, cell, sources):
    N, G = tg.getNumFSRs(), tg.getNumEnergyGroups()
    fixed = np.zeros((N, G))
    sel = tg.getFSRCells() == cell[0] * 64 + cell[1]
    for g, v in sources:
        fixed[sel, g - 1] = v
    return fixed


def _maxrel(a, b):
    b = np.ravel(b)
    return float(np.max(np.abs(np.ravel(a) - b)) / np.max(np.abs(b)))


def test_water_box_source_cell_is_the_tests_lattice_cell(tracks, golden):
    """The source universe sits at universes[0][3][3] of the reference test
    (row 3 from the top, column 3): the one FSR of that cell is where the
    reference's own FSR point lies."""
    tg = tracks("water_box_source")
    cells = tg.getFSRCells()
    assert (cells == 64).sum() == 1
    r = int(np.nonzero(cells == 64)[0][0])
    x, y = golden("water_box_source")["fsr_point"].reshape(tg.getNumFSRs(), -1)[r, :2]
    assert -0.5 < x < -0.25 and 0.25 < y < 0.5


def test_oracle_reproduces_reference_compute_flux_results(tracks):
    """tests/test_compute_flux/results_true.dat, byte for byte."""
    tg = tracks("water_box_source")
    op = orc.OracleProblem.from_track_generator(
        tg, fixed_source=_cell_sources(tg, SOURCE_CELL, WATER_BOX_FLUX_SOURCES))
    sol, _ = op.solve_state(1, max_iters=500, tol=1e-5)
    expect = (GOLDEN / "test_compute_flux_results_true.dat").read_text()
    assert _results_string(sol.num_iterations, sol.flux) == expect


def test_oracle_reproduces_reference_compute_source_results(tracks):
    """tests/test_compute_source/results_true.dat, byte for byte
    (computeSource(500, k=1, TOTAL_SOURCE))."""
    tg = tracks("water_box_source")
    op = orc.OracleProblem.from_track_generator(
        tg, fixed_source=_cell_sources(tg, SOURCE_CELL, WATER_BOX_SOURCE_SOURCE))
    sol, _ = op.solve_state(2, k_eff=1.0, max_iters=500, tol=1e-5, res_type=orc.TOTAL_SOURCE)
    expect = (GOLDEN / "test_compute_source_results_true.dat").read_text()
    assert _results_string(sol.num_iterations, sol.flux) == expect


def test_oracle_fixed_source_modes_match_reference(tracks, golden):
    """Flux from zero, flux continued (only_fixed_source=False), and source
    solves: same iteration counts and fluxes to 1e-12 as the compiled
    reference, on the water box (cell source), the C5G7 pin cell (per-FSR
    sources) and the leaking 3x3 lattice (a material source overlapping
    per-FSR sources)."""
    tg = tracks("water_box_source")
    g = golden("water_box_source")
    op = orc.OracleProblem.from_track_generator(
        tg, fixed_source=_cell_sources(tg, SOURCE_CELL, WATER_BOX_FLUX_SOURCES))
    s1, psi = op.solve_state(1, max_iters=500, tol=1e-5)
    s2, _ = op.solve_state(1, phi0=s1.flux, psi0=psi, max_iters=500, tol=1e-5)
    assert (s1.num_iterations, s2.num_iterations) == (g["flux_num_iterations"],
                                                      g["flux0_num_iterations"])
    assert _maxrel(s1.flux, g["flux_flux"]) < 1e-12
    assert _maxrel(s2.flux, g["flux0_flux"]) < 1e-12

    tg = tracks("c5g7_pin_cell")
    g = golden("c5g7_pin_cell_fixed_source")
    op = orc.OracleProblem.from_track_generator(tg, fixed_source=g["fixed"])
    s1, psi = op.solve_state(1, max_iters=1000, tol=1e-7)
    s2, _ = op.solve_state(1, phi0=s1.flux, psi0=psi, max_iters=1000, tol=1e-7)
    s3, _ = op.solve_state(2, k_eff=3.0, max_iters=1000, tol=1e-7, res_type=orc.TOTAL_SOURCE)
    for s, tag in ((s1, "flux"), (s2, "flux0"), (s3, "source")):
        assert s.num_iterations == g[f"{tag}_num_iterations"], tag
        assert _maxrel(s.flux, g[f"{tag}_flux"]) < 1e-12, tag

    tg = tracks("lattice_3x3_vac_7g")
    g = golden("lattice_3x3_vac_7g_fixed_source")
    fixed = fixed_sources_as_the_reference_applies_them(tg, g["sources"])
    op = orc.OracleProblem.from_track_generator(tg, fixed_source=fixed)
    s1, psi = op.solve_state(2, k_eff=0.8, max_iters=1000, tol=1e-7, res_type=orc.TOTAL_SOURCE)
    s2, _ = op.solve_state(1, phi0=s1.flux, psi0=psi, max_iters=1000, tol=1e-7)
    assert (s1.num_iterations, s2.num_iterations) == (g["source_num_iterations"],
                                                      g["flux0_num_iterations"])
    assert _maxrel(s1.flux, g["source_flux"]) < 1e-12
    assert _maxrel(s2.flux, g["flux0_flux"]) < 1e-12


def fixed_sources_as_the_reference_applies_them(tg, sources):
    """[N, G] fixed sources from (kind, id, group, value) rows, kind 0 per FSR
    and 1 per material, in the order of Solver::initializeFixedSources
    (src/Solver.cpp:1089-1137): material entries overwrite the per-FSR ones."""
    N, G = tg.getNumFSRs(), tg.getNumEnergyGroups()
    per_fsr = {}
    for kind, a, g, v in sources:
        if int(kind) == 0:
            per_fsr[(int(a), int(g))] = v
    for kind, a, g, v in sources:
        if int(kind) == 1:
            for r in np.nonzero(tg.fsr_material == int(a))[0]:
                per_fsr[(int(r), int(g))] = v
    fixed = np.zeros((N, G))
    for (r, g), v in per_fsr.items():
        fixed[r, g - 1] = v
    return fixed


def test_material_source_wins_over_overlapping_fsr_sources(golden):
    """The fixture sets a per-FSR source in a moderator FSR before and one
    after the moderator's material source: the reference's fluxes are those
    of the material value in both (checked through the oracle above); here,
    that the fixture does overlap."""
    g = golden("lattice_3x3_vac_7g_fixed_source")
    rows = g["sources"]
    assert rows[0][0] == 0 and rows[1][0] == 1 and rows[2][0] == 0
    assert rows[0][2] == rows[1][2] == rows[2][2] == 1


def test_oracle_flux_continues_an_eigenvalue_solve(tracks, golden):
    """computeEigenvalue, then computeFlux(only_fixed_source=False) without
    fixed sources: k = 1 sources of the eigenvalue flux, started from its
    scalar and angular fluxes."""
    tg = tracks("pin_cell_7g")
    g = golden("pin_cell_7g_eig_flux0")
    op = orc.OracleProblem.from_track_generator(tg)
    s1, psi = op.solve_state(0, max_iters=1000, tol=1e-5)
    s2, _ = op.solve_state(1, phi0=s1.flux, psi0=psi, max_iters=1000, tol=1e-5)
    assert s1.num_iterations == g["eig_num_iterations"]
    assert abs(s1.keff - g["eig_keff"]) < 1e-12
    assert s2.num_iterations == g["flux0_num_iterations"]
    assert _maxrel(s2.flux, g["flux0_flux"]) < 1e-12
