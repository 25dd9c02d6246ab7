"""The benchmark configurations at their own sizes on the GPU, against the
FP64 oracle (oracle/moc_oracle.cpp, pinned to the compiled reference in
test_oracle.py) on the same tracks:

  * 3x3 mini-core, 70 groups, linear source with the matrix-core scattering
    source (N = 38,437 FSRs, 4.7 M segments): the first two source iterations
    against the oracle's CPULSSolver iterations, and a converged solve on the
    matrix cores against the same solve with the FP64 contraction;
  * 3D axially extruded 17x17 lattice (1.09 M tracks, 66.6 M segments, 92,480
    FSRs): one z-stack sweep against the oracle's 3D sweep, and keff over the
    first ten iterations;
  * the serving mix with a 17x17 lattice and a 70-group mini-core among pin
    cells: every served result against its plain solve.

Bars: FP32 sweep against FP64 (the angular fluxes are FP32 in both, as in
the reference): flux relative L2 <= 2e-5 after a sweep or two iterations;
keff within 1 pcm (0.1 pcm between the matrix-core and FP64 contractions)."""
import os
import time

import numpy as np
import pytest

import openmoc_amd as om
from openmoc_amd import serve, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

THREADS = min(16, len(os.sched_getaffinity(0)))


def rel_l2(a, b):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def log(msg):
    print(msg, flush=True)


def test_minicore_70g_linear_source_iterations_match_oracle(tracks):
    tg = tracks("minicore_70g")
    s = om.HipLSSolver(tg)
    assert s.setScatterPrecision(om.SCATTER_AUTO) == om.SCATTER_MFMA_F16
    s.computeEigenvalue(2)
    assert s.getNumIterations() == 2
    t0 = time.time()
    sol, mom, _ = orc.OracleProblem.from_track_generator(tg).eigenvalue_ls(
        max_iters=2, tol=1e-12, threads=THREADS)
    dk = abs(s.getKeff() - sol.keff)
    rf = rel_l2(s.getFluxes(), sol.flux)
    rm = rel_l2(s.getFluxMoments(), mom)
    log(f"mini-core LS, 2 iterations: |dk| {dk * 1e5:.4f} pcm, flux rel L2 {rf:.2e}, "
        f"moments rel L2 {rm:.2e} (oracle {time.time() - t0:.1f}s on {THREADS} threads)")
    assert dk < 1e-6
    assert rf < 2e-5
    assert rm < 1e-4


def test_minicore_70g_linear_source_matrix_cores_match_fp64_contraction(tracks):
    tg = tracks("minicore_70g")
    res = {}
    for mode in (om.SCATTER_MFMA_F16, om.SCATTER_FP64):
        s = om.HipLSSolver(tg)
        s.setScatterPrecision(mode)
        s.computeEigenvalue(1000)
        res[mode] = (s.getKeff(), s.getNumIterations(), s.getFluxes())
    (k1, i1, f1), (k2, i2, f2) = res[om.SCATTER_MFMA_F16], res[om.SCATTER_FP64]
    log(f"mini-core LS solve: matrix cores k={k1:.8f} ({i1} it), FP64 k={k2:.8f} ({i2} it), "
        f"|dk| {abs(k1 - k2) * 1e5:.4f} pcm, flux rel L2 {rel_l2(f1, f2):.2e}")
    assert abs(k1 - k2) < 1e-6
    assert rel_l2(f1, f2) < 1e-4


def test_lattice_17x17_3d_sweep_and_iterations_match_oracle(tracks):
    tg = tracks("lattice_17x17_3d_7g")
    N, G, T = tg.getNumFSRs(), tg.getNumEnergyGroups(), tg.getNumTracks()
    s = om.HipSolver(tg)
    rng = np.random.default_rng(5)
    Q = rng.uniform(0.01, 1.0, N * G)
    psi = rng.uniform(0.0, 2.0, 2 * T * G).astype(np.float32)
    tally, psi_out = s.transportSweep(Q, psi)
    op = orc.OracleProblem.from_track_generator(tg)
    t0 = time.time()
    t_ref, p_ref = op.transport_sweep(Q, psi, threads=THREADS)
    log(f"3D 17x17 sweep: tally rel L2 {rel_l2(tally, t_ref):.2e}, psi rel L2 "
        f"{rel_l2(psi_out, p_ref):.2e} (oracle {time.time() - t0:.1f}s)")
    assert rel_l2(tally, t_ref) < 2e-5
    assert rel_l2(psi_out, p_ref) < 2e-5
    s2 = om.HipSolver(tg)
    s2.computeEigenvalue(10)
    t0 = time.time()
    sol = op.eigenvalue(max_iters=10, tol=1e-12, threads=THREADS)
    log(f"3D 17x17, 10 iterations: k gpu {s2.getKeff():.8f} oracle {sol.keff:.8f} "
        f"|dk| {abs(s2.getKeff() - sol.keff) * 1e5:.4f} pcm, flux rel L2 "
        f"{rel_l2(s2.getFluxes(), sol.flux):.2e} (oracle {time.time() - t0:.1f}s)")
    assert s2.getNumIterations() == 10 == sol.num_iterations
    assert abs(s2.getKeff() - sol.keff) < 1e-5
    assert rel_l2(s2.getFluxes(), sol.flux) < 1e-4


def _perturbed(name, seed):
    p = workloads.get(name)
    rng = np.random.default_rng(seed)
    p.materials = [workloads.perturb_material(m, rng) for m in p.materials]
    return p


def test_serving_mix_with_17x17_and_70_group_minicore(tracks):
    """The served results of a 17x17 lattice, a 70-group mini-core and pin
    cells, all in flight together, against their plain solves: keff within
    1 pcm and flux relative L2 <= 1e-4 (7-group results are bitwise equal:
    their tallies are exact integer sums)."""
    names = ["minicore_70g", "lattice_17x17_7g"] + ["pin_cell_7g"] * 6
    probs = [_perturbed(n, 300 + i) for i, n in enumerate(names)]
    plain = []
    for n, p in zip(names, probs):
        s = om.HipSolver(tracks(n))
        s.setMaterials(*serve.material_arrays(p))
        s.computeEigenvalue(1000)
        plain.append((s.getKeff(), s.getNumIterations(), s.getFluxes()))
    with serve.ServeEngine(max_active=8) as eng:
        gids = {n: eng.add_geometry(tracks(n)) for n in set(names)}
        ids = [eng.submit(gids[n], p, want_flux=True) for n, p in zip(names, probs)]
        res = [eng.wait(i, timeout=300) for i in ids]
        st = eng.stats()
    assert st["failed"] == 0
    for n, r, (k, it, f) in zip(names, res, plain):
        log(f"served {n}: |dk| {abs(r.keff - k) * 1e5:.5f} pcm, iterations {r.num_iterations} "
            f"(plain {it}), flux rel L2 {rel_l2(r.flux, f):.2e}")
        assert abs(r.keff - k) < 1e-5
        assert rel_l2(r.flux, f) < 1e-4
        if tracks(n).getNumEnergyGroups() == 7:
            assert r.keff == k and r.num_iterations == it
            np.testing.assert_array_equal(r.flux.ravel(), f)
