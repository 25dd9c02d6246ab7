"""Serving engine (continuous batching) through the C ABI: each result equals
the plain single solve of the same inputs, alone and in a mix, and stays
within the parity bar of the oracle; the memory budget holds across
geometries."""
import numpy as np
import pytest

import openmoc_amd as om
from openmoc_amd import serve, workloads
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def perturbed(name, seed):
    p = workloads.get(name)
    rng = np.random.default_rng(seed)
    p.materials = [workloads.perturb_material(m, rng) for m in p.materials]
    return p


def plain_solve(tg, problem, tol=1e-5):
    s = om.HipSolver(tg)
    st, ss, nsf, chi = serve.material_arrays(problem)
    s.setMaterials(st, ss, nsf, chi)
    s.setConvergenceThreshold(tol)
    s.computeEigenvalue(1000)
    return s.getKeff(), s.getNumIterations(), s.getFluxes()


def test_served_results_equal_plain_solves(tracks):
    tg = tracks("pin_cell_7g")
    probs = [perturbed("pin_cell_7g", seed) for seed in range(6)]
    with serve.ServeEngine(max_active=4) as eng:
        gid = eng.add_geometry(tg)
        ids = [eng.submit(gid, p, want_flux=True) for p in probs]
        res = [eng.wait(i, timeout=120) for i in ids]
    for p, r in zip(probs, res):
        k, it, flux = plain_solve(tg, p)
        assert r.keff == k
        assert r.num_iterations == it
        np.testing.assert_array_equal(r.flux.ravel(), flux)
        assert r.latency_s >= r.solve_s > 0


def test_result_same_alone_and_in_the_mix(tracks):
    tg_pin = tracks("pin_cell_7g")
    tg_lat = tracks("lattice_3x3_vac_7g")
    probe = perturbed("pin_cell_7g", 100)
    with serve.ServeEngine(max_active=1) as eng:
        g = eng.add_geometry(tg_pin)
        alone = eng.wait(eng.submit(g, probe, want_flux=True), timeout=120)
    with serve.ServeEngine(max_active=8) as eng:
        g_pin = eng.add_geometry(tg_pin)
        g_lat = eng.add_geometry(tg_lat)
        ids = []
        for seed in range(12):
            if seed == 5:
                probe_id = eng.submit(g_pin, probe, want_flux=True)
            ids.append(eng.submit(g_lat if seed % 2 else g_pin,
                                  perturbed("lattice_3x3_vac_7g" if seed % 2 else "pin_cell_7g",
                                            seed)))
        mixed = eng.wait(probe_id, timeout=300)
        for i in ids:
            eng.wait(i, timeout=300)
        st = eng.stats()
    assert mixed.keff == alone.keff
    assert mixed.num_iterations == alone.num_iterations
    np.testing.assert_array_equal(mixed.flux, alone.flux)
    assert st["failed"] == 0


def test_served_result_matches_oracle(tracks):
    tg = tracks("pin_cell_7g")
    p = perturbed("pin_cell_7g", 7)
    with serve.ServeEngine() as eng:
        gid = eng.add_geometry(tg)
        r = eng.wait(eng.submit(gid, p, want_flux=True), timeout=120)
    st, ss, nsf, chi = serve.material_arrays(p)
    op = orc.OracleProblem(
        num_azim=tg.info.num_azim, num_polar=tg.info.num_polar, seg_offset=tg.seg_offset,
        seg_length=tg.seg_length, seg_fsr=tg.seg_fsr, azim=tg.azim, next_fwd=tg.next_fwd,
        next_bwd=tg.next_bwd, next_fwd_is_fwd=tg.next_fwd_is_fwd,
        next_bwd_is_fwd=tg.next_bwd_is_fwd, bc_fwd=tg.bc_fwd, bc_bwd=tg.bc_bwd,
        total_weight=tg.total_weight, sin_theta=tg.sin_theta, sigma_t=st, sigma_s=ss,
        nu_sigma_f=nsf, chi=chi, fsr_material=tg.fsr_material, fsr_volume=tg.fsr_volume)
    ref = op.eigenvalue(max_iters=1000, tol=1e-5)
    assert abs(r.keff - ref.keff) < 1e-5
    assert np.linalg.norm(r.flux - ref.flux) / np.linalg.norm(ref.flux) < 1e-4


def test_memory_admission_serialises_when_budget_is_one_solve(tracks):
    tg = tracks("pin_cell_7g")
    probs = [perturbed("pin_cell_7g", 20 + s) for s in range(4)]
    # a budget below two solves: requests are admitted one at a time
    with serve.ServeEngine(max_active=8, mem_budget_bytes=1) as eng:
        gid = eng.add_geometry(tg)
        res = [eng.wait(eng.submit(gid, p), timeout=120) for p in probs]
    for p, r in zip(probs, res):
        assert r.keff == plain_solve(tg, p)[0]


def test_memory_budget_holds_across_two_geometries(tracks):
    """A budget of one solve of each of two geometries: the solvers of both
    batches, idle ones included, never hold more than the budget, and every
    result still equals its plain solve."""
    tg_pin, tg_lat = tracks("pin_cell_7g"), tracks("lattice_3x3_vac_7g")
    with serve.ServeEngine(max_active=8) as eng:
        g_pin, g_lat = eng.add_geometry(tg_pin), eng.add_geometry(tg_lat)
        eng.wait(eng.submit(g_pin, perturbed("pin_cell_7g", 40)), timeout=120)
        eng.wait(eng.submit(g_lat, perturbed("lattice_3x3_vac_7g", 41)), timeout=120)
        one_each = eng.memory()["held"]
    reqs = [("pin_cell_7g", 50 + i) if i % 2 else ("lattice_3x3_vac_7g", 50 + i)
            for i in range(8)]
    with serve.ServeEngine(max_active=8, mem_budget_bytes=one_each) as eng:
        gids = {"pin_cell_7g": eng.add_geometry(tg_pin),
                "lattice_3x3_vac_7g": eng.add_geometry(tg_lat)}
        probs = [perturbed(n, seed) for n, seed in reqs]
        ids = [eng.submit(gids[n], p) for (n, _), p in zip(reqs, probs)]
        res = [eng.wait(i, timeout=300) for i in ids]
        mem = eng.memory()
        st = eng.stats()
    assert st["failed"] == 0
    assert mem["budget"] == one_each
    assert 0 < mem["held_max"] <= one_each
    for (n, _), p, r in zip(reqs, probs, res):
        assert r.keff == plain_solve(tracks(n), p)[0]


def test_idle_solvers_of_one_geometry_give_way_to_another(tracks):
    """Budget of one lattice solve: pin-cell requests fill their batch with
    solvers first; the lattice request that follows gets the memory once those
    solvers are idle (they are freed), and the budget is never exceeded after
    the first solver."""
    tg_pin, tg_lat = tracks("pin_cell_7g"), tracks("lattice_3x3_vac_7g")
    with serve.ServeEngine(max_active=4) as eng:
        g = eng.add_geometry(tg_lat)
        eng.wait(eng.submit(g, perturbed("lattice_3x3_vac_7g", 60)), timeout=120)
        lat_bytes = eng.memory()["held"]
    with serve.ServeEngine(max_active=4, mem_budget_bytes=lat_bytes) as eng:
        g_pin, g_lat = eng.add_geometry(tg_pin), eng.add_geometry(tg_lat)
        pins = [eng.submit(g_pin, perturbed("pin_cell_7g", 70 + i)) for i in range(4)]
        for i in pins:
            eng.wait(i, timeout=120)
        held_pins = eng.memory()["held"]
        p = perturbed("lattice_3x3_vac_7g", 61)
        r = eng.wait(eng.submit(g_lat, p), timeout=120)
        mem = eng.memory()
    assert 0 < held_pins <= lat_bytes
    assert mem["held"] <= lat_bytes
    assert r.keff == plain_solve(tg_lat, p)[0]


def test_submit_validation(tracks):
    tg = tracks("pin_cell_7g")
    with serve.ServeEngine() as eng:
        gid = eng.add_geometry(tg)
        st, ss, nsf, chi = serve.material_arrays(workloads.get("pin_cell_7g"))
        with pytest.raises(RuntimeError):
            eng.submit(gid + 5, (st, ss, nsf, chi))
        with pytest.raises(RuntimeError):
            eng.submit(gid, (st[:, :3], ss, nsf, chi))


def test_small_geometries_share_one_batch_with_plain_results(tracks):
    """Pin cells at 32 and 64 azimuthal angles and a 3x3 lattice: small
    geometries whose solves share every launch of one batch, rows of
    different sizes side by side; each result equals its plain solve bit for
    bit."""
    kinds = [("pin_cell_7g", {"num_azim": 32}), ("pin_cell_7g", {"num_azim": 64}),
             ("lattice_3x3_vac_7g", {})]
    tgs = [tracks(n, **kw) for n, kw in kinds]
    reqs = [(i % 3, 200 + i) for i in range(12)]
    probs = []
    for k, seed in reqs:
        n, kw = kinds[k]
        p = workloads.get(n, **kw)
        rng = np.random.default_rng(seed)
        p.materials = [workloads.perturb_material(m, rng) for m in p.materials]
        probs.append(p)
    with serve.ServeEngine(max_active=12) as eng:
        gids = [eng.add_geometry(tg) for tg in tgs]
        ids = [eng.submit(gids[k], p, want_flux=True) for (k, _), p in zip(reqs, probs)]
        res = [eng.wait(i, timeout=300) for i in ids]
        assert eng.stats()["failed"] == 0
    for (k, _), p, r in zip(reqs, probs, res):
        kk, it, flux = plain_solve(tgs[k], p)
        assert (r.keff, r.num_iterations) == (kk, it)
        np.testing.assert_array_equal(r.flux.ravel(), flux)


def test_mini_core_served_with_pin_cells_equals_plain_solves(tracks):
    # the serving stream's 7-group 3x3 mini-core (global-tally sweep, its own
    # batch) next to pin cells (LDS sweep): two rows of the mini-core batch
    # and the pin-cell rows give the bits of their plain solves
    tg_mc = tracks("minicore_7g")
    tg_pin = tracks("pin_cell_7g")
    probs = [("mc", perturbed("minicore_7g", 40)), ("pin", perturbed("pin_cell_7g", 41)),
             ("mc", perturbed("minicore_7g", 42)), ("pin", perturbed("pin_cell_7g", 43))]
    with serve.ServeEngine(max_active=4) as eng:
        gids = {"mc": eng.add_geometry(tg_mc), "pin": eng.add_geometry(tg_pin)}
        ids = [eng.submit(gids[kind], p, want_flux=True) for kind, p in probs]
        res = [eng.wait(i, timeout=120) for i in ids]
    for (kind, p), r in zip(probs, res):
        print(f"mini-core serving: {kind} keff {r.keff:.8f} iterations {r.num_iterations}", flush=True)
        k, it, flux = plain_solve(tg_mc if kind == "mc" else tg_pin, p)
        assert r.keff == k
        assert r.num_iterations == it
        np.testing.assert_array_equal(r.flux.ravel(), flux)
