"""Generates the golden fixtures under tests/golden/ from the compiled
reference OpenMOC (oracle/_ref/ref_driver_f64, built by `make -C oracle ref`
from the reference's own C++ sources).  Run in the build container only:

    python tests/golden/make_golden.py

Outputs (data only: inputs and the reference's outputs):
  pin_cell_7g_ref.npz       full tracks/segments/links/volumes + keff/flux
  lattice_17x17_7g_ref.npz  per-track checksums + FSR volumes + keff/flux (f32)
  c5g7_pin_cell_ref.npz     keff/flux/iterations (matches the reference's
                            tests/test_forward_pin_cell/results_true.dat)
  hom_inf_medium_ref.npz    keff/flux/iterations
  lattice_3x3_vac_7g_ref.npz  default and neutron-balance keff/flux/iterations,
                            Krylov operator sweeps (fission, scatter, fission)
                            from the converged flux
  pin_cell_7g_tc_ref.npz    keff/flux/iterations without and with each
                            transport stabilisation type
  *_ls_ref.npz              linear-source (CPULSSolver) keff/flux/iterations,
                            flux moments [N, G, 2] and FSR centroids for
                            pin_cell_7g, lattice_3x3_vac_7g and pin_cell_70g
  pin_cell_70g_ref.npz      70-group keff/flux/iterations
  pin_cell_3d_7g_ref.npz, pin_cell_3d_vac_7g_ref.npz, lattice_3x3_3d_7g_ref.npz
                            explicit 3D tracks (TrackGenerator3D, EXPLICIT_3D):
                            per-track polar index, end points, links, boundary
                            conditions and segment count, segment lengths and
                            FSR ids, corrected polar angles and weights, FSR
                            volumes, keff/flux/iterations
  lattice_3x3_vac_7g_ref.fluxes, lattice_3x3_vac_7g_it10.fluxes
                            flux files written by the reference's
                            Solver::dumpFSRFluxes after the converged solve and
                            after 10 iterations
  pin_cell_3d_7g_ref.fluxes, pin_cell_70g_ref.fluxes and *_fluxfile_ref.npz
                            the same restart and residual-by-reference runs
                            from the converged file on the 3D z-stack sweep and
                            at 70 groups
  lattice_3x3_vac_7g_fluxfile_ref.npz
                            restarts from the 10-iteration file
                            (loadInitialFSRFluxes: 5 iterations, and to
                            convergence), the solve with residuals against the
                            converged file (setResidualByReference), solves
                            continued with setRestartStatus after 10 iterations
                            and after convergence, and the fluxes the
                            reference's loadFSRFluxes reads from a file this
                            library wrote with its FSRs shuffled
  lattice_3x3_vac_7g_ls_restart_ref.npz
                            linear source: setRestartStatus after 10 iterations
  lattice_3x3_vac_7g_chi_spectrum_ref.npz
                            solve from the chi spectrum of the first fissile
                            material (setChiSpectrumMaterial)
  *_adjoint_ref.npz         adjoint solves (setSolverMode(ADJOINT)) of
                            pin_cell_7g, lattice_3x3_vac_7g and pin_cell_70g
  water_box_source_ref.npz, c5g7_pin_cell_fixed_source_ref.npz,
  lattice_3x3_vac_7g_fixed_source_ref.npz, pin_cell_7g_eig_flux0_ref.npz
                            computeFlux (from zero and continued) and
                            computeSource with fixed sources by cell, by FSR
                            and by material (the water box is the reference's
                            tests/test_compute_flux and test_compute_source
                            problem)
  c5g7_pin_cell_ref_segments.data, pin_cell_7g_ref_segments.data
                            segment files written by the reference's
                            TrackGenerator::dumpSegmentsToFile
  *.txt                     the "moc-problem 1" inputs (synthetic XS bytes)

`python tests/golden/make_golden.py flux_files` regenerates only the flux-file
fixtures, `... fixed_source` only the fixed-source ones.
"""
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from openmoc_amd import workloads  # noqa: E402
from oracle import mocdump, oracle  # noqa: E402

OUT = Path(__file__).resolve().parent


def run(name: str, threads: int = 8, extra=(), fp: str = "f64"):
    p = workloads.get(name)
    txt = OUT / f"{name}.txt"
    txt.write_text(p.to_text())
    with tempfile.TemporaryDirectory() as td:
        dump = Path(td) / "ref.bin"
        oracle.run_reference(str(txt), str(dump), threads=threads, fp=fp, extra=extra)
        return p, mocdump.read(str(dump))


F32_THREADS = (1, 2, 4, 8)


def f32_iterations(name: str, extra, steps: int):
    """Iteration counts [steps][len(F32_THREADS)] of the same solve sequence
    in the reference's FP32 build (its default), on 1, 2, 4 and 8 threads.
    Near a 1e-7 threshold the FP32 sweep's rounding decides the last
    iterations, so these counts spread; the spread is the bar for this
    library's FP32 sweep."""
    out = np.zeros((steps, len(F32_THREADS)), np.int64)
    for j, th in enumerate(F32_THREADS):
        _, d = run(name, threads=th, extra=extra, fp="f32")
        for i in range(steps):
            out[i, j] = int(d[f"step{i}_num_iterations"])
    return out


def base(d):
    return dict(keff=np.float64(d["keff"]), num_iterations=np.int64(d["num_iterations"]),
                num_fsrs=np.int64(d["num_fsrs"]), num_segments=np.int64(len(d["seg_length"])),
                num_tracks=np.int64(len(d["trk_azim"])))


def flux_files():
    """Flux-file fixtures (Solver::dumpFSRFluxes / loadInitialFSRFluxes /
    setResidualByReference / loadFSRFluxes, src/Solver.cpp:2056-2318)."""
    import openmoc_amd as om

    name = "lattice_3x3_vac_7g"
    conv, part = OUT / f"{name}_ref.fluxes", OUT / f"{name}_it10.fluxes"
    _, d = run(name, threads=1, extra=("--dump-fluxes", str(conv)))
    _, d10 = run(name, threads=1, extra=("--max-iters", "10", "--dump-fluxes", str(part)))
    _, r5 = run(name, threads=1, extra=("--restart-from", str(part), "--max-iters", "5"))
    _, rc = run(name, threads=1, extra=("--restart-from", str(part)))
    _, rr = run(name, threads=1, extra=("--residual-reference", str(conv)))
    # setRestartStatus after 10 iterations, and after a converged solve
    _, rs10 = run(name, threads=1, extra=("--restart-after", "10"))
    _, rs40 = run(name, threads=1, extra=("--restart-after", "40"))
    # a file written by this library: synthetic fluxes on its own centroids,
    # FSR order shuffled, read back by the reference onto its FSRs
    tg = om.TrackGenerator(workloads.get(name))
    tg.generateTracks()
    xyz = om.fsr_centroids(tg)
    N, G = len(xyz), tg.getNumEnergyGroups()
    rng = np.random.default_rng(11)
    flux = rng.uniform(0.5, 2.0, (N, G))
    perm = rng.permutation(N)
    with tempfile.TemporaryDirectory() as td:
        ours = Path(td) / "ours.fluxes"
        om.write_flux_file(ours, 1.2345, xyz[perm], flux[perm])
        _, dl = run(name, threads=1, extra=("--load-fluxes", str(ours)))
    np.savez_compressed(
        OUT / f"{name}_fluxfile_ref.npz",
        keff=np.float64(d["keff"]), flux=d["flux"], num_iterations=np.int64(d["num_iterations"]),
        it10_keff=np.float64(d10["keff"]), it10_flux=d10["flux"],
        restart5_keff=np.float64(r5["keff"]), restart5_flux=r5["flux"],
        restart5_num_iterations=np.int64(r5["num_iterations"]),
        restart_keff=np.float64(rc["keff"]), restart_flux=rc["flux"],
        restart_num_iterations=np.int64(rc["num_iterations"]),
        residual_ref_keff=np.float64(rr["keff"]), residual_ref_flux=rr["flux"],
        residual_ref_num_iterations=np.int64(rr["num_iterations"]),
        restart10_keff=np.float64(rs10["keff"]), restart10_flux=rs10["flux"],
        restart10_num_iterations=np.int64(rs10["num_iterations"]),
        restart40_keff=np.float64(rs40["keff"]), restart40_flux=rs40["flux"],
        restart40_num_iterations=np.int64(rs40["num_iterations"]),
        written_flux=flux, written_perm=perm, loaded_flux=dl["loaded_flux"].reshape(N, G),
        fsr_centroid=d["fsr_centroid"].reshape(N, 2))


def flux_files_other_paths():
    """Restart (loadInitialFSRFluxes) and residual-by-reference runs from the
    reference's own converged file, on the 3D pin cell (z-stack sweep) and the
    70-group pin cell (per-FSR source residual terms)."""
    for name in ("pin_cell_3d_7g", "pin_cell_70g"):
        conv = OUT / f"{name}_ref.fluxes"
        _, d = run(name, threads=1, extra=("--dump-fluxes", str(conv)))
        _, rc = run(name, threads=1, extra=("--restart-from", str(conv)))
        _, rr = run(name, threads=1, extra=("--residual-reference", str(conv)))
        np.savez_compressed(
            OUT / f"{name}_fluxfile_ref.npz",
            keff=np.float64(d["keff"]), num_iterations=np.int64(d["num_iterations"]),
            restart_keff=np.float64(rc["keff"]), restart_flux=rc["flux"],
            restart_num_iterations=np.int64(rc["num_iterations"]),
            residual_ref_keff=np.float64(rr["keff"]), residual_ref_flux=rr["flux"],
            residual_ref_num_iterations=np.int64(rr["num_iterations"]))


def ls_restart():
    """Linear source: setRestartStatus after 10 iterations."""
    _, d = run("lattice_3x3_vac_7g", threads=1, extra=("--linear-source", "--restart-after", "10"))
    N, G = int(d["num_fsrs"]), int(d["num_groups"])
    mom = d["flux_xyz"].reshape(N, 3, G)[:, :2, :].transpose(0, 2, 1)
    np.savez_compressed(OUT / "lattice_3x3_vac_7g_ls_restart_ref.npz", flux=d["flux"],
                        flux_moments=mom, keff=np.float64(d["keff"]),
                        num_iterations=np.int64(d["num_iterations"]))


def adjoint():
    """Adjoint solves (setSolverMode(ADJOINT)): the 2D LDS path at 7 groups
    and the matrix-core contraction at 70 groups."""
    for name in ("pin_cell_7g", "lattice_3x3_vac_7g", "pin_cell_70g"):
        _, d = run(name, threads=1, extra=("--adjoint",))
        np.savez_compressed(OUT / f"{name}_adjoint_ref.npz", flux=d["flux"],
                            keff=np.float64(d["keff"]),
                            num_iterations=np.int64(d["num_iterations"]))


def chi_spectrum():
    """Initial guess from a material's chi (setChiSpectrumMaterial)."""
    name = "lattice_3x3_vac_7g"
    p = workloads.get(name)
    m = next(i for i, mat in enumerate(p.materials) if any(c > 0 for c in mat.chi))
    _, d = run(name, threads=1, extra=("--chi-spectrum", str(m)))
    np.savez_compressed(OUT / f"{name}_chi_spectrum_ref.npz", flux=d["flux"],
                        keff=np.float64(d["keff"]), num_iterations=np.int64(d["num_iterations"]),
                        material=np.int64(m))


WATER_BOX_FLUX_SOURCES = ((1, 1.0), (2, 0.5), (3, 0.25))   # test_compute_flux.py:86-88
WATER_BOX_SOURCE_SOURCES = ((1, 1.0),)                      # test_compute_source.py:84


def fsr_source_args(fixed):
    """--fixed-source-fsr arguments for a [N, G] array of per-FSR sources."""
    out = []
    for r, g in zip(*np.nonzero(fixed)):
        out += ["--fixed-source-fsr", str(r), str(g + 1), repr(float(fixed[r, g]))]
    return out


def pin_cell_fixed_sources(N, G):
    """Per-FSR sources used on the C5G7 pin cell: group 1 everywhere and
    group 3 in the first half of the FSRs."""
    fixed = np.zeros((N, G))
    fixed[:, 0] = 1.0
    fixed[: N // 2, 2] = 0.5
    return fixed


def fixed_source():
    """computeFlux / computeSource (src/Solver.cpp:1352-1516) of the compiled
    reference: the water box of the reference's own tests/test_compute_flux
    and tests/test_compute_source (a fixed source by cell), per-FSR sources on
    the C5G7 pin cell, a material source overlapping per-FSR sources on the
    leaking 3x3 lattice, and computeFlux(only_fixed_source=False) continuing
    an eigenvalue solve without fixed sources."""
    out = {}
    cell = lambda srcs: sum((["--fixed-source-cell", "1", "0", str(g), repr(v)] for g, v in srcs), [])
    _, d = run("water_box_source", threads=1,
               extra=(*cell(WATER_BOX_FLUX_SOURCES), "--sequence", "flux,flux0"))
    out.update(flux_flux=d["step0_flux"], flux_num_iterations=np.int64(d["step0_num_iterations"]),
               flux0_flux=d["step1_flux"], flux0_num_iterations=np.int64(d["step1_num_iterations"]),
               fsr_point=d["fsr_point"], fsr_volume=d["fsr_volume"])
    _, d = run("water_box_source", threads=1,
               extra=(*cell(WATER_BOX_SOURCE_SOURCES), "--sequence", "source:1.0"))
    out.update(source_flux=d["step0_flux"], source_num_iterations=np.int64(d["step0_num_iterations"]))
    np.savez_compressed(OUT / "water_box_source_ref.npz", **out)

    _, d0 = run("c5g7_pin_cell", threads=1, extra=("--no-solve",))
    N, G = int(d0["num_fsrs"]), int(d0["num_groups"])
    fixed = pin_cell_fixed_sources(N, G)
    extra = (*fsr_source_args(fixed), "--tol", "1e-7", "--max-iters", "1000",
             "--sequence", "flux,flux0,source:3.0:TOTAL_SOURCE")
    _, d = run("c5g7_pin_cell", threads=1, extra=extra)
    f32 = f32_iterations("c5g7_pin_cell", extra, 3)
    np.savez_compressed(
        OUT / "c5g7_pin_cell_fixed_source_ref.npz", fixed=fixed,
        flux_flux=d["step0_flux"], flux_num_iterations=np.int64(d["step0_num_iterations"]),
        flux0_flux=d["step1_flux"], flux0_num_iterations=np.int64(d["step1_num_iterations"]),
        source_flux=d["step2_flux"], source_num_iterations=np.int64(d["step2_num_iterations"]),
        flux_num_iterations_f32=f32[0], flux0_num_iterations_f32=f32[1],
        source_num_iterations_f32=f32[2])

    # material 1 (moderator) group 1 overlaps a per-FSR source set before
    # and one set after it: the material's value wins in both FSRs
    name = "lattice_3x3_vac_7g"
    _, d0 = run(name, threads=1)
    mat = np.asarray(d0["fsr_material"])
    mod = np.nonzero(mat == 1)[0]
    fuel = np.nonzero(mat == 0)[0]
    srcs = [("fsr", int(mod[0]), 1, 3.0), ("material", 1, 1, 0.5), ("fsr", int(mod[5]), 1, 4.0),
            ("fsr", int(fuel[0]), 2, 1.0)]
    args = []
    for kind, a, g, v in srcs:
        args += [f"--fixed-source-{kind}", str(a), str(g), repr(v)]
    extra = (*args, "--tol", "1e-7", "--sequence", "source:0.8:TOTAL_SOURCE,flux0")
    _, d = run(name, threads=1, extra=extra)
    f32 = f32_iterations(name, extra, 2)
    np.savez_compressed(
        OUT / f"{name}_fixed_source_ref.npz",
        sources=np.array([(0 if k == "fsr" else 1, a, g, v) for k, a, g, v in srcs]),
        source_flux=d["step0_flux"], source_num_iterations=np.int64(d["step0_num_iterations"]),
        flux0_flux=d["step1_flux"], flux0_num_iterations=np.int64(d["step1_num_iterations"]),
        source_num_iterations_f32=f32[0], flux0_num_iterations_f32=f32[1])

    _, d = run("pin_cell_7g", threads=1, extra=("--sequence", "eig,flux0"))
    np.savez_compressed(
        OUT / "pin_cell_7g_eig_flux0_ref.npz",
        eig_keff=np.float64(d["step0_keff"]), eig_flux=d["step0_flux"],
        eig_num_iterations=np.int64(d["step0_num_iterations"]),
        flux0_flux=d["step1_flux"], flux0_num_iterations=np.int64(d["step1_num_iterations"]))


def segment_files():
    """Segment files written by the reference's own
    TrackGenerator::dumpSegmentsToFile (src/TrackGenerator.cpp:1388-1501) for
    the C5G7 pin cell and pin_cell_7g, kept byte for byte."""
    import glob
    import shutil
    for name in ("c5g7_pin_cell", "pin_cell_7g"):
        p = workloads.get(name)
        txt = OUT / f"{name}.txt"
        txt.write_text(p.to_text())
        with tempfile.TemporaryDirectory() as td:
            oracle.run_reference(str(txt), str(Path(td) / "ref.bin"), threads=1, fp="f64",
                                 extra=("--no-solve", "--dump-segments", td))
            (src,) = glob.glob(str(Path(td) / "tracks" / "*.data"))
            shutil.copyfile(src, OUT / f"{name}_ref_segments.data")


def main():
    if sys.argv[1:] == ["segment_files"]:
        segment_files()
        return
    if sys.argv[1:] == ["fixed_source"]:
        fixed_source()
        return
    if sys.argv[1:] == ["flux_files"]:
        flux_files()
        flux_files_other_paths()
        return
    if sys.argv[1:] == ["flux_files_other_paths"]:
        flux_files_other_paths()
        return
    if sys.argv[1:] == ["ls_restart"]:
        ls_restart()
        return
    if sys.argv[1:] == ["adjoint"]:
        adjoint()
        return
    if sys.argv[1:] == ["chi_spectrum"]:
        chi_spectrum()
        return
    for name in ("c5g7_pin_cell", "hom_inf_medium"):
        _, d = run(name, threads=1)
        np.savez_compressed(OUT / f"{name}_ref.npz", flux=d["flux"], **base(d))

    _, d = run("pin_cell_7g", threads=1)
    np.savez_compressed(
        OUT / "pin_cell_7g_ref.npz", flux=d["flux"], seg_length=d["seg_length"],
        seg_fsr=d["seg_fsr"].astype(np.int32), trk_num_segments=d["trk_num_segments"],
        trk_next_fwd=d["trk_next_fwd"], trk_next_bwd=d["trk_next_bwd"],
        trk_next_fwd_is_fwd=d["trk_next_fwd_is_fwd"], trk_next_bwd_is_fwd=d["trk_next_bwd_is_fwd"],
        trk_bc_fwd=d["trk_bc_fwd"], trk_bc_bwd=d["trk_bc_bwd"], trk_azim=d["trk_azim"],
        trk_start=d["trk_start"], quad_total_weight=d["quad_total_weight"],
        quad_phi=d["quad_phi"], fsr_volume=d["fsr_volume"], fsr_material=d["fsr_material"],
        **base(d))

    _, d = run("lattice_17x17_7g", threads=8)
    ns = d["trk_num_segments"]
    off = np.concatenate([[0], np.cumsum(ns)])
    len_sum = np.add.reduceat(d["seg_length"], off[:-1])
    fsr_sum = np.add.reduceat(d["seg_fsr"].astype(np.int64), off[:-1])
    np.savez_compressed(
        OUT / "lattice_17x17_7g_ref.npz", flux=d["flux"].astype(np.float32),
        trk_num_segments=ns.astype(np.int32), trk_length_sum=len_sum, trk_fsr_sum=fsr_sum,
        trk_next_fwd=d["trk_next_fwd"], trk_next_bwd=d["trk_next_bwd"],
        fsr_volume=d["fsr_volume"], fsr_material=d["fsr_material"], **base(d))
    _, d = run("lattice_3x3_vac_7g", threads=1, extra=("--krylov",))
    _, db = run("lattice_3x3_vac_7g", threads=1, extra=("--balance",))
    np.savez_compressed(
        OUT / "lattice_3x3_vac_7g_ref.npz", flux=d["flux"],
        trk_num_segments=d["trk_num_segments"].astype(np.int32),
        balance_keff=np.float64(db["keff"]), balance_flux=db["flux"],
        balance_num_iterations=np.int64(db["num_iterations"]),
        krylov_phi0=d["krylov_phi0"], krylov_fission_1=d["krylov_fission_1"],
        krylov_scatter=d["krylov_scatter"], krylov_fission_2=d["krylov_fission_2"], **base(d))

    _, d = run("pin_cell_7g_tc", threads=1)
    out = dict(flux=d["flux"], **base(d))
    for tag, factor, typ in (("diagonal", 1.0, 0), ("yamamoto", 0.5, 1), ("global", 0.5, 2)):
        _, ds = run("pin_cell_7g_tc", threads=1, extra=("--stabilize", str(factor), str(typ)))
        out[f"{tag}_keff"] = np.float64(ds["keff"])
        out[f"{tag}_flux"] = ds["flux"]
        out[f"{tag}_num_iterations"] = np.int64(ds["num_iterations"])
        out[f"{tag}_factor"] = np.float64(factor)
    np.savez_compressed(OUT / "pin_cell_7g_tc_ref.npz", **out)

    for name in ("pin_cell_3d_7g", "pin_cell_3d_vac_7g", "lattice_3x3_3d_7g"):
        _, d = run(name, threads=1)
        np.savez_compressed(
            OUT / f"{name}_ref.npz", flux=d["flux"], seg_length=d["seg_length"],
            seg_fsr=d["seg_fsr"].astype(np.int32), trk_num_segments=d["trk_num_segments"],
            trk_azim=d["trk_azim"], trk_xy=d["trk_xy"], trk_polar=d["trk_polar"],
            trk_start=d["trk_start"], trk_end=d["trk_end"], trk_next_fwd=d["trk_next_fwd"],
            trk_next_bwd=d["trk_next_bwd"], trk_next_fwd_is_fwd=d["trk_next_fwd_is_fwd"],
            trk_next_bwd_is_fwd=d["trk_next_bwd_is_fwd"], trk_bc_fwd=d["trk_bc_fwd"],
            trk_bc_bwd=d["trk_bc_bwd"], quad_theta=d["quad_theta"],
            quad_total_weight=d["quad_total_weight"], fsr_volume=d["fsr_volume"],
            fsr_material=d["fsr_material"], **base(d))

    _, d = run("pin_cell_70g", threads=1)
    np.savez_compressed(OUT / "pin_cell_70g_ref.npz", flux=d["flux"], **base(d))

    for name in ("pin_cell_7g", "lattice_3x3_vac_7g", "pin_cell_70g"):
        _, d = run(name, threads=1, extra=("--linear-source",))
        N, G = int(d["num_fsrs"]), int(d["num_groups"])
        mom = d["flux_xyz"].reshape(N, 3, G)[:, :2, :].transpose(0, 2, 1)
        np.savez_compressed(OUT / f"{name}_ls_ref.npz", flux=d["flux"], flux_moments=mom,
                            fsr_centroid=d["fsr_centroid_ls"].reshape(N, 2), **base(d))
    flux_files()
    flux_files_other_paths()
    ls_restart()
    adjoint()
    chi_spectrum()
    fixed_source()
    segment_files()
    print("golden fixtures written to", OUT)


if __name__ == "__main__":
    main()
