#!/usr/bin/env python
"""Benchmark of the MOC source iteration on one MI355X.

Workload (default): 2D 17x17 pin lattice, 7 synthetic groups, 64 azimuthal
angles, 0.03 cm spacing, TY 6 polar angles, reflective -- 29,144 tracks,
3,285,080 segments, 9,248 FSRs.  A *step* is one full source iteration of
Solver::computeEigenvalue (transport sweep + addSourceToScalarFlux + keff +
normalisation + residual + next sources), all on the GPU with inputs resident
in HBM.  value = segment-group integrations per second over the timed steps,
with the reference's integration count 2 * G * (num_polar/2) * segments per
sweep (src/Solver.cpp:1901-1905).

Also reported: one full keff solve (seconds, iterations, error vs the
compiled reference's converged k and fluxes), the sweep kernel against the
HBM and FP32-VALU roofs, and the reference CPUSolver (FP32 build, OpenMP) on
the host cores over a bounded number of sweeps of the same problem.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "segment-group integrations/s + keff-solves/s, p50/p99 latency at ≤1 pcm"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 (vector, packed)
FLOPS_PER_INTEGRATION = 32     # rational F1 + attenuation + tally (FMA = 2)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(problem_text: str, G: int, P: int, S: int, sweeps: int, threads: int):
    """Reference CPUSolver (FP32 build, OpenMP) timed on the host: a bounded
    sample of `sweeps` transport sweeps after 2 source iterations."""
    from oracle import mocdump, oracle as orc
    exe = orc.REF_F32
    with tempfile.TemporaryDirectory() as td:
        txt = Path(td) / "problem.txt"
        txt.write_text(problem_text)
        if exe.exists():
            dump = Path(td) / "ref.bin"
            try:
                subprocess.run([str(exe), str(txt), str(dump), "--threads", str(threads),
                                "--max-iters", "2", "--sweeps", str(sweeps)],
                               check=True, stdout=subprocess.DEVNULL, timeout=600)
                d = mocdump.read(str(dump))
                secs = float(d["timed_sweeps_s"])
                return {"value": 2.0 * G * P * S * sweeps / secs, "unit": "integrations/s",
                        "cores": threads, "kind": "reference",
                        "sample": f"{sweeps} transport sweeps of the same workload by the "
                                  f"reference CPUSolver (FP32 build, -O3 -ffast-math, "
                                  f"{threads} OpenMP threads), after 2 source iterations",
                        "seconds": secs}
            except (subprocess.SubprocessError, OSError, KeyError) as e:
                log(f"reference CPU baseline unavailable ({e}); timing the oracle port")
        # fallback: the FP64 restatement (oracle port)
        import openmoc_amd as om
        tg = om.TrackGenerator(problem_text)
        tg.generateTracks()
        op = orc.OracleProblem.from_track_generator(tg)
        sol = op.eigenvalue(max_iters=sweeps, tol=0.0, threads=threads)
        return {"value": sol.num_integrations / sol.sweep_seconds, "unit": "integrations/s",
                "cores": threads, "kind": "port",
                "sample": f"{sweeps} source iterations of the FP64 oracle restatement "
                          f"({threads} OpenMP threads)", "seconds": sol.sweep_seconds}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="lattice_17x17_7g")
    ap.add_argument("--exp", choices=["rational", "hw"], default="rational")
    ap.add_argument("--cpu-sweeps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        return 2

    import openmoc_amd as om
    from openmoc_amd import workloads

    problem = workloads.get(args.workload)
    text = problem.to_text()
    t0 = time.perf_counter()
    tg = om.TrackGenerator(problem)
    tg.generateTracks()
    log(f"tracks: {tg.getNumTracks()} tracks, {tg.getNumSegments()} segments, "
        f"{tg.getNumFSRs()} FSRs ({time.perf_counter() - t0:.2f}s)")
    G, P = tg.getNumEnergyGroups(), tg.info.num_polar // 2
    S, T, N = tg.getNumSegments(), tg.getNumTracks(), tg.getNumFSRs()
    w_sweep = 2 * G * P * S

    solver = om.HipSolver(tg, device=0)
    solver.setExponential(om.EXP_HW if args.exp == "hw" else om.EXP_RATIONAL)
    shape = solver.getLaunchShape()
    if not args.no_autotune:
        from openmoc_amd.autotune import autotune
        solver.computeEigenvalue(3)
        shape, t_best = autotune(solver, repeats=5)
        log(f"autotune: {shape} -> {t_best * 1e6:.1f} us/sweep")

    solve = None
    if not args.no_solve:
        solver.computeEigenvalue(problem.max_iters)          # warm (JIT of code objects, caches)
        t0 = time.perf_counter()
        solver.computeEigenvalue(problem.max_iters)
        solve_s = time.perf_counter() - t0
        keff = solver.getKeff()
        solve = {"seconds": solve_s, "solves_per_s": 1.0 / solve_s,
                 "iterations": solver.getNumIterations(), "keff": keff}
        golden = ROOT / "tests" / "golden" / f"{args.workload}_ref.npz"
        if golden.exists():
            g = np.load(golden)
            flux = solver.getFluxes()
            ref = g["flux"].astype(np.float64)
            solve["dk_pcm_vs_reference"] = (keff - float(g["keff"])) * 1e5
            solve["flux_rel_l2_vs_reference"] = float(np.linalg.norm(flux - ref) / np.linalg.norm(ref))
            solve["reference_iterations"] = int(g["num_iterations"])
        log(f"solve: {solve}")

    # steps: full source iterations, stopping rule disabled
    solver.runIterations(max(args.warmup, 1))
    t0 = time.perf_counter()
    sweep_s = solver.runIterations(args.steps)   # synchronises before returning
    elapsed = time.perf_counter() - t0
    value = w_sweep * args.steps / elapsed
    ms_per_step = elapsed / args.steps * 1e3

    bytes_sweep = 8 * S + 8 * N * G + 16 * T * G * P
    achieved_gbs = bytes_sweep / sweep_s / 1e9
    achieved_tflops = FLOPS_PER_INTEGRATION * w_sweep / sweep_s / 1e12
    roofline = {"bound": "hbm", "achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": None,
                "kernel": "sweep2d_lds_kernel" if shape.get("lds_groups", 0) > 0 else "sweep2d_kernel",
                "algorithmic_bytes_per_launch": bytes_sweep,
                "launch_seconds": sweep_s}
    roofline_valu = {"achieved": achieved_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP32_PEAK_TFLOPS,
                     "flops_per_integration": FLOPS_PER_INTEGRATION}

    cpu = None
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
        threads = max(1, min(threads, os.cpu_count() or 1))
        log(f"cpu baseline: {args.cpu_sweeps} sweeps on {threads} threads ...")
        cpu = cpu_baseline(text, G, P, S, args.cpu_sweeps, threads)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC, "value": value, "unit": "integrations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "tracks": T, "segments": S, "fsrs": N,
                   "groups": G, "polar_half": P, "num_azim": problem.num_azim,
                   "azim_spacing": problem.azim_spacing, "exponential": args.exp,
                   "launch_shape": shape,
                   "step": "one source iteration (sweep + per-FSR FP64 steps)"},
        "sweep_integrations_per_s": w_sweep / sweep_s,
        "roofline": roofline, "roofline_valu": roofline_valu,
        "cpu_baseline": cpu, "solve": solve,
    }
    if cpu:
        out["speedup_vs_cpu_sweep"] = (w_sweep / sweep_s) / cpu["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
