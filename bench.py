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
# the same VALUs without packed FP32 instructions, which this build does not
# emit (one FP32 FMA per lane per clock)
FP32_UNPACKED_PEAK_TFLOPS = FP32_PEAK_TFLOPS / 2
FLOPS_PER_INTEGRATION = 32     # rational F1 + attenuation + tally (FMA = 2)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(problem_text: str, G: int, P: int, S: int, sweeps: int, threads: int,
                 use_reference: bool = True):
    """Reference CPUSolver (FP32 build, OpenMP) timed on the host: a bounded
    sample of `sweeps` transport sweeps after 2 source iterations.  With
    use_reference False (3D: the reference's explicit 3D setup alone takes
    minutes), the FP64 oracle restatement is timed instead."""
    from oracle import mocdump, oracle as orc
    exe = orc.REF_F32
    with tempfile.TemporaryDirectory() as td:
        txt = Path(td) / "problem.txt"
        txt.write_text(problem_text)
        if exe.exists() and use_reference:
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


def host_cpu():
    """(threads to use, description) of the host cores this process may run
    on: the scheduler affinity, capped by the cgroup CPU quota where one is
    set (a GPU box grants each GPU's jobs a share of its cores)."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) // int(period)))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    threads = min(affinity, quota) if quota else affinity
    return threads, {"cpu_model": model, "affinity_cores": affinity, "cgroup_quota_cores": quota,
                     "os_cpu_count": os.cpu_count()}


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by what the
    counter reads for a plain device copy of known size), or None."""
    f = ROOT / "profiles" / f"pmc_{kernel}_{workload}.json"
    if not f.exists():
        return None, None
    d = json.loads(f.read_text())
    return float(d["traffic_bytes_per_launch"]), f"{f.relative_to(ROOT)}"


def copy_calibration(mib: int):
    """The library's plain device copy (copy_kernel) of `mib` MiB, three
    times: rocprofv3 counters of that kernel calibrate FETCH_SIZE/WRITE_SIZE
    (it reads and writes exactly mib MiB per launch)."""
    import ctypes as C
    from openmoc_amd._lib import check, lib
    t = C.c_double()
    check(lib().moc_device_copy_calibration(0, mib << 20, 3, C.byref(t)))
    log(f"copy calibration: 3 copies of {mib << 20} bytes, {t.value * 1e6:.1f} us each "
        f"({2 * (mib << 20) / t.value / 1e9:.0f} GB/s)")


# Serving mix: (kind, workload, keyword arguments, share of the requests).
# Pin cells and 3x3 pin lattices (the small "mini-cores") at 7 and 70 groups
# and 16 to 128 azimuthal angles, and 17x17 lattices.  The 70-group 3x3
# assembly mini-core is left out: each of its solves (about 2 s, 560
# iterations of 3 ms) fills the GPU alone, so it would time one solve, not
# the serving of a stream; its served result is checked in the test suite.
SERVING_MIX = [
    ("pin_cell_7g@32", "pin_cell_7g", {"num_azim": 32}, 0.22),
    ("pin_cell_7g@64", "pin_cell_7g", {"num_azim": 64}, 0.14),
    ("pin_cell_7g@128", "pin_cell_7g", {"num_azim": 128}, 0.10),
    ("pin_cell_70g@32", "pin_cell_70g", {"num_azim": 32}, 0.10),
    ("lattice_3x3_vac_7g@32", "lattice_3x3_vac_7g", {"num_azim": 32}, 0.14),
    ("lattice_3x3_vac_7g@64", "lattice_3x3_vac_7g", {"num_azim": 64}, 0.14),
    ("minicore_7g@32", "minicore_7g", {"num_azim": 32}, 0.06),
    ("lattice_17x17_7g@64", "lattice_17x17_7g", {"num_azim": 64}, 0.10),
]


def serving_bench(n_requests: int, seed: int = 0, max_active: int = 16):
    """Serving stream: eigenvalue solves of the SERVING_MIX kinds, each with
    per-request seeded +-10% per-group cross-section perturbations.  Measures
    (1) back-to-back plain solves, (2) the engine's sustained solves/s with
    all requests queued, (3) p50/p99 latency under Poisson arrivals at 70% of
    that rate, and compares every served result with its plain solve."""
    import openmoc_amd as om
    from openmoc_amd import serve, workloads
    rng = np.random.default_rng(seed)
    tgs, problems = {}, {}
    for kind, name, kw, _ in SERVING_MIX:
        problems[kind] = (name, kw)
        tg = om.TrackGenerator(workloads.get(name, **kw))
        tg.generateTracks()
        tgs[kind] = tg
    names = rng.choice([k for k, _, _, _ in SERVING_MIX], size=n_requests,
                       p=[w for _, _, _, w in SERVING_MIX])
    reqs = []
    for i, kind in enumerate(names):
        name, kw = problems[kind]
        p = workloads.get(name, **kw)
        r = np.random.default_rng(1000 + i)
        p.materials = [workloads.perturb_material(m, r) for m in p.materials]
        reqs.append((kind, serve.material_arrays(p)))

    # (1) back-to-back plain solves (one solver per geometry, re-armed per request)
    solvers = {n: om.HipSolver(tg) for n, tg in tgs.items()}
    for n, s in solvers.items():       # warm (code objects, graphs)
        s.computeEigenvalue(1000)
    plain, t_req = [], []
    t0 = time.perf_counter()
    for name, xs in reqs:
        t1 = time.perf_counter()
        s = solvers[name]
        s.setMaterials(*xs)
        s.computeEigenvalue(1000)
        plain.append((s.getKeff(), s.getFluxes()))
        t_req.append(time.perf_counter() - t1)
    t_plain = time.perf_counter() - t0

    def run_engine(arrivals, subset=None):
        stream = [reqs[i] for i in subset] if subset is not None else reqs
        with serve.ServeEngine(max_active=max_active) as eng:
            gids = {n: eng.add_geometry(tg) for n, tg in tgs.items()}
            # warm-up: as many solves of each kind at once as the stream will
            # have in flight, so the engine's row solvers (one per solve in
            # flight, kept between requests) exist before the timed stream
            warm = []
            for n in gids:
                name, kw = problems[n]
                xs = serve.material_arrays(workloads.get(name, **kw))
                for _ in range(min(sum(1 for k, _ in stream if k == n), max_active)):
                    warm.append(eng.submit(gids[n], xs))
            for i in warm:
                eng.wait(i, 300)
            ids = []
            t0 = time.perf_counter()
            for (name, xs), ta in zip(stream, arrivals):
                d = t0 + ta - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids.append(eng.submit(gids[name], xs, want_flux=True))
            res = [eng.wait(i, 600) for i in ids]
            return time.perf_counter() - t0, res

    # (2) saturation: everything queued at t = 0
    t_sat, res = run_engine(np.zeros(n_requests))
    rate_sat = n_requests / t_sat
    # (3) Poisson arrivals at 70% of the sustained rate
    gaps = rng.exponential(1.0 / (0.7 * rate_sat), size=n_requests)
    t_poi, res_p = run_engine(np.cumsum(gaps) - gaps[0])
    lat = np.array([r.latency_s for r in res_p])

    # (4) the same stream without its mini-cores (the longest solves,
    # latency-bound on the device: they set the ratio of the whole stream)
    sub = [i for i, (k, _) in enumerate(reqs) if not k.startswith("minicore")]
    t_sub, _ = run_engine(np.zeros(len(sub)), sub)
    t_sub_plain = float(sum(t_req[i] for i in sub))
    by_kind = {}
    for (k, _), t in zip(reqs, t_req):
        by_kind[k] = by_kind.get(k, 0.0) + t

    dk = max(abs(r.keff - k) for r, (k, _) in zip(res + res_p, plain + plain))
    dphi = max(float(np.linalg.norm(r.flux.ravel() - f) / np.linalg.norm(f))
               for r, (_, f) in zip(res + res_p, plain + plain))
    counts = {k: int(np.sum(names == k)) for k, _, _, _ in SERVING_MIX}
    return {
        "requests": n_requests, "mix": counts, "max_active": max_active,
        "plain_solves_per_s": n_requests / t_plain,
        "served_solves_per_s": rate_sat,
        "speedup_vs_plain": rate_sat / (n_requests / t_plain),
        "poisson_rate_per_s": 0.7 * rate_sat,
        "poisson_solves_per_s": n_requests / t_poi,
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "max_dk_pcm_vs_plain": dk * 1e5, "max_flux_rel_l2_vs_plain": dphi,
        "plain_seconds_by_kind": by_kind,
        "without_minicores": {"requests": len(sub),
                              "speedup_vs_plain": t_sub_plain / t_sub if t_sub > 0 else None},
        "note": "each served result compared with the plain single solve of the same "
                "inputs (bitwise equal expected); parity of plain solves vs the "
                "compiled reference is in the solve object and the test suite",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="lattice_17x17_7g")
    ap.add_argument("--exp", choices=["rational", "hw"], default="rational")
    ap.add_argument("--linear-source", action="store_true",
                    help="linear-source solver (CPULSSolver form) instead of the flat source")
    ap.add_argument("--scatter", choices=["auto", "fp64", "mfma", "fp32"], default="auto",
                    help="scattering-source arithmetic (auto: matrix cores from 32 groups on)")
    ap.add_argument("--cpu-sweeps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--launch-shape", default="", metavar="K=V,...",
                    help="fixed sweep launch shape (keys of getLaunchShape), no autotuning")
    ap.add_argument("--serving", type=int, default=48, metavar="N",
                    help="requests in the serving-stream measurement (0: skip)")
    ap.add_argument("--max-active", type=int, default=48, metavar="K",
                    help="solves in flight at once in the serving measurement (48: the whole "
                         "stream may be in flight; 16 and 32 measured 2.84-2.89x and 2.92x)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="after the run, copy a MIB-MiB device buffer 3 times (a plain copy kernel "
                         "of known size, the reference for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
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
    # fluxes per track end: G x num_polar/2 in 2D, G in 3D (src/Solver.cpp:433-441)
    G, P = tg.getNumEnergyGroups(), (1 if tg.info.is_3d else tg.info.num_polar // 2)
    S, T, N = tg.getNumSegments(), tg.getNumTracks(), tg.getNumFSRs()
    w_sweep = 2 * G * P * S

    solver = (om.HipLSSolver if args.linear_source else om.HipSolver)(tg, device=0)
    solver.setExponential(om.EXP_HW if args.exp == "hw" else om.EXP_RATIONAL)
    scatter = solver.setScatterPrecision({"auto": om.SCATTER_AUTO, "fp64": om.SCATTER_FP64,
                                          "mfma": om.SCATTER_MFMA_F16,
                                          "fp32": om.SCATTER_FP32}[args.scatter])
    scatter_name = {om.SCATTER_FP64: "fp64", om.SCATTER_MFMA_F16: "mfma_f16_split",
                    om.SCATTER_FP32: "fp32"}[scatter]
    shape = solver.getLaunchShape()
    if args.launch_shape:
        kv = dict(item.split("=") for item in args.launch_shape.split(","))
        solver.setLaunchShape(**{k: int(v) for k, v in kv.items()})
        shape = solver.getLaunchShape()
    elif not args.no_autotune:
        from openmoc_amd.autotune import autotune
        solver.computeEigenvalue(3)
        shape, t_best = autotune(solver, repeats=5)
        log(f"autotune: {shape} -> {t_best * 1e6:.1f} us/iteration")

    solve = None
    if not args.no_solve:
        solver.computeEigenvalue(problem.max_iters)          # warm (JIT of code objects, caches)
        t0 = time.perf_counter()
        solver.computeEigenvalue(problem.max_iters)
        solve_s = time.perf_counter() - t0
        keff = solver.getKeff()
        solve = {"seconds": solve_s, "solves_per_s": 1.0 / solve_s,
                 "iterations": solver.getNumIterations(), "keff": keff}
        golden = ROOT / "tests" / "golden" / (
            f"{args.workload}_ls_ref.npz" if args.linear_source else f"{args.workload}_ref.npz")
        if golden.exists():
            g = np.load(golden)
            flux = solver.getFluxes()
            ref = g["flux"].astype(np.float64)
            solve["dk_pcm_vs_reference"] = (keff - float(g["keff"])) * 1e5
            solve["flux_rel_l2_vs_reference"] = float(np.linalg.norm(flux - ref) / np.linalg.norm(ref))
            solve["reference_iterations"] = int(g["num_iterations"])
        log(f"solve: {solve}")

    # steps: full source iterations, stopping rule disabled.  The warm-up
    # call also captures the iteration graphs (both start-flux parities, with
    # the stopping rule off), which the timed call then replays
    solver.runIterations(max(args.warmup, 1), time_sweep=False)
    # exactly `steps` iterations; runIterations synchronises the device on
    # entry (state read) and before returning
    t0 = time.perf_counter()
    step_dev_s, _ = solver.runIterations(args.steps, time_sweep=False)
    elapsed = time.perf_counter() - t0
    # the sweep kernel alone (HIP events on its stream), for the roofline
    _, sweep_s = solver.runIterations(8)
    value = w_sweep * args.steps / elapsed
    ms_per_step = elapsed / args.steps * 1e3

    bytes_sweep = 8 * S + 8 * N * G + 16 * T * G * P
    achieved_gbs = bytes_sweep / sweep_s / 1e9
    achieved_tflops = FLOPS_PER_INTEGRATION * w_sweep / sweep_s / 1e12
    if args.linear_source:
        kernel = "sweep2d_ls_kernel"
    elif tg.info.is_3d and shape.get("lds_groups", 0) > 0:
        kernel = "sweep3d_stack_kernel"
    else:
        kernel = "sweep2d_lds_kernel" if shape.get("lds_groups", 0) > 0 else "sweep2d_kernel"
    traffic, traffic_src = measured_traffic(kernel, args.workload)
    # the binding roof is the lower of the two ceilings on integrations/s:
    # BW_HBM / B_int (algorithmic bytes) and P_FP32 / F_int
    b_int = bytes_sweep / w_sweep
    ceil_hbm = HBM_PEAK_GBS * 1e9 / b_int
    ceil_valu = FP32_PEAK_TFLOPS * 1e12 / FLOPS_PER_INTEGRATION
    roofline_hbm = {"achieved": achieved_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved_gbs / HBM_PEAK_GBS, "traffic": traffic,
                    "traffic_source": traffic_src, "kernel": kernel,
                    "algorithmic_bytes_per_launch": bytes_sweep,
                    "bytes_per_integration": b_int, "ceiling_integrations_per_s": ceil_hbm,
                    "launch_seconds": sweep_s}
    roofline_valu = {"achieved": achieved_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP32_PEAK_TFLOPS,
                     "peak_unpacked": FP32_UNPACKED_PEAK_TFLOPS,
                     "frac_unpacked": achieved_tflops / FP32_UNPACKED_PEAK_TFLOPS,
                     "flops_per_integration": FLOPS_PER_INTEGRATION,
                     "ceiling_integrations_per_s": ceil_valu, "kernel": kernel,
                     "launch_seconds": sweep_s}
    if ceil_valu <= ceil_hbm:
        roofline = dict(roofline_valu, bound="valu", traffic=traffic, traffic_source=traffic_src,
                        algorithmic_bytes_per_launch=bytes_sweep)
    else:
        roofline = dict(roofline_hbm, bound="hbm")

    serving = None
    if args.serving > 0:
        log(f"serving: {args.serving} requests ...")
        serving = serving_bench(args.serving, max_active=args.max_active)
        log(f"serving: {serving}")

    cpu = None
    if not args.no_cpu_baseline:
        threads, host = host_cpu()
        log(f"cpu baseline: {args.cpu_sweeps} sweeps on {threads} threads ({host}) ...")
        three_d = bool(tg.info.is_3d)
        sweeps = max(2, args.cpu_sweeps // 4) if three_d else args.cpu_sweeps
        cpu = cpu_baseline(text, G, P, S, sweeps, threads, use_reference=not three_d)
        cpu.update(host)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC, "value": value, "unit": "integrations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "tracks": T, "segments": S, "fsrs": N,
                   "groups": G, "polar_half": P, "num_azim": problem.num_azim,
                   "azim_spacing": problem.azim_spacing, "exponential": args.exp,
                   "source": "linear" if args.linear_source else "flat",
                   "scatter": scatter_name,
                   "launch_shape": shape,
                   "step": "one source iteration (sweep + per-FSR FP64 steps)"},
        "sweep_integrations_per_s": w_sweep / sweep_s,
        "roofline": roofline, "roofline_hbm": roofline_hbm, "roofline_valu": roofline_valu,
        "cpu_baseline": cpu, "solve": solve, "serving": serving,
    }
    if cpu:
        out["speedup_vs_cpu_sweep"] = (w_sweep / sweep_s) / cpu["value"]
    print(json.dumps(out), flush=True)
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    return 0


if __name__ == "__main__":
    sys.exit(main())
