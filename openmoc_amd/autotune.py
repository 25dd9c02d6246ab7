"""Sweep launch-shape autotuner.

Times a set of candidate launch shapes of the transport sweep on the solver's
own problem (HIP events around whole source iterations, which include the
reduction of the sweep's partial tallies, moc_solver_run_iterations) and
keeps the fastest.  Candidates cover both tally paths:

  * LDS sweep (lds_groups > 0): groups per block subset (= lanes per track),
    whether the subset's Q/sigma_t slice is staged in LDS (lds_q), threads per
    block, blocks per subset, and how blocks hand their tally on (lds_flush:
    per-block partial slices, or exact atomics into one int64 tally);
  * global-tally sweep (lds_groups = 0): groups per lane, block size and how
    segments hand their tally on (seg_store: stored per segment and summed
    per FSR in a fixed order, or float atomics into private tally copies).

Shapes the device rejects (LDS size) are skipped.  Results are cached per
problem signature for the life of the process.
"""
from __future__ import annotations

import itertools
from typing import Dict, List, Optional, Tuple

_CACHE: Dict[Tuple, Tuple[dict, float]] = {}


def _signature(solver) -> Tuple:
    tg = solver._tg
    if tg is None:
        return (solver._num_fsrs, solver._num_groups, id(solver))
    i = tg.info
    return (i.num_tracks, i.num_segments, i.num_fsrs, i.num_groups, i.num_materials,
            i.num_azim, i.num_polar)


def candidates(num_groups: int, num_cus: int = 256, is_3d: bool = False) -> List[dict]:
    out = []
    if is_3d:
        # z-stack sweep: one block per (track group, group subset), its
        # sources in LDS, so only the subset width and the block size vary
        for gb in sorted({1, 2, 3, 4, 7, min(num_groups, 64)}):
            if gb > num_groups:
                continue
            for block in (512, 1024):
                out.append(dict(lds_groups=gb, lds_q=1, lds_block=block))
        return out
    for gb in sorted({1, 2, 3, 4, 7, min(num_groups, 64)}):
        if gb > num_groups:
            continue
        nsub = -(-num_groups // gb)
        # Blocks per subset, multiples of 8: the subsets of one track slot
        # then run on one XCD and share its L2 for the segment stream (other
        # counts fetch several times the stream: on the 17x17 lattice 7 x 73
        # blocks sweep 3% faster than 7 x 72 but fetch 357 MB against 47).
        # A slice that fills most of a CU's LDS leaves one block per CU, so
        # the sweep runs in rounds of num_cus blocks: counts rounded down
        # fill whole rounds, counts rounded up give smaller blocks.
        counts = set()
        for per_cu in (1, 2, 4):
            counts.add(max(8, ((num_cus * per_cu) // nsub + 7) // 8 * 8))
            counts.add(max(8, (num_cus * per_cu) // nsub // 8 * 8))
        counts.add(max(8, (4 * num_cus) // (3 * nsub) // 8 * 8))
        for q, block, bps, flush in itertools.product((1, 0), (512, 1024), sorted(counts), (0, 1)):
            out.append(dict(lds_groups=gb, lds_q=q, lds_block=block, lds_blocks_per_sub=bps,
                            lds_flush=flush))
    for gpl in (1, 2, 4, 9):
        lpt = -(-num_groups // gpl)
        if lpt > 64:
            continue
        # a full wave of tracks, or half of one: a problem of a few thousand
        # tracks has only a few waves per CU, and twice as many half-empty
        # waves hide more of the gathers' latency
        tpws = sorted({64 // lpt, max(1, (64 // lpt) // 2)})
        for block, (store, copies), tpw in itertools.product((64, 256), ((1, 1), (0, 1), (0, 8)), tpws):
            out.append(dict(lds_groups=0, groups_per_lane=gpl, lanes_per_track=lpt,
                            tracks_per_wave=tpw, block=block, tally_copies=copies,
                            seg_store=store))
    return out


def autotune(solver, repeats: int = 5, num_cus: int = 256, verbose: bool = False,
             cands: Optional[List[dict]] = None) -> Tuple[dict, float]:
    """Pick the fastest launch shape for `solver`'s problem and apply it.
    Returns (shape, seconds per source iteration)."""
    key = _signature(solver)
    if key in _CACHE:
        shape, t = _CACHE[key]
        solver.setLaunchShape(**shape)
        return shape, t
    best: Optional[Tuple[dict, float]] = None
    start = solver.getLaunchShape()
    is_3d = bool(solver._tg is not None and solver._tg.info.is_3d)
    for c in (cands or candidates(solver._num_groups, num_cus, is_3d)):
        shape = dict(start)
        shape.update(c)
        try:
            solver.setLaunchShape(**shape)
        except RuntimeError:
            continue
        # a whole source iteration (the sweep, the reduction of its partial
        # tallies and the per-FSR steps), as a solve runs it
        solver.runIterations(2, time_sweep=False)
        t, _ = solver.runIterations(max(8, repeats), time_sweep=False)
        if verbose:
            print(f"autotune {t * 1e6:9.1f} us {c}", flush=True)
        if best is None or t < best[1]:
            best = (solver.getLaunchShape(), t)
    if best is None:
        solver.setLaunchShape(**start)
        return start, float("nan")
    solver.setLaunchShape(**best[0])
    _CACHE[key] = best
    return best
